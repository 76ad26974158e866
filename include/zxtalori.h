/* zxtalori C ABI, MI355X (gfx950) build: libzxtalori.so.
 *
 * Drop-in for the reference's exported table (src/exports.zig, src/native_api.zig), which
 * its Python adapter resolves at load time (python/zxtalori/api.py:29-53).  Every entry
 * keeps the reference's signature, return convention (0 success, -1 failure with the
 * message in zo_error; NULL from zo_create with the message in zo_create_error),
 * ownership (returned pointers belong to the engine and stay valid until the next
 * mutating call) and threading (one engine per thread; zo_cancel from any thread).
 *
 * Group ids 0..10 are C2 C3 C4 C6 D2 D3 D4 D6 T O I as in the reference
 * (src/exports.zig:10, src/groups.zig:2); id 11 is C1 (identity only), an extension.
 */
#ifndef ZXTALORI_H
#define ZXTALORI_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zxt_engine zxt_engine;

/* src/exports.zig:4-6 */
const char* zo_version(void);

/* src/native_api.zig:9-21.  The argument (the reference's wgpu-native directory) is
 * ignored; the engine takes HIP device 0. */
zxt_engine* zo_create(const char* wgpu_dir);
/* src/native_api.zig:20-22 */
const char* zo_create_error(void);
/* src/native_api.zig:23-29 */
void zo_destroy(zxt_engine* e);

/* src/exports.zig:83-87: seeded Haar-random orientations, folded into the zone. */
int32_t zo_random(zxt_engine* e, uint32_t group, uint32_t count, uint32_t seed);
/* src/exports.zig:23-28: import n quaternions (scalar first, copied), folded. */
int32_t zo_set(zxt_engine* e, uint32_t group, const float* q, uint32_t n);
/* src/exports.zig:9-15 (src/engine.zig:242-282): kind 0 or 1, the inverse Hopf product grid
 * at frequency h (1..512), or with h = 0 the frequency whose count is nearest `target`
 * (2048 when 0), folded into the zone.  Errors: "Generation failed: GridTooLarge" (more than
 * 10,000,000 points), "...: InvalidGridResolution".  kind 2: the FCC transport of T, O or I
 * at semi-edge max(1, h) <= 128 in lattice order, not folded (src/grids.zig:40-77), or
 * "...: TransportNeedsTOrOOrI" for other groups. */
int32_t zo_generate(zxt_engine* e, uint32_t group, uint32_t h, uint32_t target, uint32_t seed, uint32_t kind);

/* src/exports.zig:29-40 */
uint32_t zo_count(zxt_engine* e);
uint32_t zo_order(zxt_engine* e);
const float* zo_operations(zxt_engine* e); /* zo_order() rows of 4 floats */
const float* zo_quaternions(zxt_engine* e); /* zo_count() rows of 4 floats */

/* src/exports.zig:41-44: energy and separation (and the energy ratios below the ratio limit). */
int32_t zo_measure(zxt_engine* e);
/* src/exports.zig:49-52: up to 1000 BB steps. */
int32_t zo_relax(zxt_engine* e, uint32_t steps);
/* src/exports.zig:55-58: 0 (all pairs) or 2..1000 mean orientation radii. */
int32_t zo_relax_neighbourhood(zxt_engine* e, float delta);
/* src/exports.zig:60-63: each point's nearest orbit neighbour and the separation. */
int32_t zo_spacing(zxt_engine* e);
/* src/exports.zig:66-68: kind 0 nearest-neighbour SO(3) degrees; NULL if not measured. */
const float* zo_point_values(zxt_engine* e, uint32_t kind);
/* src/exports.zig:69-72: bit 0 add the identity, bit 1 hold it fixed. */
int32_t zo_identity(zxt_engine* e, uint32_t flags);
/* src/exports.zig:73-81 */
const char* zo_stats(zxt_engine* e);
const char* zo_error(zxt_engine* e);
/* Any thread: a running zo_relax or zo_cover_optimize stops between steps (the relax waits
 * for the device every ~2e10 pair evaluations, as engine.zig:896-915 does). */
void zo_cancel(zxt_engine* e);

/* src/exports.zig:45-48 (engine.zig:378-398): the covering radius of the sample, the
 * largest empty cap over every signed orbit image, into the stats record
 * (covering_radius_deg, covering_witness, ...) and point values 1 (cap radius, degrees) and
 * 2 (Voronoi faces); reused while the sample is unchanged.  Errors: "Covering radius:
 * TooFewOrientations" (fewer than 4), "...: CoveringDomainNotEnclosed". */
int32_t zo_cover(zxt_engine* e);
/* src/exports.zig:104-120 (engine.zig:399-420): `steps` covering optimisation steps, each
 * pulling the points of the largest empty cap toward its centre and kept when the radius
 * drops (mode & 3: 0 "tet", 1 "face" = half steps; other bits as the reference's variants);
 * `finish` leaves the best sample seen; moves never bring a pair nearer than floor_deg
 * (SO(3) degrees).  zo_cover_kicks: plateau patience, relative gain, kick size (times the best
 * radius) and restart trust.  Errors: "Covering optimization: InvalidCoveringMode",
 * "...: InvalidSeparationFloor", "Covering kicks: InvalidKickPatience" and the like. */
int32_t zo_cover_optimize(zxt_engine* e, uint32_t steps, uint32_t mode, uint32_t finish, float floor_deg);
int32_t zo_cover_kicks(zxt_engine* e, uint32_t patience, float progress, float size, float trust);
/* src/exports.zig:88-103 (engine.zig:352-358): a covering measurement that also keeps the
 * Delaunay tetrahedra around the sample's orbit points: vertices as stereographic
 * (x, y, z, 0) of the sample (first, in order) and of the other orbit images used, tetrahedra
 * as 4 vertex indices; valid until the next mutating call.  Errors as zo_cover, prefixed
 * "DT display". */
int32_t zo_mesh(zxt_engine* e);
const float* zo_mesh_vertices(zxt_engine* e);
const uint32_t* zo_mesh_tets(zxt_engine* e);
uint32_t zo_mesh_vertex_count(zxt_engine* e);
uint32_t zo_mesh_tet_count(zxt_engine* e);
/* src/exports.zig:16-22 (engine.zig:199-232): exactly `target` (32..1,000,000) orientations,
 * the largest base grid of `kind` (flags bits 8-15) that fits and then centres of the largest
 * empty caps; flags bit 0 adds the identity.  Errors: "Exact-count generation:
 * ExactCountRange32To1000000", "...: BaseGridExceedsTarget", "...: InsufficientDelaunayCentres". */
int32_t zo_generate_exact(zxt_engine* e, uint32_t group, uint32_t target, uint32_t flags);

/* Development hooks (src/native_api.zig:31-43). */
/* Per-kernel HIP-event timing on the engine's stream; returns the new state. */
bool zo_profile(zxt_engine* e, bool enabled);
/* Tab-separated "kernel count total_ms" lines; returns the bytes written. */
size_t zo_profile_report(zxt_engine* e, char* output, size_t capacity);
/* Scales the all-pairs work split (tests shrink it: only rounding may change). */
void zo_pair_budget(zxt_engine* e, float scale);
/* src/native_api.zig:37-39: picks among the reference's Delaunay schedulers.  This build
 * has one covering method (per-point Voronoi cells), so the mode is accepted and ignored. */
void zo_delaunay_mode(zxt_engine* e, uint32_t mode);
/* No counterpart in the reference: the autotuner's hook (zxtalori_amd/tune.py).  Sets, for
 * every engine of the process, the all-pairs force pass's tile plan: about `target`
 * workgroups, `ops_per_split` operators per workgroup (0: automatic); target 0 returns to the
 * compiled-in tuned table. */
void zo_tune_allpairs(uint32_t target, uint32_t ops_per_split);

/* No counterpart in the reference: the chip's issue rates, measured on device 0 with long
 * chains of one instruction kind on 8 waves per SIMD (kind 0 non-packed FP32 FMA, kind 1
 * transcendentals v_rsq_f32/v_rcp_f32, kind 2 v_mfma_f32_16x16x4f32 FP32 Gram tiles), in
 * wave64 instructions per second; the peaks the pair kernels' counters are read against.
 * 0, or -1 on failure. */
int32_t zo_issue_probe(uint32_t kind, uint32_t iters, double* rate_out);

/* Additions of this build (no reference counterpart). */
/* Per point, the rank j*2*order + 2*k + sign of its nearest orbit point (source j,
 * operator k, sign 0 for +q_j*g_k), from the last zo_spacing; NULL before one. */
const uint32_t* zo_point_indices(zxt_engine* e);
/* Runs `passes` force passes of the current neighbourhood on the current positions and
 * waits; nothing is read back (throughput measurement). */
int32_t zo_force_passes(zxt_engine* e, uint32_t passes);
/* One force pass of the current neighbourhood on the current positions, its rows copied
 * out: per point (fx, fy, fz, curvature) in force4, the energy sum and the nearest
 * chord^2 (the reference's Row columns force and extra.y/extra.z, pairs.zig:11-15). */
int32_t zo_force_rows(zxt_engine* e, float* force4, double* energy, float* nearest);
/* The last near-field force pass's accepted pair evaluations (out2[0]: source and orbit
 * point within the cutoff, not of the source's own orbit) and sites tested (out2[1]). */
int32_t zo_near_pairs(zxt_engine* e, uint64_t* out2);
/* The last near-field pass's site count, shell count and cell grid dimension. */
int32_t zo_near_info(zxt_engine* e, uint32_t* out3);
/* Host only, no device: the inverse Hopf grid's rows before the zone fold (src/hopf.zig:516-604
 * points()), 4 floats each, in the reference's row order.  Returns the row count and writes the
 * rows when `capacity` rows fit; returns 0 for an unknown group, f outside 1..512 or more than
 * 10,000,000 rows. */
uint32_t zo_hopf_points(uint32_t group, uint32_t f, float* out, uint32_t capacity);
/* Host only: the frequency zo_generate takes for h = 0 and `target` (src/hopf.zig:82-97). */
uint32_t zo_hopf_frequency(uint32_t group, uint32_t target);
/* Development hook, the counterpart of the reference's capacity growth (src/near.zig:102-116,
 * :191-195; src/pairs.zig:391-394): the next near-field sizing leaves room for exactly the
 * site count it finds plus `sites` (may be negative; never fewer than the sources) instead
 * of a quarter more, so that a pass overflows and is redone, as a shell that outgrows its
 * buffers is.  Used once, then the usual headroom returns. */
int32_t zo_near_squeeze(zxt_engine* e, int32_t sites);
/* Near-field passes redone after their shell overflowed, over the engine's life. */
uint32_t zo_near_resizes(zxt_engine* e);
/* The last near-field pass's cutoff: chord^2, S3 radius r, and the number of Taylor terms
 * of the pair kernel's arctangent for it (0: the minimax polynomial). */
int32_t zo_near_cut(zxt_engine* e, float* out3);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops and harnesses that forbid hipMalloc): zo_spacing's
 * all-pairs path on caller-owned device buffers, allocating nothing.  q: n device rows
 * (w, x, y, z), 16-byte aligned, already in the fundamental zone (as zo_quaternions returns
 * them; they are not folded here); outputs, all device memory: rank_out[n] the nearest orbit
 * point's rank j*2|K| + 2k + sign, degrees_out[n] its SO(3) angle in degrees, least_out[1] the
 * separation (the least angle; 0 for duplicate orientations, where zo_spacing fails).  Launches
 * on `stream` (a hipStream_t, NULL for the null stream) and returns without waiting.  Limit:
 * 2 <= n <= 20,000 (zo_spacing's all-pairs range); the scratch is one static 10 MiB device
 * array per device, shared by every call on that device, so calls that overlap on one device
 * race.  Returns 0, or -1 with the message in zo_create_error() ("Spacing:
 * CountExceedsStaticScratch", "...: TooFewOrientations", "...: UnknownGroup", "...: NullPointer",
 * "...: MisalignedBuffer", "...: GpuError: ..."). */
int32_t zo_spacing_static(uint32_t group, const float* q, uint32_t n, uint32_t* rank_out, float* degrees_out,
                          float* least_out, void* stream);

/* ---- Batches: many samples relaxed together ------------------------------------------
 * No counterpart in the reference, which runs one sample per engine and one call at a time
 * (src/native_api.zig:9-19, python/zxtalori/api.py:95-122).  A batch holds jobs, each a
 * sample with its own resident state; zo_batch_step advances every running relax job by
 * one BB step with the all-pairs passes, BB steps and folds of all jobs as one launch per
 * stage (near-field jobs add their own passes), so a job's result is bit-identical alone
 * and in any mix.  Same int32 / error-string rules as the engine calls; jobs are uint64
 * handles.  One thread per batch. */
typedef struct zxt_batch zxt_batch;
/* flags bit 0: replay steady-state steps (same running jobs) from a captured HIP graph;
 * bits 8-15: extra streams (at most 8) that the near-field passes of different jobs are
 * spread over, so that small passes overlap on the device (0: one stream).
 * NULL on failure, message in zo_create_error(). */
zxt_batch* zo_batch_create(uint32_t flags);
void zo_batch_destroy(zxt_batch* b);
const char* zo_batch_error(zxt_batch* b);
/* kind 0 relax (`steps` BB steps, at most 1000, then the end-of-relax measure), 1 measure,
 * 2 spacing.  q: n rows of (w,x,y,z) as zo_set takes them, or NULL for n seeded Haar
 * orientations (zo_random).  delta: the relax neighbourhood (0 all pairs, else 2..1000). */
int32_t zo_batch_add(zxt_batch* b, uint32_t group, const float* q, uint32_t n, uint32_t seed, float delta,
                     uint32_t kind, uint32_t steps, uint64_t* id_out);
/* One step of every running job (pending measure and spacing jobs run first); completed
 * receives how many jobs ended (done or failed) in it. */
int32_t zo_batch_step(zxt_batch* b, uint32_t* completed);
/* As zo_batch_step, but steps that end no job may run as one launch of a graph of up to
 * `most` steps (graph replay on, the same jobs running; at most 64): taken receives how
 * many steps every running job advanced (1 when no such span applies). */
int32_t zo_batch_advance(zxt_batch* b, uint32_t most, uint32_t* completed, uint32_t* taken);
/* 0 running, 1 done, 2 failed (zo_batch_job_error), -1 unknown id. */
int32_t zo_batch_status(zxt_batch* b, uint64_t id);
const char* zo_batch_job_error(zxt_batch* b, uint64_t id);
/* A job's engine views, as zo_stats / zo_quaternions / zo_count / zo_point_values /
 * zo_point_indices give them for an engine; valid until the job's next step or removal. */
const char* zo_batch_stats(zxt_batch* b, uint64_t id);
const float* zo_batch_quaternions(zxt_batch* b, uint64_t id);
uint32_t zo_batch_count(zxt_batch* b, uint64_t id);
const float* zo_batch_point_values(zxt_batch* b, uint64_t id, uint32_t kind);
const uint32_t* zo_batch_point_indices(zxt_batch* b, uint64_t id);
/* Device bytes a job holds now (admission by HBM budget). */
uint64_t zo_batch_bytes(zxt_batch* b, uint64_t id);
int32_t zo_batch_remove(zxt_batch* b, uint64_t id);
uint32_t zo_batch_running(zxt_batch* b);
/* Test hooks, as zo_near_squeeze and zo_near_resizes for a batch job's engine: the job's
 * next near-field sizing leaves room for the site count it finds plus `sites`. */
int32_t zo_batch_near_squeeze(zxt_batch* b, uint64_t id, int32_t sites);
uint32_t zo_batch_near_resizes(zxt_batch* b, uint64_t id);
/* Blocks until at most `most` steps of this batch are unfinished on the device (a serving
 * loop keeps each batch a step or two ahead of the device); in_flight receives how many are. */
int32_t zo_batch_wait(zxt_batch* b, uint32_t most, uint32_t* in_flight);
/* out2[0]: steps captured into a graph, out2[1]: steps replayed from one. */
int32_t zo_batch_graphs(zxt_batch* b, uint64_t* out2);

#ifdef __cplusplus
}
#endif
#endif
