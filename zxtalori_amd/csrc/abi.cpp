// The zo_* C ABI (include/zxtalori.h): the reference's export table
// (src/exports.zig:4-120, src/native_api.zig:9-55) over the HIP engine.
#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "../../include/zxtalori.h"
#include "batch.h"
#include "engine.h"
#include "groups_table.h"
#include "hopf.h"
#include "qvm_device.h"

struct zxt_engine {
  zxt::Engine engine;
};

struct zxt_batch {
  zxt_batch(bool graphs, uint32_t streams) : batch(graphs, streams) {}
  zxt::Batch batch;
  std::string text;
};

namespace {

char g_create_error[1024] = {0};

// Runs f; a thrown EngineError becomes -1 with "prefix: message" in zo_error.
template <typename F>
int32_t guarded(zxt_engine* h, const char* prefix, F&& f) {
  try {
    f();
    return 0;
  } catch (const zxt::EngineError& err) {
    h->engine.err = std::string(prefix) + ": " + err.message;
  } catch (const std::bad_alloc&) {
    h->engine.err = std::string(prefix) + ": OutOfMemory";
  } catch (...) {
    h->engine.err = std::string(prefix) + ": Unexpected";
  }
  return -1;
}

template <typename F>
int32_t guarded_batch(zxt_batch* b, const char* prefix, F&& f) {
  try {
    f();
    return 0;
  } catch (const zxt::EngineError& err) {
    b->batch.err = std::string(prefix) + ": " + err.message;
  } catch (const std::bad_alloc&) {
    b->batch.err = std::string(prefix) + ": OutOfMemory";
  } catch (...) {
    b->batch.err = std::string(prefix) + ": Unexpected";
  }
  return -1;
}

bool valid_group(uint32_t id) { return id < zxt::kGroupCount; }

}  // namespace

void zxt::set_create_error(const char* text) { snprintf(g_create_error, sizeof(g_create_error), "%s", text); }

extern "C" {

const char* zo_version(void) { return "0.3.0-mi355x"; }

zxt_engine* zo_create(const char* /*wgpu_dir*/) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    snprintf(g_create_error, sizeof(g_create_error), "GPU init: NoHipDevice");
    return nullptr;
  }
  try {
    return new zxt_engine();
  } catch (const zxt::EngineError& err) {
    snprintf(g_create_error, sizeof(g_create_error), "Engine init: %s", err.message.c_str());
  } catch (...) {
    snprintf(g_create_error, sizeof(g_create_error), "Engine init: Unexpected");
  }
  return nullptr;
}

const char* zo_create_error(void) { return g_create_error; }

void zo_destroy(zxt_engine* e) { delete e; }

int32_t zo_random(zxt_engine* e, uint32_t group, uint32_t count, uint32_t seed) {
  if (!valid_group(group)) {
    e->engine.err = "Unknown group";
    return -1;
  }
  return guarded(e, "Haar initialization", [&] { e->engine.generate_random(group, count, seed); });
}

int32_t zo_set(zxt_engine* e, uint32_t group, const float* q, uint32_t n) {
  if (!valid_group(group)) {
    e->engine.err = "Unknown group";
    return -1;
  }
  return guarded(e, "Import failed", [&] { e->engine.set(group, q, size_t(n) * 4); });
}

int32_t zo_generate(zxt_engine* e, uint32_t group, uint32_t h, uint32_t target, uint32_t seed, uint32_t kind) {
  if (!valid_group(group)) {
    e->engine.err = "Unknown group";
    return -1;
  }
  if (kind > 2) {
    e->engine.err = "Unknown grid kind";
    return -1;
  }
  return guarded(e, "Generation failed", [&] { e->engine.generate(group, h, target, seed, kind); });
}

uint32_t zo_hopf_points(uint32_t group, uint32_t f, float* out, uint32_t capacity) {
  if (!valid_group(group) || f < 1 || f > zxt::hopf::kMaxFrequency) return 0;
  if (zxt::hopf::raw_count(group, f) > zxt::hopf::kMaxCount) return 0;
  const uint32_t n = uint32_t(zxt::hopf::raw_count(group, f));
  if (!out || capacity < n) return n;
  try {
    zxt::hopf::points(group, f, out);
    return n;
  } catch (...) {
    return 0;
  }
}

uint32_t zo_hopf_frequency(uint32_t group, uint32_t target) {
  if (!valid_group(group)) return 0;
  return zxt::hopf::nearest_frequency(group, target == 0 ? 2048 : target);
}

uint32_t zo_count(zxt_engine* e) { return e->engine.count(); }
uint32_t zo_order(zxt_engine* e) { return e->engine.order(); }
const float* zo_operations(zxt_engine* e) { return e->engine.operations(); }
const float* zo_quaternions(zxt_engine* e) { return e->engine.quaternions(); }

int32_t zo_measure(zxt_engine* e) {
  return guarded(e, "Measure failed", [&] { e->engine.measure(); });
}
int32_t zo_relax(zxt_engine* e, uint32_t steps) {
  return guarded(e, "Relaxation failed", [&] { e->engine.relax(steps); });
}
int32_t zo_relax_neighbourhood(zxt_engine* e, float delta) {
  return guarded(e, "Relaxation neighbourhood", [&] { e->engine.set_neighbourhood(double(delta)); });
}
int32_t zo_spacing(zxt_engine* e) {
  return guarded(e, "Spacing failed", [&] { e->engine.spacing(); });
}
const float* zo_point_values(zxt_engine* e, uint32_t kind) { return e->engine.point_values(kind); }
int32_t zo_identity(zxt_engine* e, uint32_t flags) {
  return guarded(e, "Identity constraint", [&] { e->engine.identity((flags & 1u) != 0, (flags & 2u) != 0); });
}
const char* zo_stats(zxt_engine* e) {
  try {
    e->engine.stats();
  } catch (...) {
    e->engine.text = "{}";
  }
  return e->engine.text.c_str();
}
const char* zo_error(zxt_engine* e) { return e->engine.err.c_str(); }
void zo_cancel(zxt_engine* e) { e->engine.cancelled.store(true); }

int32_t zo_cover(zxt_engine* e) {
  return guarded(e, "Covering radius", [&] { e->engine.cover(); });
}
int32_t zo_cover_optimize(zxt_engine* e, uint32_t steps, uint32_t mode, uint32_t finish, float floor_deg) {
  return guarded(e, "Covering optimization",
                 [&] { e->engine.optimize_covering(steps, mode, finish != 0, floor_deg); });
}
int32_t zo_cover_kicks(zxt_engine* e, uint32_t patience, float progress, float size, float trust) {
  return guarded(e, "Covering kicks", [&] { e->engine.set_kicks(patience, progress, size, trust); });
}
int32_t zo_mesh(zxt_engine* e) {
  return guarded(e, "DT display", [&] { e->engine.build_mesh(); });
}
const float* zo_mesh_vertices(zxt_engine* e) { return e->engine.mesh_vertices(); }
const uint32_t* zo_mesh_tets(zxt_engine* e) { return e->engine.mesh_tets(); }
uint32_t zo_mesh_vertex_count(zxt_engine* e) { return e->engine.mesh_vertex_count(); }
uint32_t zo_mesh_tet_count(zxt_engine* e) { return e->engine.mesh_tet_count(); }
int32_t zo_generate_exact(zxt_engine* e, uint32_t group, uint32_t target, uint32_t flags) {
  if (!valid_group(group)) {
    e->engine.err = "Unknown group";
    return -1;
  }
  const uint32_t kind = (flags >> 8) & 0xffu;
  if (kind > 2) {
    e->engine.err = "Unknown grid kind";
    return -1;
  }
  return guarded(e, "Exact-count generation",
                 [&] { e->engine.generate_exact(group, target, (flags & 1u) != 0, kind); });
}

bool zo_profile(zxt_engine* e, bool enabled) {
  e->engine.prof.enable(enabled);
  return e->engine.prof.on();
}
size_t zo_profile_report(zxt_engine* e, char* output, size_t capacity) {
  const std::string r = e->engine.prof.report();
  if (!output || capacity == 0) return r.size();
  const size_t n = std::min(capacity - 1, r.size());
  std::memcpy(output, r.data(), n);
  output[n] = 0;
  return n;
}
void zo_pair_budget(zxt_engine* e, float scale) { e->engine.pair_budget_scale = double(scale); }
void zo_delaunay_mode(zxt_engine*, uint32_t) {}
void zo_tune_allpairs(uint32_t target, uint32_t ops_per_split) { zxt::set_allpairs_override(target, ops_per_split); }

int32_t zo_issue_probe(uint32_t kind, uint32_t iters, double* rate_out) {
  if (kind > 2 || iters == 0 || !rate_out) return -1;
  int sms = 0;
  if (hipDeviceGetAttribute(&sms, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || sms < 1) return -1;
  const uint32_t blocks = uint32_t(sms) * 8;  // 8 waves per SIMD
  float* scratch = nullptr;
  if (hipMalloc(&scratch, size_t(blocks) * 256 * sizeof(float)) != hipSuccess) return -1;
  *rate_out = zxt::probe_issue_rate(nullptr, kind, iters, scratch, blocks);
  (void)hipFree(scratch);
  return *rate_out > 0 ? 0 : -1;
}

const uint32_t* zo_point_indices(zxt_engine* e) { return e->engine.point_indices(); }
int32_t zo_force_passes(zxt_engine* e, uint32_t passes) {
  return guarded(e, "Force passes", [&] { e->engine.force_passes(passes); });
}
int32_t zo_force_rows(zxt_engine* e, float* force4, double* energy, float* nearest) {
  return guarded(e, "Force rows", [&] { e->engine.force_rows(force4, energy, nearest); });
}
int32_t zo_near_pairs(zxt_engine* e, uint64_t* out2) {
  out2[0] = e->engine.last_near_counts[0];
  out2[1] = e->engine.last_near_counts[1];
  return 0;
}
int32_t zo_near_info(zxt_engine* e, uint32_t* out3) {
  out3[0] = e->engine.last_near.sites;
  out3[1] = e->engine.last_near.sites >= e->engine.count() ? e->engine.last_near.sites - e->engine.count() : 0;
  out3[2] = e->engine.last_near_sdim;
  return 0;
}

int32_t zo_near_squeeze(zxt_engine* e, int32_t sites) {
  e->engine.near_squeeze = sites;
  e->engine.near_squeeze_pending = true;
  return 0;
}
uint32_t zo_near_resizes(zxt_engine* e) { return e->engine.near_resizes; }

int32_t zo_near_cut(zxt_engine* e, float* out3) {
  out3[0] = e->engine.last_near_cut.chord2;
  out3[1] = float(e->engine.last_near_radius);
  out3[2] = float(zxt::atan_terms(e->engine.last_near_cut.chord2));
  return 0;
}

// ---- batches ----------------------------------------------------------------------
zxt_batch* zo_batch_create(uint32_t flags) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    snprintf(g_create_error, sizeof(g_create_error), "GPU init: NoHipDevice");
    return nullptr;
  }
  try {
    return new zxt_batch((flags & 1u) != 0, (flags >> 8) & 0xffu);
  } catch (const zxt::EngineError& err) {
    snprintf(g_create_error, sizeof(g_create_error), "Batch init: %s", err.message.c_str());
  } catch (...) {
    snprintf(g_create_error, sizeof(g_create_error), "Batch init: Unexpected");
  }
  return nullptr;
}

void zo_batch_destroy(zxt_batch* b) { delete b; }
const char* zo_batch_error(zxt_batch* b) { return b->batch.err.c_str(); }

int32_t zo_batch_add(zxt_batch* b, uint32_t group, const float* q, uint32_t n, uint32_t seed, float delta,
                     uint32_t kind, uint32_t steps, uint64_t* id_out) {
  if (!valid_group(group)) {
    b->batch.err = "Unknown group";
    return -1;
  }
  if (kind > 2) {
    b->batch.err = "Batch add: UnknownJobKind";
    return -1;
  }
  return guarded_batch(b, "Batch add", [&] {
    *id_out = b->batch.add(group, q, n, seed, double(delta), zxt::JobKind(kind), steps);
  });
}

int32_t zo_batch_step(zxt_batch* b, uint32_t* completed) {
  return guarded_batch(b, "Batch step", [&] {
    const uint32_t done = b->batch.step();
    if (completed) *completed = done;
  });
}

int32_t zo_batch_advance(zxt_batch* b, uint32_t most, uint32_t* completed, uint32_t* taken) {
  return guarded_batch(b, "Batch advance", [&] {
    uint32_t steps = 0;
    const uint32_t done = b->batch.step(std::max<uint32_t>(most, 1), &steps);
    if (completed) *completed = done;
    if (taken) *taken = steps;
  });
}

int32_t zo_batch_status(zxt_batch* b, uint64_t id) {
  int32_t s = -1;
  if (guarded_batch(b, "Batch status", [&] { s = b->batch.status(id); }) != 0) return -1;
  return s;
}

const char* zo_batch_job_error(zxt_batch* b, uint64_t id) {
  if (guarded_batch(b, "Batch job error", [&] { b->text = b->batch.error(id); }) != 0) return "";
  return b->text.c_str();
}

const char* zo_batch_stats(zxt_batch* b, uint64_t id) {
  if (guarded_batch(b, "Batch stats", [&] { b->text = b->batch.stats(id); }) != 0) return "";
  return b->text.c_str();
}

const float* zo_batch_quaternions(zxt_batch* b, uint64_t id) {
  const float* p = nullptr;
  guarded_batch(b, "Batch quaternions", [&] { p = b->batch.quaternions(id); });
  return p;
}

uint32_t zo_batch_count(zxt_batch* b, uint64_t id) {
  uint32_t c = 0;
  guarded_batch(b, "Batch count", [&] { c = b->batch.count(id); });
  return c;
}

const float* zo_batch_point_values(zxt_batch* b, uint64_t id, uint32_t kind) {
  const float* p = nullptr;
  guarded_batch(b, "Batch point values", [&] { p = b->batch.engine(id).point_values(kind); });
  return p;
}

const uint32_t* zo_batch_point_indices(zxt_batch* b, uint64_t id) {
  const uint32_t* p = nullptr;
  guarded_batch(b, "Batch point indices", [&] { p = b->batch.engine(id).point_indices(); });
  return p;
}

uint64_t zo_batch_bytes(zxt_batch* b, uint64_t id) {
  uint64_t v = 0;
  guarded_batch(b, "Batch bytes", [&] { v = b->batch.bytes(id); });
  return v;
}

int32_t zo_batch_remove(zxt_batch* b, uint64_t id) {
  return guarded_batch(b, "Batch remove", [&] { b->batch.remove(id); });
}

uint32_t zo_batch_running(zxt_batch* b) { return b->batch.running(); }

int32_t zo_batch_near_squeeze(zxt_batch* b, uint64_t id, int32_t sites) {
  return guarded_batch(b, "Batch near squeeze", [&] {
    zxt::Engine& e = b->batch.engine(id);
    e.near_squeeze = sites;
    e.near_squeeze_pending = true;
  });
}

uint32_t zo_batch_near_resizes(zxt_batch* b, uint64_t id) {
  uint32_t v = 0;
  guarded_batch(b, "Batch near resizes", [&] { v = b->batch.engine(id).near_resizes; });
  return v;
}

int32_t zo_batch_wait(zxt_batch* b, uint32_t most, uint32_t* in_flight) {
  return guarded_batch(b, "Batch wait", [&] {
    b->batch.wait(most);
    if (in_flight) *in_flight = b->batch.in_flight();
  });
}

int32_t zo_batch_graphs(zxt_batch* b, uint64_t* out2) {
  out2[0] = b->batch.captures();
  out2[1] = b->batch.replays();
  return 0;
}

}  // extern "C"
