"""ctypes surface over libzxtalori.so (the gfx950 build).

Mirrors the reference's Python adapter (python/zxtalori/api.py): the same names,
argument meanings, error behaviour (RuntimeError carrying zo_error's text) and JSON stats.
The QVM relaxation and spacing path, inverse Hopf and FCC transport generation (also with
exact counts), the covering radius, its Delaunay mesh and its optimisation run here.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path

import numpy as np

GROUPS = ("C2", "C3", "C4", "C6", "D2", "D3", "D4", "D6", "T", "O", "I")
# C1 (identity only) is accepted as an extension with id 11.
EXTENDED_GROUPS = GROUPS + ("C1",)
HERE = Path(__file__).resolve().parent
_LIB = None


def library_path():
    if os.environ.get("ZXTALORI_LIBRARY"):
        return Path(os.environ["ZXTALORI_LIBRARY"])
    return HERE / "_lib" / "libzxtalori.so"


def _library():
    global _LIB
    if _LIB is not None:
        return _LIB
    selected = library_path()
    if not selected.is_file():
        raise RuntimeError(f"zxtalori engine missing at {selected}: build it with `python -m zxtalori_amd.build`.")
    lib = C.CDLL(str(selected))
    ptr = C.c_void_p
    signatures = {
        "zo_version": ([], C.c_char_p),
        "zo_create": ([C.c_char_p], ptr), "zo_create_error": ([], C.c_char_p),
        "zo_destroy": ([ptr], None),
        "zo_generate": ([ptr, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32], C.c_int),
        "zo_random": ([ptr, C.c_uint32, C.c_uint32, C.c_uint32], C.c_int),
        "zo_set": ([ptr, C.c_uint32, C.POINTER(C.c_float), C.c_uint32], C.c_int),
        "zo_count": ([ptr], C.c_uint32), "zo_order": ([ptr], C.c_uint32),
        "zo_quaternions": ([ptr], C.POINTER(C.c_float)), "zo_operations": ([ptr], C.POINTER(C.c_float)),
        "zo_measure": ([ptr], C.c_int), "zo_spacing": ([ptr], C.c_int), "zo_relax": ([ptr, C.c_uint32], C.c_int),
        "zo_relax_neighbourhood": ([ptr, C.c_float], C.c_int),
        "zo_point_values": ([ptr, C.c_uint32], C.POINTER(C.c_float)),
        "zo_identity": ([ptr, C.c_uint32], C.c_int),
        "zo_cover": ([ptr], C.c_int),
        "zo_cover_optimize": ([ptr, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float], C.c_int),
        "zo_cover_kicks": ([ptr, C.c_uint32, C.c_float, C.c_float, C.c_float], C.c_int),
        "zo_mesh": ([ptr], C.c_int),
        "zo_mesh_vertices": ([ptr], C.POINTER(C.c_float)),
        "zo_mesh_tets": ([ptr], C.POINTER(C.c_uint32)),
        "zo_mesh_vertex_count": ([ptr], C.c_uint32),
        "zo_mesh_tet_count": ([ptr], C.c_uint32),
        "zo_generate_exact": ([ptr, C.c_uint32, C.c_uint32, C.c_uint32], C.c_int),
        "zo_stats": ([ptr], C.c_char_p), "zo_error": ([ptr], C.c_char_p), "zo_cancel": ([ptr], None),
        "zo_profile": ([ptr, C.c_bool], C.c_bool),
        "zo_profile_report": ([ptr, C.c_char_p, C.c_size_t], C.c_size_t),
        "zo_pair_budget": ([ptr, C.c_float], None),
        "zo_tune_allpairs": ([C.c_uint32, C.c_uint32], None),
        "zo_point_indices": ([ptr], C.POINTER(C.c_uint32)),
        "zo_force_passes": ([ptr, C.c_uint32], C.c_int),
        "zo_near_info": ([ptr, C.POINTER(C.c_uint32)], C.c_int),
        "zo_force_rows": ([ptr, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
        "zo_near_pairs": ([ptr, C.POINTER(C.c_uint64)], C.c_int),
        "zo_near_cut": ([ptr, C.POINTER(C.c_float)], C.c_int),
        "zo_hopf_points": ([C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32], C.c_uint32),
        "zo_hopf_frequency": ([C.c_uint32, C.c_uint32], C.c_uint32),
        # batches (no reference counterpart)
        "zo_batch_create": ([C.c_uint32], ptr), "zo_batch_destroy": ([ptr], None),
        "zo_batch_error": ([ptr], C.c_char_p),
        "zo_batch_add": ([ptr, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                          C.c_uint32, C.POINTER(C.c_uint64)], C.c_int),
        "zo_batch_step": ([ptr, C.POINTER(C.c_uint32)], C.c_int),
        "zo_batch_status": ([ptr, C.c_uint64], C.c_int),
        "zo_batch_job_error": ([ptr, C.c_uint64], C.c_char_p),
        "zo_batch_stats": ([ptr, C.c_uint64], C.c_char_p),
        "zo_batch_quaternions": ([ptr, C.c_uint64], C.POINTER(C.c_float)),
        "zo_batch_count": ([ptr, C.c_uint64], C.c_uint32),
        "zo_batch_point_values": ([ptr, C.c_uint64, C.c_uint32], C.POINTER(C.c_float)),
        "zo_batch_point_indices": ([ptr, C.c_uint64], C.POINTER(C.c_uint32)),
        "zo_batch_bytes": ([ptr, C.c_uint64], C.c_uint64),
        "zo_batch_remove": ([ptr, C.c_uint64], C.c_int),
        "zo_batch_running": ([ptr], C.c_uint32),
        "zo_batch_wait": ([ptr, C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
        "zo_batch_graphs": ([ptr, C.POINTER(C.c_uint64)], C.c_int),
    }
    for name, (args, result) in signatures.items():
        function = getattr(lib, name)
        function.argtypes, function.restype = args, result
    _LIB = lib
    return lib


def _group_id(group):
    try:
        return EXTENDED_GROUPS.index(str(group).upper())
    except ValueError:
        raise ValueError(f"group must be one of {', '.join(GROUPS)}") from None


KINDS = (None, "hopf", "transport")


def _kind_id(kind):
    try:
        return KINDS.index(None if kind is None else str(kind).lower())
    except ValueError:
        raise ValueError("kind must be 'hopf' or 'transport'") from None


class Grid:
    """A GPU-resident sample of orientation orbits on SO(3)/K. Close it to release the device.

    Quaternions have shape (N,4), scalar first, and right action q*g.  Construct samples
    with haar(), from_quaternions() or generate() (inverse Hopf product grids, or the FCC
    transports of T, O and I with kind="transport"; exact=True fills the base grid up to
    the count with centres of its largest empty caps).
    """

    def __init__(self, group="O", *, h=None, count=2048, quaternions=None, exact=False, identity=False,
                 fixed_identity=False, kind=None, _empty=False):
        self._lib = _library()
        self._handle = self._lib.zo_create(b"")
        if not self._handle:
            raise RuntimeError(self._lib.zo_create_error().decode(errors="replace"))
        self.history = []
        self._group = str(group).upper()
        try:
            if _empty:
                _group_id(group)
            elif quaternions is None:
                kind_id = _kind_id(kind)
                if kind_id == 2 and _group_id(group) < 8:
                    raise ValueError("transport grids exist for T, O and I")
                if h is None:
                    h = 6 if kind_id == 2 else 0
                elif not 1 <= int(h) <= (64 if kind_id == 2 else 512):
                    raise ValueError(f"h must be 1–{64 if kind_id == 2 else 512}")
                if not 1 <= int(count) <= 10_000_000:
                    raise ValueError("count must be 1–10,000,000")
                if exact:
                    self._check(self._lib.zo_generate_exact(self._handle, _group_id(group), int(count),
                                                            int(identity or fixed_identity) | kind_id << 8))
                else:
                    self._check(self._lib.zo_generate(self._handle, _group_id(group), int(h), int(count), 0, kind_id))
            else:
                self.set(quaternions, group)
            if identity or fixed_identity:
                self.identity(add=True, fixed=fixed_identity)
        except BaseException:
            self.close()
            raise

    def _check(self, code):
        if code != 0:
            raise RuntimeError(self._lib.zo_error(self._handle).decode(errors="replace"))

    def _open(self):
        if not self._handle:
            raise RuntimeError("Grid is closed")

    def close(self):
        if getattr(self, "_handle", None):
            self._lib.zo_destroy(self._handle)
            self._handle = None

    def __enter__(self):
        self._open()
        return self

    def __exit__(self, *_):
        self.close()

    def __del__(self):
        self.close()

    def __len__(self):
        self._open()
        return int(self._lib.zo_count(self._handle))

    @property
    def quaternions(self):
        self._open()
        n = len(self)
        if n == 0:
            return np.zeros((0, 4), np.float32)
        return np.ctypeslib.as_array(self._lib.zo_quaternions(self._handle), shape=(n, 4)).copy()

    def point_values(self, kind):
        """Per-orientation values, or None before the pass that makes them.

        "separation": angle to the nearest other orbit point (SO(3) degrees), from every
        energy pass and spacing().  "covering": radius of the largest empty cap the point's
        Voronoi cell reaches (degrees) and "faces": faces of that cell, from covering().
        """
        self._open()
        kinds = {"separation": 0, "covering": 1, "faces": 2}
        if kind not in kinds:
            raise ValueError(f"kind must be one of {', '.join(kinds)}")
        pointer = self._lib.zo_point_values(self._handle, kinds[kind])
        if not pointer:
            return None
        return np.ctypeslib.as_array(pointer, shape=(len(self),)).copy()

    def nearest_indices(self):
        """From the last spacing(): per point its nearest orbit point as (source, operator,
        sign) arrays, sign 0 for +q[source]*g[operator]; None before spacing()."""
        self._open()
        pointer = self._lib.zo_point_indices(self._handle)
        if not pointer:
            return None
        rank = np.ctypeslib.as_array(pointer, shape=(len(self),)).copy().astype(np.int64)
        order = int(self._lib.zo_order(self._handle))
        return rank // (2 * order), (rank % (2 * order)) // 2, rank % 2

    @property
    def operations(self):
        self._open()
        count = self._lib.zo_order(self._handle)
        return np.ctypeslib.as_array(self._lib.zo_operations(self._handle), shape=(count, 4)).copy()

    @property
    def stats(self):
        self._open()
        return json.loads(self._lib.zo_stats(self._handle))

    def set(self, quaternions, group=None):
        self._open()
        q = np.asarray(quaternions, dtype=np.float32, order="C")
        if q.ndim != 2 or q.shape[1] != 4:
            raise ValueError("quaternions must have shape (N,4), scalar first")
        group = group or self.stats["group"]
        self._check(self._lib.zo_set(self._handle, _group_id(group), q.ctypes.data_as(C.POINTER(C.c_float)), len(q)))
        self.history.clear()
        return self

    def measure(self):
        """Quey energy and separation; below the ratio limit also the Riesz, Green and log ratios."""
        self._open()
        self._check(self._lib.zo_measure(self._handle))
        return self.stats

    def spacing(self):
        """Each orientation's nearest orbit neighbour, point_values("separation"), and the
        separation, in time linear in the count; energy measures are left as they are."""
        self._open()
        self._check(self._lib.zo_spacing(self._handle))
        return self.stats

    def identity(self, *, add=False, fixed=False):
        """Optionally add the identity and hold it exactly fixed during relaxation."""
        self._open()
        self._check(self._lib.zo_identity(self._handle, int(add) | (int(fixed) << 1)))
        return self

    def covering(self):
        """Covering radius (stats covering_radius_deg): the largest angle from any
        orientation to the sample's nearest orbit point.  Reused while the sample is
        unchanged (stats covering_cached)."""
        self._open()
        self._check(self._lib.zo_cover(self._handle))
        return self.stats

    def triangulation(self):
        """The Delaunay tetrahedra around the sample's orbit points: stereographic xyz of the
        sample (first, in order) and of the other orbit images that close its cells, and
        tetrahedron indices.  Measures the covering radius too; the arrays are copies."""
        self._open()
        self._check(self._lib.zo_mesh(self._handle))
        n = self._lib.zo_mesh_vertex_count(self._handle)
        t = self._lib.zo_mesh_tet_count(self._handle)
        vertices = np.ctypeslib.as_array(self._lib.zo_mesh_vertices(self._handle), shape=(n, 4))[:, :3].copy()
        tetrahedra = np.ctypeslib.as_array(self._lib.zo_mesh_tets(self._handle), shape=(t, 4)).copy()
        return {"stereographic": vertices, "tetrahedra": tetrahedra, "stats": self.stats}

    def relax(self, steps=50, *, callback=None, neighbourhood=None):
        """Quey geodesic force relaxation with BB steps; updates this grid.

        callback(grid, stats) runs after each step; return False to stop.  Without a
        callback the steps run on the GPU in batches of up to 1,000, with one readback and
        one history entry per batch.  neighbourhood: forces only from orbit points within
        this many mean orientation radii (20 holds about 7.5k neighbours), or 0 for all
        pairs; it stays set, None keeps the current setting.
        """
        self._open()
        if int(steps) < 0:
            raise ValueError("steps must be nonnegative")
        if neighbourhood is not None:
            self._check(self._lib.zo_relax_neighbourhood(self._handle, float(neighbourhood)))
        remaining = int(steps)
        while remaining > 0:
            count = 1 if callback is not None else min(remaining, 1000)
            self._check(self._lib.zo_relax(self._handle, count))
            remaining -= count
            stats = self.stats
            self.history.append(stats)
            if callback is not None and callback(self, stats) is False:
                break
        return self

    def optimize_covering(self, steps=2000, *, mode="face", batch=100, callback=None, local=True, check=False,
                          local_rounds=None, min_separation_deg=None, live=True, patience=150, plateau_gain=5e-4,
                          kick=0.04, kick_trust=0.03):
        """Lower the covering radius; keeps the best sample (python/zxtalori/api.py:269-336).

        Each step pulls the points spanning the largest empty cap toward its centre ("tet";
        "face" takes half steps) and keeps the move when the measured covering radius
        drops.  After `patience` steps without a relative gain of `plateau_gain`, the best
        sample is restored and every point turned by an angle of `kick` times the best
        radius; steps restart at `kick_trust`.  min_separation_deg rejects moves that bring
        any pair nearer than that SO(3) angle.  callback(grid, stats) runs after each batch
        of steps; return False to stop.  local, check, live and local_rounds select the
        reference's local re-tessellation variants and are accepted for compatibility (every
        step here measures the whole sample).  Energy and separation are not measured:
        stats["measures_current"] is False until measure().
        """
        self._open()
        modes = {"tet": 0, "face": 1}
        floor = float(min_separation_deg or 0)
        if not 0 <= floor < 180:
            raise ValueError("min_separation_deg must be in [0, 180)")
        if local_rounds is not None and not 1 <= int(local_rounds) <= 255:
            raise ValueError("local_rounds must be 1-255")
        variant = (0 if local else 4) | (8 if check and local else 0) | (0 if live else 16) | (int(local_rounds or 0) << 16)
        if mode not in modes:
            raise ValueError(f"mode must be one of {', '.join(modes)}")
        if int(steps) < 0 or int(batch) < 1:
            raise ValueError("steps must be nonnegative and batch positive")
        if not (1 <= int(patience) <= 1000000 and 0 <= plateau_gain < 1 and 0 <= kick <= 1 and 0 < kick_trust <= 1):
            raise ValueError("patience must be 1-1000000, plateau_gain in [0, 1), kick in [0, 1] and kick_trust in (0, 1]")
        self._check(self._lib.zo_cover_kicks(self._handle, int(patience), float(plateau_gain), float(kick), float(kick_trust)))
        done = 0
        while done < int(steps):
            count = min(int(batch), int(steps) - done)
            done += count
            last = done >= int(steps)
            self._check(self._lib.zo_cover_optimize(self._handle, count, modes[mode] | variant, int(last), floor))
            stats = self.stats
            self.history.append(stats)
            if not last and callback is not None and callback(self, stats) is False:
                self._check(self._lib.zo_cover_optimize(self._handle, 0, modes[mode] | variant, 1, floor))
                break
        return self

    def save(self, path):
        """Save JSON shared with the browser (format zxtalori-1), or NPZ."""
        path = Path(path)
        stats = self.stats
        if path.suffix.lower() == ".npz":
            np.savez_compressed(path, quaternions=self.quaternions, metadata=json.dumps(stats))
        else:
            path.write_text(json.dumps({"format": "zxtalori-1", "group": stats["group"], "metadata": stats,
                                        "quaternions": self.quaternions.tolist()}, indent=2) + "\n", encoding="utf-8")
        return path

    # ---- development hooks (src/native_api.zig:31-43) -------------------------------
    def profile(self, enabled=True):
        return bool(self._lib.zo_profile(self._handle, bool(enabled)))

    def profile_report(self):
        """{kernel: (launches, total ms)} of the launches timed since profiling began."""
        size = self._lib.zo_profile_report(self._handle, None, 0)
        buf = C.create_string_buffer(size + 1)
        self._lib.zo_profile_report(self._handle, buf, size + 1)
        out = {}
        for line in buf.value.decode().splitlines():
            name, count, ms = line.split("\t")
            out[name] = (int(count), float(ms))
        return out

    def pair_budget(self, scale):
        self._lib.zo_pair_budget(self._handle, float(scale))

    def force_passes(self, passes=1):
        self._check(self._lib.zo_force_passes(self._handle, int(passes)))

    def force_rows(self):
        """One force pass of the current neighbourhood: (force (N,4) fx fy fz curvature,
        energy (N,) float64, nearest chord^2 (N,))."""
        n = len(self)
        force = np.empty((n, 4), np.float32)
        energy = np.empty(n, np.float64)
        nearest = np.empty(n, np.float32)
        self._check(self._lib.zo_force_rows(self._handle, force.ctypes.data, energy.ctypes.data,
                                            nearest.ctypes.data))
        return force, energy, nearest

    def near_info(self):
        out = (C.c_uint32 * 3)()
        self._lib.zo_near_info(self._handle, out)
        pairs = (C.c_uint64 * 2)()
        self._lib.zo_near_pairs(self._handle, pairs)
        cut = (C.c_float * 3)()
        self._lib.zo_near_cut(self._handle, cut)
        return {"sites": out[0], "shell": out[1], "sdim": out[2], "pairs": pairs[0], "candidates": pairs[1],
                "chord2": cut[0], "radius": cut[1], "atan_terms": int(cut[2])}


def hopf_points(group, f):
    """The frequency-f inverse Hopf grid of `group` before the zone fold, (N,4) float32 in the
    reference's row order, computed on the host (no device)."""
    lib = _library()
    gid = _group_id(group)
    n = lib.zo_hopf_points(gid, int(f), None, 0)
    if n == 0:
        raise ValueError(f"no inverse Hopf grid for {group} at frequency {f} (1-512, at most 10,000,000 points)")
    out = np.empty((n, 4), np.float32)
    if lib.zo_hopf_points(gid, int(f), out.ctypes.data, n) != n:
        raise RuntimeError("inverse Hopf grid construction failed")
    return out


def hopf_frequency(group, count=2048):
    """The frequency generate() takes for `count` when h is not given."""
    return int(_library().zo_hopf_frequency(_group_id(group), int(count)))


def generate(group="O", *, h=None, count=2048, exact=False, identity=False, fixed_identity=False, kind=None):
    return Grid(group, h=h, count=count, exact=exact, identity=identity, fixed_identity=fixed_identity, kind=kind)


def from_quaternions(quaternions, group="O"):
    return Grid(group, quaternions=quaternions)


def load(path):
    path = Path(path)
    if path.suffix.lower() == ".npz":
        with np.load(path, allow_pickle=False) as data:
            metadata = json.loads(str(data["metadata"]))
            return from_quaternions(data["quaternions"], metadata["group"])
    data = json.loads(path.read_text(encoding="utf-8"))
    return from_quaternions(data["quaternions"], data["group"])


def haar(group="O", count=1000, seed=193):
    """Seeded GPU Haar-uniform random orientations folded into SO(3)/K."""
    grid = Grid(group, _empty=True)
    try:
        grid._check(grid._lib.zo_random(grid._handle, _group_id(group), int(count), int(seed)))
        return grid
    except BaseException:
        grid.close()
        raise
