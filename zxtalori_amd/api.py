"""ctypes surface over libzxtalori.so (the gfx950 build).

The public names, argument meanings, error texts and the stats record are those of the
reference's Python adapter (python/zxtalori/api.py), so that code written against it runs
unchanged: Grid, generate, from_quaternions, haar and load; RuntimeError carrying the
engine's error text; the zxtalori-1 JSON and NPZ sample files.  Everything behind those
names is this build's: the library is resolved from ZXTALORI_LIBRARY or the in-tree build,
and every call goes to the HIP engine.
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
KINDS = (None, "hopf", "transport")
HERE = Path(__file__).resolve().parent
_LIB = None

# ---- the export table ------------------------------------------------------------
# One line per entry: name, argument codes, result code.  Codes: e engine or batch handle,
# u uint32, i int32, q uint64, f float, b bool, z size_t, s char*, p void*, F float*,
# U uint32*, Q uint64*; "-" no arguments, "v" no result.  The first block is the table the
# reference's adapter binds at load time (src/exports.zig, src/native_api.zig); the rest
# are this build's development hooks and batch entries.
_EXPORTS = """
zo_create s e;  zo_create_error - s;  zo_destroy e v;  zo_version - s
zo_generate euuuuu i;  zo_random euuu i;  zo_set euFu i;  zo_count e u;  zo_order e u
zo_quaternions e F;  zo_operations e F;  zo_measure e i;  zo_spacing e i;  zo_relax eu i
zo_relax_neighbourhood ef i;  zo_point_values eu F;  zo_identity eu i;  zo_cover e i
zo_cover_optimize euuuf i;  zo_cover_kicks eufff i;  zo_mesh e i;  zo_mesh_vertices e F
zo_mesh_tets e U;  zo_mesh_vertex_count e u;  zo_mesh_tet_count e u;  zo_generate_exact euuu i
zo_stats e s;  zo_error e s;  zo_cancel e v
zo_profile eb b;  zo_profile_report esz z;  zo_pair_budget ef v;  zo_tune_allpairs uu v
zo_point_indices e U;  zo_force_passes eu i;  zo_near_info eU i;  zo_force_rows eppp i
zo_near_pairs eQ i;  zo_near_cut eF i;  zo_near_squeeze ei i;  zo_near_resizes e u
zo_hopf_points uupu u;  zo_hopf_frequency uu u;  zo_issue_probe uup i
zo_batch_create u e;  zo_batch_destroy e v;  zo_batch_error e s;  zo_batch_add eupuufuuQ i
zo_batch_step eU i;  zo_batch_advance euUU i;  zo_batch_status eq i;  zo_batch_job_error eq s;  zo_batch_stats eq s
zo_batch_quaternions eq F;  zo_batch_count eq u;  zo_batch_point_values equ F
zo_batch_point_indices eq U;  zo_batch_bytes eq q;  zo_batch_remove eq i;  zo_batch_running e u
zo_batch_wait euU i;  zo_batch_graphs eQ i;  zo_batch_near_squeeze eqi i;  zo_batch_near_resizes eq u
"""
_CTYPES = {"e": C.c_void_p, "u": C.c_uint32, "i": C.c_int32, "q": C.c_uint64, "f": C.c_float, "b": C.c_bool,
           "z": C.c_size_t, "s": C.c_char_p, "p": C.c_void_p, "F": C.POINTER(C.c_float),
           "U": C.POINTER(C.c_uint32), "Q": C.POINTER(C.c_uint64), "v": None}


def _exports():
    for line in _EXPORTS.strip().splitlines():
        for entry in filter(None, (e.strip() for e in line.split(";"))):
            name, args, result = entry.split()
            yield name, [] if args == "-" else [_CTYPES[c] for c in args], _CTYPES[result]


def library_path():
    override = os.environ.get("ZXTALORI_LIBRARY")
    return Path(override) if override else HERE / "_lib" / "libzxtalori.so"


def _library():
    """The loaded library with every export's argument and result types set (loaded once)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not path.is_file():
            raise RuntimeError(f"zxtalori engine missing at {path}: build it with `python -m zxtalori_amd.build`.")
        lib = C.CDLL(str(path))
        for name, args, result in _exports():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = result
        _LIB = lib
    return _LIB


def _group_id(group):
    name = str(group).upper()
    if name not in EXTENDED_GROUPS:
        raise ValueError(f"group must be one of {', '.join(GROUPS)}")
    return EXTENDED_GROUPS.index(name)


def _kind_id(kind):
    name = None if kind is None else str(kind).lower()
    if name not in KINDS:
        raise ValueError("kind must be 'hopf' or 'transport'")
    return KINDS.index(name)


def _rows(pointer, shape, dtype=None):
    """A copy of engine-owned memory (valid only until the engine's next mutating call)."""
    view = np.ctypeslib.as_array(pointer, shape=shape)
    return view.astype(dtype) if dtype is not None else view.copy()


def _generation_args(group, kind, h, count):
    """(group id, kind id, frequency or semi-edge) after the adapter's argument checks."""
    gid, kid = _group_id(group), _kind_id(kind)
    transport = kid == 2
    if transport and gid < 8:
        raise ValueError("transport grids exist for T, O and I")
    top = 64 if transport else 512
    if h is None:
        h = 6 if transport else 0
    elif not 1 <= int(h) <= top:
        raise ValueError(f"h must be 1–{top}")
    if not 1 <= int(count) <= 10_000_000:
        raise ValueError("count must be 1–10,000,000")
    return gid, kid, int(h)


def _covering_flags(mode, steps, batch, local, check, local_rounds, min_separation_deg, live, patience,
                    plateau_gain, kick, kick_trust):
    """(mode word, separation floor) for zo_cover_optimize, after the adapter's checks."""
    floor = float(min_separation_deg or 0)
    problems = [
        (not 0 <= floor < 180, "min_separation_deg must be in [0, 180)"),
        (local_rounds is not None and not 1 <= int(local_rounds) <= 255, "local_rounds must be 1-255"),
        (mode not in ("tet", "face"), "mode must be one of tet, face"),
        (int(steps) < 0 or int(batch) < 1, "steps must be nonnegative and batch positive"),
        (not (1 <= int(patience) <= 1_000_000 and 0 <= plateau_gain < 1 and 0 <= kick <= 1 and 0 < kick_trust <= 1),
         "patience must be 1-1000000, plateau_gain in [0, 1), kick in [0, 1] and kick_trust in (0, 1]"),
    ]
    for failed, message in problems:
        if failed:
            raise ValueError(message)
    word = ("tet", "face").index(mode)
    word |= 0 if local else 4
    word |= 8 if (check and local) else 0
    word |= 0 if live else 16
    word |= int(local_rounds or 0) << 16
    return word, floor


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
        try:
            self._fill(group, h, count, quaternions, exact, identity or fixed_identity, kind, _empty)
            if identity or fixed_identity:
                self.identity(add=True, fixed=fixed_identity)
        except BaseException:
            self.close()
            raise

    def _fill(self, group, h, count, quaternions, exact, with_identity, kind, empty):
        if empty:
            _group_id(group)
        elif quaternions is not None:
            self.set(quaternions, group)
        else:
            gid, kid, h = _generation_args(group, kind, h, count)
            if exact:
                self._check(self._lib.zo_generate_exact(self._handle, gid, int(count), int(with_identity) | kid << 8))
            else:
                self._check(self._lib.zo_generate(self._handle, gid, h, int(count), 0, kid))

    # ---- handle -----------------------------------------------------------------
    def _check(self, code):
        if code != 0:
            raise RuntimeError(self._lib.zo_error(self._handle).decode(errors="replace"))

    def _open(self):
        if not self._handle:
            raise RuntimeError("Grid is closed")

    def _call(self, name, *args):
        """Calls a status-returning export on this engine, raising its error text."""
        self._open()
        self._check(getattr(self._lib, name)(self._handle, *args))
        return self

    def close(self):
        handle = getattr(self, "_handle", None)
        if handle:
            self._handle = None
            self._lib.zo_destroy(handle)

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

    # ---- views ------------------------------------------------------------------
    @property
    def quaternions(self):
        n = len(self)
        return _rows(self._lib.zo_quaternions(self._handle), (n, 4)) if n else np.zeros((0, 4), np.float32)

    @property
    def operations(self):
        self._open()
        return _rows(self._lib.zo_operations(self._handle), (int(self._lib.zo_order(self._handle)), 4))

    @property
    def stats(self):
        self._open()
        return json.loads(self._lib.zo_stats(self._handle))

    def point_values(self, kind):
        """Per-orientation values, or None before the pass that makes them.

        "separation": angle to the nearest other orbit point (SO(3) degrees), from every
        energy pass and spacing().  "covering": radius of the largest empty cap the point's
        Voronoi cell reaches (degrees) and "faces": faces of that cell, from covering().
        """
        self._open()
        names = ("separation", "covering", "faces")
        if kind not in names:
            raise ValueError(f"kind must be one of {', '.join(names)}")
        pointer = self._lib.zo_point_values(self._handle, names.index(kind))
        return _rows(pointer, (len(self),)) if pointer else None

    def nearest_indices(self):
        """From the last spacing(): per point its nearest orbit point as (source, operator,
        sign) arrays, sign 0 for +q[source]*g[operator]; None before spacing()."""
        self._open()
        pointer = self._lib.zo_point_indices(self._handle)
        if not pointer:
            return None
        code = _rows(pointer, (len(self),), np.int64)
        source, image = np.divmod(code, 2 * int(self._lib.zo_order(self._handle)))
        return source, image >> 1, image & 1

    # ---- operations -------------------------------------------------------------
    def set(self, quaternions, group=None):
        q = np.ascontiguousarray(quaternions, dtype=np.float32)
        if q.ndim != 2 or q.shape[1] != 4:
            raise ValueError("quaternions must have shape (N,4), scalar first")
        gid = _group_id(group or self.stats["group"])
        self._call("zo_set", gid, q.ctypes.data_as(C.POINTER(C.c_float)), len(q))
        self.history.clear()
        return self

    def measure(self):
        """Quey energy and separation; below the ratio limit also the Riesz, Green and log ratios."""
        return self._call("zo_measure").stats

    def spacing(self):
        """Each orientation's nearest orbit neighbour, point_values("separation"), and the
        separation, in time linear in the count; energy measures are left as they are."""
        return self._call("zo_spacing").stats

    def identity(self, *, add=False, fixed=False):
        """Optionally add the identity and hold it exactly fixed during relaxation."""
        return self._call("zo_identity", (1 if add else 0) | (2 if fixed else 0))

    def covering(self):
        """Covering radius (stats covering_radius_deg): the largest angle from any
        orientation to the sample's nearest orbit point.  Reused while the sample is
        unchanged (stats covering_cached)."""
        return self._call("zo_cover").stats

    def triangulation(self):
        """The Delaunay tetrahedra around the sample's orbit points: stereographic xyz of the
        sample (first, in order) and of the other orbit images that close its cells, and
        tetrahedron indices.  Measures the covering radius too; the arrays are copies."""
        self._call("zo_mesh")
        lib, h = self._lib, self._handle
        vertices = _rows(lib.zo_mesh_vertices(h), (lib.zo_mesh_vertex_count(h), 4))[:, :3].copy()
        tetrahedra = _rows(lib.zo_mesh_tets(h), (lib.zo_mesh_tet_count(h), 4))
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
        total = int(steps)
        if total < 0:
            raise ValueError("steps must be nonnegative")
        if neighbourhood is not None:
            self._call("zo_relax_neighbourhood", float(neighbourhood))
        chunk = 1 if callback is not None else 1000
        for first in range(0, total, chunk):
            self._call("zo_relax", min(chunk, total - first))
            record = self.stats
            self.history.append(record)
            if callback is not None and callback(self, record) is False:
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
        word, floor = _covering_flags(mode, steps, batch, local, check, local_rounds, min_separation_deg, live,
                                      patience, plateau_gain, kick, kick_trust)
        self._call("zo_cover_kicks", int(patience), float(plateau_gain), float(kick), float(kick_trust))
        total, per_call = int(steps), int(batch)
        for first in range(0, total, per_call):
            count = min(per_call, total - first)
            final = first + count >= total
            self._call("zo_cover_optimize", count, word, int(final), floor)
            record = self.stats
            self.history.append(record)
            if not final and callback is not None and callback(self, record) is False:
                self._call("zo_cover_optimize", 0, word, 1, floor)  # keep the best sample
                break
        return self

    def save(self, path):
        """Save JSON shared with the browser (format zxtalori-1), or NPZ."""
        path = Path(path)
        writer = _write_npz if path.suffix.lower() == ".npz" else _write_json
        writer(path, self.quaternions, self.stats)
        return path

    # ---- development hooks (src/native_api.zig:31-43) -------------------------------
    def profile(self, enabled=True):
        return bool(self._lib.zo_profile(self._handle, bool(enabled)))

    def profile_report(self):
        """{kernel: (launches, total ms)} of the launches timed since profiling began."""
        size = self._lib.zo_profile_report(self._handle, None, 0)
        text = C.create_string_buffer(size + 1)
        self._lib.zo_profile_report(self._handle, text, size + 1)
        report = {}
        for line in filter(None, text.value.decode().splitlines()):
            name, launches, ms = line.split("\t")
            report[name] = (int(launches), float(ms))
        return report

    def pair_budget(self, scale):
        self._lib.zo_pair_budget(self._handle, float(scale))

    def force_passes(self, passes=1):
        self._call("zo_force_passes", int(passes))

    def force_rows(self):
        """One force pass of the current neighbourhood: (force (N,4) fx fy fz curvature,
        energy (N,) float64, nearest chord^2 (N,))."""
        n = len(self)
        out = (np.empty((n, 4), np.float32), np.empty(n, np.float64), np.empty(n, np.float32))
        self._call("zo_force_rows", *(a.ctypes.data for a in out))
        return out

    def near_info(self):
        self._open()
        sizes, pairs, cut = (C.c_uint32 * 3)(), (C.c_uint64 * 2)(), (C.c_float * 3)()
        self._lib.zo_near_info(self._handle, sizes)
        self._lib.zo_near_pairs(self._handle, pairs)
        self._lib.zo_near_cut(self._handle, cut)
        info = dict(zip(("sites", "shell", "sdim"), sizes))
        info.update(zip(("pairs", "candidates"), pairs))
        info.update(chord2=cut[0], radius=cut[1], atan_terms=int(cut[2]))
        return info


# ---- sample files (the zxtalori-1 JSON and the NPZ of python/zxtalori/api.py) --------
def _write_npz(path, quaternions, stats):
    np.savez_compressed(path, quaternions=quaternions, metadata=json.dumps(stats))


def _write_json(path, quaternions, stats):
    record = {"format": "zxtalori-1", "group": stats["group"], "metadata": stats,
              "quaternions": quaternions.tolist()}
    path.write_text(json.dumps(record, indent=2) + "\n", encoding="utf-8")


def _read_sample(path):
    """(quaternions, group) of a saved sample."""
    if path.suffix.lower() == ".npz":
        with np.load(path, allow_pickle=False) as data:
            return data["quaternions"], json.loads(str(data["metadata"]))["group"]
    record = json.loads(path.read_text(encoding="utf-8"))
    return record["quaternions"], record["group"]


# ---- constructors --------------------------------------------------------------------
def generate(group="O", *, h=None, count=2048, exact=False, identity=False, fixed_identity=False, kind=None):
    return Grid(group, h=h, count=count, exact=exact, identity=identity, fixed_identity=fixed_identity, kind=kind)


def from_quaternions(quaternions, group="O"):
    return Grid(group, quaternions=quaternions)


def load(path):
    quaternions, group = _read_sample(Path(path))
    return from_quaternions(quaternions, group)


def haar(group="O", count=1000, seed=193):
    """Seeded GPU Haar-uniform random orientations folded into SO(3)/K."""
    grid = Grid(group, _empty=True)
    try:
        grid._call("zo_random", _group_id(group), int(count), int(seed))
    except BaseException:
        grid.close()
        raise
    return grid


def hopf_points(group, f):
    """The frequency-f inverse Hopf grid of `group` before the zone fold, (N,4) float32 in the
    reference's row order, computed on the host (no device)."""
    lib, gid = _library(), _group_id(group)
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
