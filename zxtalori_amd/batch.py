"""Many samples relaxed together on one GPU (zo_batch_* in libzxtalori.so).

The reference runs one sample per engine, one synchronous call at a time
(src/native_api.zig:9-19, python/zxtalori/api.py:95-122); this module adds a batch of
jobs whose per-step launches are shared: all all-pairs force passes, all BB steps and all
folds of the running relax jobs go out as one launch per stage, and steady-state steps
replay from a captured HIP graph.  A job's arithmetic is the one a Grid runs alone, so
its positions, stats and per-point values are bit-identical alone and in any mix.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from .api import _group_id, _library

KINDS = {"relax": 0, "measure": 1, "spacing": 2}
STATES = {0: "running", 1: "done", 2: "failed"}


class Batch:
    """A set of jobs advanced together by step().  Close it to release the device."""

    def __init__(self, *, graphs=True, streams=3):
        self._lib = _library()
        self._handle = self._lib.zo_batch_create((1 if graphs else 0) | (min(int(streams), 8) << 8))
        if not self._handle:
            raise RuntimeError(self._lib.zo_create_error().decode(errors="replace"))

    def _open(self):
        if not self._handle:
            raise RuntimeError("batch is closed")

    def _check(self, code):
        if code != 0:
            raise RuntimeError(self._lib.zo_batch_error(self._handle).decode(errors="replace"))

    def close(self):
        if getattr(self, "_handle", None):
            self._lib.zo_batch_destroy(self._handle)
            self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, group="O", *, quaternions=None, count=None, seed=0, kind="relax", steps=50, neighbourhood=0.0):
        """Adds a job and returns its id.

        quaternions (N,4) scalar first, or count seeded Haar orientations (as haar() draws
        them).  kind: "relax" (steps BB steps, at most 1,000, then the end-of-relax
        measure), "measure" or "spacing".  neighbourhood as Grid.relax takes it.
        """
        self._open()
        if kind not in KINDS:
            raise ValueError(f"kind must be one of {', '.join(KINDS)}")
        if int(steps) < 0:
            raise ValueError("steps must be nonnegative")
        q = None
        if quaternions is not None:
            q = np.ascontiguousarray(quaternions, dtype=np.float32)
            if q.ndim != 2 or q.shape[1] != 4:
                raise ValueError("quaternions must have shape (N,4), scalar first")
            n, pointer = len(q), q.ctypes.data
        else:
            if count is None or not 1 <= int(count) <= 10_000_000:
                raise ValueError("count must be 1–10,000,000")
            n, pointer = int(count), None
        job = C.c_uint64(0)
        self._check(self._lib.zo_batch_add(self._handle, _group_id(group), pointer, n, int(seed) & 0xFFFFFFFF,
                                           float(neighbourhood), KINDS[kind], int(steps), C.byref(job)))
        return int(job.value)

    def step(self):
        """One step of every running job; returns how many jobs ended in it."""
        self._open()
        done = C.c_uint32(0)
        self._check(self._lib.zo_batch_step(self._handle, C.byref(done)))
        return int(done.value)

    def advance(self, most=64):
        """Steps every running job: up to `most` steps that end no job go as one launch of a
        captured graph (with graph replay on and the same jobs running).  Returns (jobs
        ended, steps taken)."""
        self._open()
        done, taken = C.c_uint32(0), C.c_uint32(0)
        self._check(self._lib.zo_batch_advance(self._handle, int(most), C.byref(done), C.byref(taken)))
        return int(done.value), int(taken.value)

    def run(self, most=64):
        """Steps until no job is running, spans of steps that end no job as single graph
        launches."""
        while self.running:
            self.advance(most)
        return self

    def wait(self, most=1):
        """Blocks until at most `most` steps are unfinished on the device; returns how many are."""
        self._open()
        left = C.c_uint32(0)
        self._check(self._lib.zo_batch_wait(self._handle, int(most), C.byref(left)))
        return int(left.value)

    @property
    def running(self):
        self._open()
        return int(self._lib.zo_batch_running(self._handle))

    def status(self, job):
        self._open()
        code = self._lib.zo_batch_status(self._handle, int(job))
        if code < 0:
            raise KeyError(job)
        return STATES[code]

    def error(self, job):
        self._open()
        return self._lib.zo_batch_job_error(self._handle, int(job)).decode(errors="replace")

    def stats(self, job):
        self._open()
        return json.loads(self._lib.zo_batch_stats(self._handle, int(job)))

    def quaternions(self, job):
        self._open()
        n = int(self._lib.zo_batch_count(self._handle, int(job)))
        pointer = self._lib.zo_batch_quaternions(self._handle, int(job))
        if n == 0 or not pointer:
            return np.zeros((0, 4), np.float32)
        return np.ctypeslib.as_array(pointer, shape=(n, 4)).copy()

    def point_values(self, job, kind="separation"):
        self._open()
        kinds = {"separation": 0, "covering": 1, "faces": 2}
        if kind not in kinds:
            raise ValueError(f"kind must be one of {', '.join(kinds)}")
        n = int(self._lib.zo_batch_count(self._handle, int(job)))
        pointer = self._lib.zo_batch_point_values(self._handle, int(job), kinds[kind])
        if not pointer:
            return None
        return np.ctypeslib.as_array(pointer, shape=(n,)).copy()

    def point_indices(self, job):
        """Per point the rank source*2*order + 2*operator + sign of its nearest orbit point
        (after a spacing job), or None."""
        self._open()
        n = int(self._lib.zo_batch_count(self._handle, int(job)))
        pointer = self._lib.zo_batch_point_indices(self._handle, int(job))
        if not pointer:
            return None
        return np.ctypeslib.as_array(pointer, shape=(n,)).copy()

    def device_bytes(self, job):
        self._open()
        return int(self._lib.zo_batch_bytes(self._handle, int(job)))

    def remove(self, job):
        self._open()
        self._check(self._lib.zo_batch_remove(self._handle, int(job)))

    @property
    def graphs(self):
        """(steps captured into a graph, steps replayed from one)."""
        self._open()
        out = (C.c_uint64 * 2)()
        self._lib.zo_batch_graphs(self._handle, out)
        return int(out[0]), int(out[1])
