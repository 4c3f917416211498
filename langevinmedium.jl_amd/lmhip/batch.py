"""Many independent 2+1D runs advanced together (extension).

`Batch` wraps liblmhip's lmh_batch_*: media and tables are added once and referenced by index
from the runs, so runs that share a medium or a table (e.g. the Juttner table of one mass over
a common temperature range) hold it once on the device.  Run r draws its Philox noise with key
(seed, stream + r): it reproduces the run alone (`Context.run` with stream + r) bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi
from ._ffi import dptr
from .context import LMHError


class Batch:
    def __init__(self, device: int = 0):
        self._lib = _ffi.lib()
        h = ctypes.c_void_p()
        rc = self._lib.lmh_batch_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise LMHError(rc, "lmh_batch_create failed")
        self._h = h
        self.N = 0
        self.pdim = 0
        self.nruns = 0
        self.counts = []

    def close(self):
        if self._h:
            self._lib.lmh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != 0:
            raise LMHError(rc, self._lib.lmh_batch_last_error(self._h).decode())

    @property
    def stream(self) -> int:
        return int(self._lib.lmh_batch_stream(self._h) or 0)

    def add_medium(self, g) -> int:
        idx = ctypes.c_int32(-1)
        T, vx, vy = (np.asfortranarray(a, dtype=np.float64) for a in (g.T, g.vx, g.vy))
        tg = np.ascontiguousarray(g.tgrid, dtype=np.float64)
        self._check(self._lib.lmh_batch_add_medium_grid2p1(self._h, g.x0, g.dx, g.nx, g.y0, g.dy, g.ny,
                                                           dptr(tg), tg.size, dptr(T), dptr(vx), dptr(vy),
                                                           ctypes.byref(idx)))
        return idx.value

    def add_time_table(self, Tmin, invdT, vals) -> int:
        idx = ctypes.c_int32(-1)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self._check(self._lib.lmh_batch_add_time_table(self._h, float(Tmin), float(invdT), dptr(v), v.size,
                                                       ctypes.byref(idx)))
        return idx.value

    def add_juttner_table(self, invcdf, nU, nT, Tmin, invdT) -> int:
        idx = ctypes.c_int32(-1)
        v = np.ascontiguousarray(invcdf, dtype=np.float64)
        self._check(self._lib.lmh_batch_add_juttner_table(self._h, dptr(v), int(nU), int(nT), float(Tmin),
                                                          float(invdT), ctypes.byref(idx)))
        return idx.value

    def set_runs(self, runs, pos, mom):
        """runs: dicts with count, mass, DsT, medium and optional drag_table, rta_table,
        juttner_table; pos (2, N) and mom (pdim, N), run after run."""
        arr = (_ffi.lmh_run_desc * len(runs))()
        for k, r in enumerate(runs):
            arr[k] = _ffi.lmh_run_desc(int(r["count"]), float(r["mass"]), float(r.get("DsT", 0.0)),
                                       int(r["medium"]), int(r.get("drag_table", -1)),
                                       int(r.get("rta_table", -1)), int(r.get("juttner_table", -1)))
        pos = np.asfortranarray(pos, dtype=np.float64)
        mom = np.asfortranarray(mom, dtype=np.float64)
        N = sum(int(r["count"]) for r in runs)
        if pos.shape != (2, N) or mom.ndim != 2 or mom.shape[1] != N:
            raise ValueError(f"pos must be (2, {N}) and mom (pdim, {N})")
        self._check(self._lib.lmh_batch_set_runs(self._h, arr, len(runs), dptr(pos), dptr(mom), mom.shape[0]))
        self.N, self.pdim, self.nruns = N, mom.shape[0], len(runs)
        self.counts = [int(r["count"]) for r in runs]

    def get_particles(self):
        pos = np.empty((2, self.N), dtype=np.float64, order="F")
        mom = np.empty((self.pdim, self.N), dtype=np.float64, order="F")
        self._check(self._lib.lmh_batch_get_particles(self._h, dptr(pos), dptr(mom)))
        return pos, mom

    def set_tuning(self, block=256, steps_per_launch=1, persistent=0, graph_steps=0, sort_every=0, slice_min=0):
        t = _ffi.lmh_batch_tuning(int(block), int(steps_per_launch), int(persistent), int(graph_steps),
                                  int(sort_every), int(slice_min))
        self._check(self._lib.lmh_batch_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self) -> dict:
        t = _ffi.lmh_batch_tuning()
        self._check(self._lib.lmh_batch_get_tuning(self._h, ctypes.byref(t)))
        return {"block": t.block, "steps_per_launch": t.steps_per_launch, "persistent": t.persistent,
                "graph_steps": t.graph_steps, "sort_every": t.sort_every, "slice_min": t.slice_min}

    def boost_to_lab(self, params, step: int = 0):
        self._check(self._lib.lmh_batch_boost_to_lab(self._h, ctypes.byref(params), int(step)))

    def run(self, params, step_begin: int, nsteps: int):
        self._check(self._lib.lmh_batch_run(self._h, ctypes.byref(params), int(step_begin), int(nsteps)))

    def autotune(self, params, trial_steps: int = 4):
        t = _ffi.lmh_batch_tuning()
        ms = ctypes.c_double(0.0)
        self._check(self._lib.lmh_batch_autotune(self._h, ctypes.byref(params), int(trial_steps),
                                                 ctypes.byref(t), ctypes.byref(ms)))
        return {"block": t.block, "steps_per_launch": t.steps_per_launch, "persistent": t.persistent,
                "graph_steps": t.graph_steps, "sort_every": t.sort_every, "slice_min": t.slice_min}, ms.value

    def observables(self, nbins=40, pt_min=0.0, pt_max=10.0, sigma=0.0):
        spec = _ffi.lmh_obs_spec(int(nbins), -1, float(pt_min), float(pt_max), float(sigma))
        out = np.zeros((self.nruns, 5 + int(nbins)), dtype=np.float64)
        self._check(self._lib.lmh_batch_observables(self._h, ctypes.byref(spec), dptr(out)))
        return out

    def nonfinite_step(self) -> int:
        return int(self._lib.lmh_batch_nonfinite_step(self._h))
