"""A device context: particle state, medium and tables resident in HBM behind liblmhip.so.

Arrays follow the reference's Julia layout: particle matrices are (rows, N), medium matrices
(n_r, n_t) and grid fields (nx, ny, nt); they are handed to the library in column-major order.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi
from ._ffi import dptr


class LMHError(RuntimeError):
    """A non-zero status of liblmhip.so, carrying lmh_last_error()."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"liblmhip status {code}: {msg}")
        self.code = code


def _f64(a, order="F"):
    return np.require(np.asarray(a, dtype=np.float64), dtype=np.float64, requirements=[order, "A"])


class Context:
    def __init__(self, device: int = 0):
        self._lib = _ffi.lib()
        h = ctypes.c_void_p()
        rc = self._lib.lmh_create(ctypes.byref(h), int(device))
        if rc != _ffi.LMH_OK:
            raise LMHError(rc, f"lmh_create(device={device}) failed")
        self._h = h
        self.dims = 0
        self.pdim = 0
        self.N = 0
        self.counts = []

    # ---- lifetime ----
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lmh_destroy(self._h)
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
        if rc != _ffi.LMH_OK:
            msg = self._lib.lmh_last_error(self._h)
            raise LMHError(rc, msg.decode() if msg else "")

    @property
    def stream(self) -> int:
        return int(self._lib.lmh_stream(self._h) or 0)

    def synchronize(self):
        self._check(self._lib.lmh_synchronize(self._h))

    # ---- launch shape ----
    def set_tuning(self, block=256, steps_per_launch=1, slice=1, defer=0, specialize=1, sort_every=64):
        t = _ffi.lmh_tuning(int(block), int(steps_per_launch), int(slice), int(defer), int(specialize),
                            int(sort_every))
        self._check(self._lib.lmh_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self) -> dict:
        t = _ffi.lmh_tuning()
        self._check(self._lib.lmh_get_tuning(self._h, ctypes.byref(t)))
        return {"block": t.block, "steps_per_launch": t.steps_per_launch, "slice": t.slice,
                "defer": t.defer, "specialize": t.specialize, "sort_every": t.sort_every}

    # ---- medium ----
    def set_medium_radial(self, rgrid, tgrid, T, v, v2=None):
        rg, tg = _f64(rgrid).ravel(), _f64(tgrid).ravel()
        T, v = _f64(T), _f64(v)
        if T.shape != (rg.size, tg.size) or v.shape != T.shape:
            raise ValueError(f"medium matrices must be (n_r, n_t) = {(rg.size, tg.size)}")
        v2a = None
        if v2 is not None:
            v2a = _f64(v2)
            if v2a.shape != T.shape:
                raise ValueError("V2Evolution must have the shape of T")
        self._keep = (rg, tg, T, v, v2a)
        self._check(self._lib.lmh_set_medium_radial(self._h, dptr(rg), rg.size, dptr(tg), tg.size,
                                                     dptr(T), dptr(v), dptr(v2a)))

    def set_medium_grid2p1(self, x0, dx, nx, y0, dy, ny, tgrid, T, vx, vy):
        tg = _f64(tgrid).ravel()
        arrs = [_f64(a) for a in (T, vx, vy)]
        for a in arrs:
            if a.shape != (nx, ny, tg.size):
                raise ValueError(f"grid fields must be (nx, ny, nt) = {(nx, ny, tg.size)}")
        self._check(self._lib.lmh_set_medium_grid2p1(self._h, float(x0), float(dx), int(nx), float(y0),
                                                      float(dy), int(ny), dptr(tg), tg.size,
                                                      dptr(arrs[0]), dptr(arrs[1]), dptr(arrs[2])))

    # ---- tables ----
    def set_time_table(self, which: int, species: int, Tmin: float, invdT: float, vals):
        v = _f64(vals).ravel()
        self._check(self._lib.lmh_set_time_table(self._h, int(which), int(species), float(Tmin),
                                                  float(invdT), dptr(v), v.size))

    def set_juttner_table(self, species: int, invcdf, nU: int, nT: int, Tmin: float, invdT: float):
        v = _f64(invcdf).ravel()
        if v.size != nU * nT:
            raise ValueError("Juttner table must hold nU*nT values")
        self._check(self._lib.lmh_set_juttner_table(self._h, int(species), dptr(v), int(nU), int(nT),
                                                     float(Tmin), float(invdT)))

    # ---- particles ----
    def set_particles(self, pos, mom, masses, counts=None):
        pos, mom = _f64(pos), _f64(mom)
        if pos.ndim != 2 or mom.ndim != 2 or pos.shape[1] != mom.shape[1]:
            raise ValueError("positions (dims, N) and momenta (pdim, N) must share N")
        dims, N = pos.shape
        pdim = mom.shape[0]
        masses = np.atleast_1d(np.asarray(masses, dtype=np.float64))
        if counts is None:
            if masses.size != 1:
                raise ValueError("species counts are needed with several masses")
            counts = [N]
        cnt = np.ascontiguousarray(np.asarray(counts, dtype=np.int64))
        masses = np.ascontiguousarray(masses)
        self._check(self._lib.lmh_set_particles(self._h, dptr(pos), dims, dptr(mom), pdim, N,
                                                 cnt.size, cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                 dptr(masses)))
        self.dims, self.pdim, self.N = dims, pdim, N
        self.counts = [int(c) for c in cnt]

    def get_particles(self):
        pos = np.empty((self.dims, self.N), dtype=np.float64, order="F")
        mom = np.empty((self.pdim, self.N), dtype=np.float64, order="F")
        self._check(self._lib.lmh_get_particles(self._h, dptr(pos), dptr(mom)))
        return pos, mom

    # ---- initial-state samplers on the device ----
    def sample_fonll(self, tables, N, antithetic=False, seed=0, stream=0):
        """sample_particles_from_FONLL (utils.jl:33-172) on the device from the tables of
        lmhip.utils.fonll_tables: (x (2, N), p (2, N))"""
        keep = {k: (None if tables[k] is None else np.asfortranarray(tables[k], dtype=np.float64))
                for k in ("r_grid", "p_grid", "cdf_r", "n_r", "cdf_p", "p_ok")}
        nr, np_ = keep["r_grid"].size, keep["p_grid"].size
        if keep["cdf_p"].shape != (np_, nr) or keep["p_ok"].size != nr:
            raise ValueError("fonll tables: cdf_p must be (np, nr) and p_ok (nr,)")
        null = ctypes.POINTER(ctypes.c_double)()
        t = _ffi.lmh_fonll_tables(dptr(keep["r_grid"]), dptr(keep["p_grid"]),
                                  dptr(keep["cdf_r"]) if keep["cdf_r"] is not None else null,
                                  dptr(keep["n_r"]) if keep["n_r"] is not None else null,
                                  dptr(keep["cdf_p"]), dptr(keep["p_ok"]), nr, np_,
                                  1 if tables["cartesian"] else 0, 0, float(tables["rmax"]), float(tables["fmax"]))
        x = np.empty((2, int(N)), dtype=np.float64, order="F")
        p = np.empty((2, int(N)), dtype=np.float64, order="F")
        self._check(self._lib.lmh_sample_fonll(self._h, ctypes.byref(t), int(N), 1 if antithetic else 0,
                                               int(seed), int(stream), dptr(x), dptr(p)))
        return x, p

    def append_thermal_pz(self, momenta, positions, m, t0, antithetic=False, seed=0, stream=0):
        """append_thermal_pz (utils.jl:533-558) on the device, T(r) from the radial medium at t0:
        (2, N) transverse momenta -> (3, N)"""
        mom = np.asfortranarray(momenta, dtype=np.float64)
        pos = np.asfortranarray(positions, dtype=np.float64)
        if mom.ndim != 2 or mom.shape[0] != 2 or pos.ndim != 2 or pos.shape[1] != mom.shape[1]:
            raise ValueError(f"append_thermal_pz: expected 2 transverse momentum rows, got {mom.shape[0]}")
        out = np.empty((3, mom.shape[1]), dtype=np.float64, order="F")
        self._check(self._lib.lmh_append_thermal_pz(self._h, dptr(mom), dptr(pos), pos.shape[0], mom.shape[1],
                                                    float(m), float(t0), 1 if antithetic else 0, int(seed),
                                                    int(stream), dptr(out)))
        return out

    # ---- stepping ----
    def noise_channels(self, params) -> int:
        return int(self._lib.lmh_noise_channels(ctypes.byref(params), self.dims, self.pdim))

    def boost_to_lab(self, params, step: int = 0):
        self._check(self._lib.lmh_boost_to_lab(self._h, ctypes.byref(params), int(step)))

    def run(self, params, step_begin: int, nsteps: int, noise=None, steps_per_launch: int = 0):
        nz = None
        if params.noise == _ffi.LMH_NOISE_FIXED:
            if noise is None:
                raise ValueError("fixed-noise mode needs a noise record (nsteps, C, N)")
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            C = self.noise_channels(params)
            if nz.shape != (nsteps, C, self.N):
                raise ValueError(f"noise record must be (nsteps, C, N) = {(nsteps, C, self.N)}")
        self._check(self._lib.lmh_run(self._h, ctypes.byref(params), int(step_begin), int(nsteps),
                                      dptr(nz), int(steps_per_launch)))

    def run_device_noise(self, params, step_begin: int, nsteps: int, d_noise_ptr: int,
                         steps_per_launch: int = 0):
        self._check(self._lib.lmh_run_device_noise(self._h, ctypes.byref(params), int(step_begin),
                                                   int(nsteps), ctypes.c_void_p(d_noise_ptr),
                                                   int(steps_per_launch)))

    def nonfinite_step(self) -> int:
        return int(self._lib.lmh_nonfinite_step(self._h))

    # ---- launch timing ----
    def profile_start(self):
        self._check(self._lib.lmh_profile_start(self._h))

    def profile_stop(self) -> dict:
        p = _ffi.lmh_profile()
        self._check(self._lib.lmh_profile_stop(self._h, ctypes.byref(p)))
        names = ("step", "redraw", "slice", "sort")
        return {n: {"launches": int(p.launches[k]), "ms": float(p.ms[k])} for k, n in enumerate(names)}

    # ---- observables ----
    def observables(self, nbins=40, pt_min=0.0, pt_max=10.0, sigma=0.0, species=-1):
        spec = _ffi.lmh_obs_spec(int(nbins), int(species), float(pt_min), float(pt_max), float(sigma))
        out = np.zeros(5 + int(nbins), dtype=np.float64)
        self._check(self._lib.lmh_observables(self._h, ctypes.byref(spec), dptr(out)))
        return out

    # ---- snapshots ----
    def snapshot(self) -> int:
        """take the state after every step enqueued so far; the host copy overlaps later steps"""
        k = ctypes.c_int32(-1)
        self._check(self._lib.lmh_snapshot(self._h, ctypes.byref(k)))
        return k.value

    def snapshot_get(self, index: int):
        pos = np.empty((self.dims, self.N), dtype=np.float64, order="F")
        mom = np.empty((self.pdim, self.N), dtype=np.float64, order="F")
        self._check(self._lib.lmh_snapshot_get(self._h, int(index), dptr(pos), dptr(mom)))
        return pos, mom

    # ---- on-the-fly freeze-out capture ----
    def freezeout_enable(self, Tfo: float, interp: bool = True):
        self._check(self._lib.lmh_freezeout_enable(self._h, float(Tfo), int(bool(interp))))

    def freezeout_get(self):
        """(pos (dims, N), mom (pdim, N), tau (N), flag (N)) of the latched states"""
        pos = np.empty((self.dims, self.N), dtype=np.float64, order="F")
        mom = np.empty((self.pdim, self.N), dtype=np.float64, order="F")
        tau = np.empty(self.N, dtype=np.float64)
        flag = np.empty(self.N, dtype=np.float64)
        self._check(self._lib.lmh_freezeout_get(self._h, dptr(pos), dptr(mom), dptr(tau), dptr(flag)))
        return pos, mom, tau, flag

    # ---- pathwise adjoint ----
    def adjoint_forward(self, params, step_begin: int, nsteps: int, checkpoint_every: int = 50,
                        dtau_dDsT: float | None = None, keep_tape: bool = False):
        """Steps as run() does, keeping a checkpoint every `checkpoint_every` steps (with
        keep_tape, the whole tape instead, so that the pullback only runs reverse sweeps).
        dtau_dDsT is (d tau_drag/d DsT)/tau_drag of the drag table: 1/DsT (the default) for
        tables proportional to DsT, 0 for the linear effective-D_s law."""
        if dtau_dDsT is None:
            dtau_dDsT = 1.0 / params.DsT if params.DsT > 0.0 else 0.0
        spec = _ffi.lmh_adjoint_spec(int(checkpoint_every), int(bool(keep_tape)), float(dtau_dDsT))
        self._check(self._lib.lmh_adjoint_forward(self._h, ctypes.byref(params), int(step_begin),
                                                  int(nsteps), ctypes.byref(spec)))

    def adjoint_backward(self, cotangent, nbins=40, pt_min=0.0, pt_max=10.0, sigma=0.0, species=-1,
                         state_grad=False):
        """dL/d(DsT, gamma0, Lambda) for L = cotangent . observables(nbins, ...) of the final
        state; with state_grad also dL/d(positions, momenta) before the forward pass."""
        spec = _ffi.lmh_obs_spec(int(nbins), int(species), float(pt_min), float(pt_max), float(sigma))
        w = np.ascontiguousarray(cotangent, dtype=np.float64)
        if w.shape != (5 + int(nbins),):
            raise ValueError(f"cotangent must have 5 + nbins = {5 + int(nbins)} entries")
        grad = np.zeros(3, dtype=np.float64)
        dpos = dmom = None
        if state_grad:
            dpos = np.empty((self.dims, self.N), dtype=np.float64, order="F")
            dmom = np.empty((self.pdim, self.N), dtype=np.float64, order="F")
        self._check(self._lib.lmh_adjoint_backward(self._h, ctypes.byref(spec), dptr(w), dptr(grad),
                                                   dptr(dpos), dptr(dmom)))
        return (grad, dpos, dmom) if state_grad else grad
