"""ctypes access to the CPU oracle (oracle/liblmoracle.so) for the tests.

The oracle is test infrastructure: it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

from lmhip import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liblmoracle.so")

PD = POINTER(c_double)


class RadialMedium(Structure):
    _fields_ = [("xgrid", PD), ("tgrid", PD), ("T", PD), ("v", PD), ("v2", PD),
                ("nx", c_int32), ("nt", c_int32)]


class GridMedium(Structure):
    _fields_ = [("x0", c_double), ("dx", c_double), ("y0", c_double), ("dy", c_double),
                ("tgrid", PD), ("T", PD), ("vx", PD), ("vy", PD),
                ("nx", c_int32), ("ny", c_int32), ("nt", c_int32), ("_pad", c_int32)]


class Table(Structure):
    _fields_ = [("vals", PD), ("Tmin", c_double), ("invdT", c_double), ("n", c_int32), ("_pad", c_int32)]


class JTable(Structure):
    _fields_ = [("vals", PD), ("Tmin", c_double), ("invdT", c_double), ("nU", c_int32), ("nT", c_int32)]


class Species(Structure):
    _fields_ = [("nspecies", c_int32), ("_pad", c_int32), ("counts", c_int64 * 4),
                ("mass", c_double * 4), ("drag", Table * 4), ("rta", Table * 4), ("jt", JTable * 4)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: build the oracle first")
        L = ctypes.CDLL(LIB)
        P = POINTER(_ffi.lmh_params)
        sig = {
            "lmo_fmGeV": (c_double, []),
            "lmo_searchsortedlast": (c_int64, [PD, c_int64, c_double]),
            "lmo_interpolate_2d": (c_double, [PD, c_int32, PD, c_int32, PD, c_double, c_double]),
            "lmo_interpolate_2d_gc": (c_double, [PD, c_int32, PD, c_int32, PD, c_double, c_double]),
            "lmo_eval_time_table": (c_double, [c_double, c_double, c_double, PD, c_int32]),
            "lmo_tau_drag": (c_double, [c_double, c_double, c_double]),
            "lmo_effective_DsT": (c_double, [c_double, c_double, c_int32, c_double, c_double, c_double,
                                             c_int32, c_double, POINTER(c_int32)]),
            "lmo_besselkx": (c_double, [c_double, c_double]),
            "lmo_tau_n_main3": (c_double, [c_double, c_double, c_double]),
            "lmo_build_time_spline": (c_int, [c_int32, c_double, c_double, c_double, c_double, c_int32,
                                              c_int32, c_double, c_double, c_double, c_int32, c_double,
                                              c_double, PD, PD, PD]),
            "lmo_build_juttner_invcdf": (c_int, [c_double, c_int32, c_double, c_double, c_int32, c_int32,
                                                 c_int32, c_double, PD, PD, PD]),
            "lmo_juttner_invcdf_lookup": (c_double, [PD, c_int32, c_int32, c_double, c_double, c_double, c_double]),
            "lmo_draw_juttner_rejection": (c_int, [c_double, c_double, c_int32, c_uint64, c_int64, PD]),
            "lmo_philox4x32_10": (None, [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]),
            "lmo_noise_block": (None, [c_uint64, c_uint64, c_int64, c_int64, c_int32, POINTER(c_uint64)]),
            "lmo_set_threads": (None, [c_int32]),
            "lmo_freezeout_capture": (None, [c_void_p, c_void_p, PD, PD, c_int32, c_int32, c_int64, c_int64,
                                             c_double, c_double, c_double, c_int32, PD, PD, PD, PD,
                                             PD, PD, PD, PD]),
            "lmo_get_threads": (c_int32, []),
            "lmo_noise_channels": (c_int32, [P, c_int32, c_int32]),
            "lmo_noise_record": (None, [P, c_int32, c_int32, c_int64, c_int64, PD]),
            "lmo_kernel_boost_to_rest": (None, [PD, PD, POINTER(RadialMedium), c_double, c_int64, c_int64,
                                                c_double, c_double, c_int32, c_int32, c_int32, c_double, c_int32]),
            "lmo_kernel_boost_to_lab": (None, [PD, PD, POINTER(RadialMedium), c_double, c_int64, c_int64,
                                               c_double, c_double, c_int32, c_int32, c_int32, c_double, c_int32]),
            "lmo_boost_to_lab_all": (c_int, [P, POINTER(RadialMedium), POINTER(Species), c_int32, c_int32,
                                             PD, PD, c_int64, c_int64]),
            "lmo_boost_to_lab_grid": (c_int, [P, POINTER(GridMedium), POINTER(Species), c_int32, PD, PD,
                                              c_int64, c_int64]),
            "lmo_run_cartesian": (c_int, [P, POINTER(RadialMedium), POINTER(Species), c_int32, c_int32, PD, PD,
                                          c_int64, c_int64, c_int64, PD, POINTER(c_int64)]),
            "lmo_milne_deta": (c_double, [c_double, c_double, c_double, c_double, c_double, c_double]),
            "lmo_run_gc": (c_int, [P, POINTER(RadialMedium), POINTER(Species), PD, PD, c_int64, c_int64,
                                   c_int64, PD, POINTER(c_int64)]),
            "lmo_grid_sample": (None, [POINTER(GridMedium), c_double, c_double, c_double, PD]),
            "lmo_run_grid": (c_int, [P, POINTER(GridMedium), POINTER(Species), c_int32, PD, PD, c_int64,
                                     c_int64, c_int64, PD, POINTER(c_int64)]),
            "lmo_observables": (c_int, [PD, c_int32, c_int64, c_int64, POINTER(_ffi.lmh_obs_spec), PD]),
            "lmo_sample_fonll": (c_int, [PD, PD, PD, PD, PD, PD, c_int32, c_int32, c_int32, c_double, c_double,
                                         c_int64, c_int32, c_uint64, c_uint64, PD, PD]),
            "lmo_append_thermal_pz": (None, [PD, PD, c_int32, c_int64, c_double, c_double, c_int32, c_uint64,
                                             c_uint64, POINTER(RadialMedium), PD]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def dp(a):
    return None if a is None else a.ctypes.data_as(PD)


class Keep:
    """holds numpy buffers referenced by ctypes structures"""

    def __init__(self):
        self.refs = []

    def f64(self, a, order="F"):
        if a is None:
            return None
        b = np.require(np.asarray(a, dtype=np.float64), requirements=[order, "A"])
        self.refs.append(b)
        return b


def radial_medium(keep, xgrid, tgrid, T, v, v2=None):
    xg, tg = keep.f64(np.ravel(xgrid)), keep.f64(np.ravel(tgrid))
    Ta, va, v2a = keep.f64(T), keep.f64(v), keep.f64(v2)
    return RadialMedium(dp(xg), dp(tg), dp(Ta), dp(va), dp(v2a), xg.size, tg.size)


def grid_medium(keep, med):
    tg = keep.f64(np.ravel(med.tgrid))
    T, vx, vy = keep.f64(med.T), keep.f64(med.vx), keep.f64(med.vy)
    return GridMedium(med.x0, med.dx, med.y0, med.dy, dp(tg), dp(T), dp(vx), dp(vy),
                      med.nx, med.ny, tg.size, 0)


def species(keep, masses, counts, drag=None, rta=None, jt=None):
    """drag/rta: lists of (Tmin, invdT, vals); jt: list of (vals, nU, nT, Tmin, invdT)"""
    s = Species()
    s.nspecies = len(masses)
    for k, (mk, ck) in enumerate(zip(masses, counts)):
        s.counts[k] = int(ck)
        s.mass[k] = float(mk)
        for dst, src in ((s.drag, drag), (s.rta, rta)):
            if src is not None and src[k] is not None:
                t0, inv, vals = src[k]
                v = keep.f64(np.ravel(vals))
                dst[k] = Table(dp(v), t0, inv, v.size, 0)
            else:
                v = keep.f64(np.zeros(2))
                dst[k] = Table(dp(v), 0.0, 1.0, 2, 0)
        if jt is not None and jt[k] is not None:
            vals, nU, nT, Tt, inv = jt[k]
            v = keep.f64(np.ravel(vals))
            s.jt[k] = JTable(dp(v), Tt, inv, nU, nT)
        else:
            v = keep.f64(np.zeros(4))
            s.jt[k] = JTable(dp(v), 0.0, 1.0, 2, 2)
    return s


def sample_fonll(tables, N, antithetic=False, seed=0, stream=0):
    """The FONLL sampler's draws (the device sampler's definition) from lmhip.utils.fonll_tables."""
    keep = Keep()
    t = {k: keep.f64(tables[k]) for k in ("r_grid", "p_grid", "cdf_r", "n_r", "cdf_p", "p_ok")}
    x = np.zeros((2, N), order="F")
    p = np.zeros((2, N), order="F")
    rc = lib().lmo_sample_fonll(dp(t["r_grid"]), dp(t["p_grid"]), dp(t["cdf_r"]), dp(t["n_r"]), dp(t["cdf_p"]),
                                dp(t["p_ok"]), t["r_grid"].size, t["p_grid"].size, 1 if tables["cartesian"] else 0,
                                float(tables["rmax"]), float(tables["fmax"]), N, 1 if antithetic else 0,
                                seed, stream, dp(x), dp(p))
    assert rc == 0, rc
    return x, p


def append_thermal_pz(mom, pos, m, t0, med, antithetic=False, seed=0, stream=0):
    """append_thermal_pz with the device sampler's draws, T(r) of the radial medium at t0."""
    keep = Keep()
    mom, pos = keep.f64(mom), keep.f64(pos)
    out = np.zeros((3, mom.shape[1]), order="F")
    lib().lmo_append_thermal_pz(dp(mom), dp(pos), pos.shape[0], mom.shape[1], float(m), float(t0),
                                1 if antithetic else 0, seed, stream, ctypes.byref(med), dp(out))
    return out


def noise_record(params, dims, pdim, N, step_begin, nsteps):
    """The Philox noise record the kernels generate, (nsteps, C, N)."""
    L = lib()
    C = L.lmo_noise_channels(ctypes.byref(params), dims, pdim)
    out = np.zeros((nsteps, C, N), dtype=np.float64)
    for s in range(nsteps):
        L.lmo_noise_record(ctypes.byref(params), dims, pdim, N, step_begin + s, dp(out[s]))
    return out


def run_cartesian(params, med, sp, pos, mom, step_begin, nsteps, noise):
    """advance (pos, mom) in place; returns (rc, bad_step)"""
    L = lib()
    bad = c_int64(-1)
    rc = L.lmo_run_cartesian(ctypes.byref(params), ctypes.byref(med), ctypes.byref(sp), pos.shape[0],
                             mom.shape[0], dp(pos), dp(mom), pos.shape[1], step_begin, nsteps,
                             dp(noise), ctypes.byref(bad))
    return rc, bad.value


def run_gc(params, med, sp, pos, mom, step_begin, nsteps, noise):
    L = lib()
    bad = c_int64(-1)
    rc = L.lmo_run_gc(ctypes.byref(params), ctypes.byref(med), ctypes.byref(sp), dp(pos), dp(mom),
                      pos.shape[1], step_begin, nsteps, dp(noise), ctypes.byref(bad))
    return rc, bad.value


def run_grid(params, med, sp, pos, mom, step_begin, nsteps, noise):
    L = lib()
    bad = c_int64(-1)
    rc = L.lmo_run_grid(ctypes.byref(params), ctypes.byref(med), ctypes.byref(sp), mom.shape[0],
                        dp(pos), dp(mom), pos.shape[1], step_begin, nsteps, dp(noise), ctypes.byref(bad))
    return rc, bad.value


def boost_to_lab_grid(params, med, sp, pos, mom, step):
    return lib().lmo_boost_to_lab_grid(ctypes.byref(params), ctypes.byref(med), ctypes.byref(sp),
                                       mom.shape[0], dp(pos), dp(mom), pos.shape[1], step)


def boost_to_lab_all(params, med, sp, pos, mom, step):
    return lib().lmo_boost_to_lab_all(ctypes.byref(params), ctypes.byref(med), ctypes.byref(sp),
                                      pos.shape[0], mom.shape[0], dp(pos), dp(mom), pos.shape[1], step)


def observables(mom, i0, i1, nbins, pt_min, pt_max, sigma):
    spec = _ffi.lmh_obs_spec(nbins, -1, pt_min, pt_max, sigma)
    out = np.zeros(5 + nbins)
    m = np.require(mom, dtype=np.float64, requirements=["F"])
    rc = lib().lmo_observables(dp(m), m.shape[0], i0, i1, ctypes.byref(spec), dp(out))
    assert rc == 0
    return out


class Freezeout:
    """the capture arrays of lmo_freezeout_capture, all zero (the GPU driver's CUDA.zeros)"""

    def __init__(self, dims, pdim, N):
        self.dims, self.pdim, self.N = dims, pdim, N
        z = lambda r: np.zeros((r, N), order="F")
        self.fo_pos, self.fo_mom = z(dims), z(pdim)
        self.fo_tau, self.fo_flag = np.zeros(N), np.zeros(N)
        self.last_pos, self.last_mom = z(dims), z(pdim)
        self.last_tau, self.last_T = np.zeros(N), np.zeros(N)

    def capture(self, rmed, gmed, pos, mom, step, dt, t0, Tfo, interp):
        lib().lmo_freezeout_capture(ctypes.byref(rmed) if rmed is not None else None,
                                    ctypes.byref(gmed) if gmed is not None else None,
                                    dp(pos), dp(mom), self.dims, self.pdim, self.N, step, dt, t0, Tfo,
                                    int(bool(interp)), dp(self.fo_pos), dp(self.fo_mom), dp(self.fo_tau),
                                    dp(self.fo_flag), dp(self.last_pos), dp(self.last_mom),
                                    dp(self.last_tau), dp(self.last_T))
