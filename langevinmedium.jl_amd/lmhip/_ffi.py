"""ctypes binding of liblmhip.so (include/lmhip.h).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

from . import build as _build

# ---- constants of include/lmhip.h --------------------------------------------------------------
LMH_OK = 0
LMH_E_INVALID = -1
LMH_E_STATE = -2
LMH_E_HIP = -3
LMH_E_NONFINITE = -4
LMH_E_NOMEM = -5

LMH_CARTESIAN = 0
LMH_GENERAL_COORDS = 1
LMH_MILNE = 2
LMH_LANGEVIN = 0
LMH_RTA = 1
LMH_FLUID = 2
LMH_LANGEVIN_RTA = 3
LMH_NOISE_FIXED = 0
LMH_NOISE_PHILOX = 1
LMH_PREPOINT = 0
LMH_MIDPOINT = 1
LMH_POSTPOINT = 2
LMH_DRAG_TABLE = 0
LMH_DRAG_GAMMA_P = 1
LMH_TABLE_DRAG = 0
LMH_TABLE_RTA = 1
LMH_MAX_SPECIES = 4


class lmh_params(ctypes.Structure):
    _fields_ = [
        ("dt", c_double), ("t0", c_double), ("DsT", c_double), ("psi2", c_double),
        ("gamma0", c_double), ("Lambda", c_double),
        ("seed", c_uint64), ("stream", c_uint64),
        ("coords", c_int32), ("collision", c_int32), ("noise", c_int32),
        ("integrator", c_int32), ("drag_model", c_int32), ("relativistic", c_int32),
        ("bjorken_redshift", c_int32), ("position_diffusion", c_int32),
        ("reflecting_boundary", c_int32), ("check_finite", c_int32),
    ]


class lmh_tuning(ctypes.Structure):
    _fields_ = [("block", c_int32), ("steps_per_launch", c_int32), ("slice", c_int32), ("defer", c_int32),
                ("specialize", c_int32), ("sort_every", c_int32)]


class lmh_profile(ctypes.Structure):
    _fields_ = [("launches", c_int64 * 4), ("ms", c_double * 4)]


class lmh_fonll_tables(ctypes.Structure):
    _fields_ = [("r_grid", POINTER(c_double)), ("p_grid", POINTER(c_double)), ("cdf_r", POINTER(c_double)),
                ("n_r", POINTER(c_double)), ("cdf_p", POINTER(c_double)), ("p_ok", POINTER(c_double)),
                ("nr", c_int32), ("np", c_int32), ("cartesian", c_int32),
                ("pad_", c_int32), ("rmax", c_double), ("fmax", c_double)]


class lmh_obs_spec(ctypes.Structure):
    _fields_ = [("nbins", c_int32), ("species", c_int32), ("pt_min", c_double),
                ("pt_max", c_double), ("sigma", c_double)]


class lmh_adjoint_spec(ctypes.Structure):
    _fields_ = [("checkpoint_every", c_int32), ("keep_tape", c_int32), ("dtau_dDsT", c_double)]


class lmh_run_desc(ctypes.Structure):
    _fields_ = [("count", c_int64), ("mass", c_double), ("DsT", c_double), ("medium", c_int32),
                ("drag_table", c_int32), ("rta_table", c_int32), ("juttner_table", c_int32)]


class lmh_batch_tuning(ctypes.Structure):
    _fields_ = [("block", c_int32), ("steps_per_launch", c_int32), ("persistent", c_int32),
                ("graph_steps", c_int32), ("sort_every", c_int32), ("slice_min", c_int32)]


def make_params(*, dt, t0=0.0, DsT=0.0, psi2=0.0, gamma0=0.0, Lambda=1.0, seed=0, stream=0,
                coords=LMH_CARTESIAN, collision=LMH_LANGEVIN, noise=LMH_NOISE_PHILOX,
                integrator=LMH_PREPOINT, drag_model=LMH_DRAG_TABLE, relativistic=True,
                bjorken_redshift=False, position_diffusion=False, reflecting_boundary=False,
                check_finite=True) -> lmh_params:
    return lmh_params(float(dt), float(t0), float(DsT), float(psi2), float(gamma0), float(Lambda),
                      int(seed) & 0xFFFFFFFFFFFFFFFF, int(stream) & 0xFFFFFFFFFFFFFFFF,
                      int(coords), int(collision), int(noise), int(integrator), int(drag_model),
                      int(bool(relativistic)), int(bool(bjorken_redshift)),
                      int(bool(position_diffusion)), int(bool(reflecting_boundary)),
                      int(bool(check_finite)))


_PD = POINTER(c_double)
_PI64 = POINTER(c_int64)

# name -> (restype, argtypes); every function declared in include/lmhip.h
SIGNATURES = {
    "lmh_noise_channels": (c_int32, [POINTER(lmh_params), c_int32, c_int32]),
    "lmh_create": (c_int, [POINTER(c_void_p), c_int32]),
    "lmh_destroy": (None, [c_void_p]),
    "lmh_last_error": (c_char_p, [c_void_p]),
    "lmh_stream": (c_void_p, [c_void_p]),
    "lmh_synchronize": (c_int, [c_void_p]),
    "lmh_set_tuning": (c_int, [c_void_p, POINTER(lmh_tuning)]),
    "lmh_get_tuning": (c_int, [c_void_p, POINTER(lmh_tuning)]),
    "lmh_set_medium_radial": (c_int, [c_void_p, _PD, c_int32, _PD, c_int32, _PD, _PD, _PD]),
    "lmh_set_medium_grid2p1": (c_int, [c_void_p, c_double, c_double, c_int32, c_double, c_double,
                                       c_int32, _PD, c_int32, _PD, _PD, _PD]),
    "lmh_set_time_table": (c_int, [c_void_p, c_int32, c_int32, c_double, c_double, _PD, c_int32]),
    "lmh_set_juttner_table": (c_int, [c_void_p, c_int32, _PD, c_int32, c_int32, c_double, c_double]),
    "lmh_set_particles": (c_int, [c_void_p, _PD, c_int32, _PD, c_int32, c_int64, c_int32, _PI64, _PD]),
    "lmh_get_particles": (c_int, [c_void_p, _PD, _PD]),
    "lmh_sample_fonll": (c_int, [c_void_p, POINTER(lmh_fonll_tables), c_int64, c_int32, c_uint64, c_uint64,
                                 _PD, _PD]),
    "lmh_append_thermal_pz": (c_int, [c_void_p, _PD, _PD, c_int32, c_int64, c_double, c_double, c_int32,
                                      c_uint64, c_uint64, _PD]),
    "lmh_boost_to_lab": (c_int, [c_void_p, POINTER(lmh_params), c_int64]),
    "lmh_run": (c_int, [c_void_p, POINTER(lmh_params), c_int64, c_int64, _PD, c_int32]),
    "lmh_run_device_noise": (c_int, [c_void_p, POINTER(lmh_params), c_int64, c_int64, c_void_p, c_int32]),
    "lmh_nonfinite_step": (c_int64, [c_void_p]),
    "lmh_observables": (c_int, [c_void_p, POINTER(lmh_obs_spec), _PD]),
    "lmh_observables_noalloc": (c_int, [c_void_p, POINTER(lmh_obs_spec), c_void_p]),
    "lmh_profile_start": (c_int, [c_void_p]),
    "lmh_profile_stop": (c_int, [c_void_p, POINTER(lmh_profile)]),
    "lmh_snapshot": (c_int, [c_void_p, POINTER(c_int32)]),
    "lmh_snapshot_get": (c_int, [c_void_p, c_int32, _PD, _PD]),
    "lmh_freezeout_enable": (c_int, [c_void_p, c_double, c_int32]),
    "lmh_freezeout_get": (c_int, [c_void_p, _PD, _PD, _PD, _PD]),
    "lmh_adjoint_forward": (c_int, [c_void_p, POINTER(lmh_params), c_int64, c_int64,
                                    POINTER(lmh_adjoint_spec)]),
    "lmh_adjoint_backward": (c_int, [c_void_p, POINTER(lmh_obs_spec), _PD, _PD, _PD, _PD]),
    "lmh_batch_create": (c_int, [POINTER(c_void_p), c_int32]),
    "lmh_batch_destroy": (None, [c_void_p]),
    "lmh_batch_last_error": (c_char_p, [c_void_p]),
    "lmh_batch_stream": (c_void_p, [c_void_p]),
    "lmh_batch_add_medium_grid2p1": (c_int, [c_void_p, c_double, c_double, c_int32, c_double, c_double,
                                             c_int32, _PD, c_int32, _PD, _PD, _PD, POINTER(c_int32)]),
    "lmh_batch_add_time_table": (c_int, [c_void_p, c_double, c_double, _PD, c_int32, POINTER(c_int32)]),
    "lmh_batch_add_juttner_table": (c_int, [c_void_p, _PD, c_int32, c_int32, c_double, c_double,
                                            POINTER(c_int32)]),
    "lmh_batch_set_runs": (c_int, [c_void_p, POINTER(lmh_run_desc), c_int32, _PD, _PD, c_int32]),
    "lmh_batch_get_particles": (c_int, [c_void_p, _PD, _PD]),
    "lmh_batch_set_tuning": (c_int, [c_void_p, POINTER(lmh_batch_tuning)]),
    "lmh_batch_get_tuning": (c_int, [c_void_p, POINTER(lmh_batch_tuning)]),
    "lmh_batch_boost_to_lab": (c_int, [c_void_p, POINTER(lmh_params), c_int64]),
    "lmh_batch_run": (c_int, [c_void_p, POINTER(lmh_params), c_int64, c_int64]),
    "lmh_batch_autotune": (c_int, [c_void_p, POINTER(lmh_params), c_int32, POINTER(lmh_batch_tuning),
                                   POINTER(c_double)]),
    "lmh_batch_observables": (c_int, [c_void_p, POINTER(lmh_obs_spec), _PD]),
    "lmh_batch_nonfinite_step": (c_int64, [c_void_p]),
}

_LIB = None


def header_symbols() -> list[str]:
    """Function names declared in include/lmhip.h."""
    path = os.path.join(_build.REPO_ROOT, "include", "lmhip.h")
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmh_[a-z0-9_]+)\s*\(", text)))


def lib_path() -> str:
    return _build.LIB_PATH


def lib() -> ctypes.CDLL:
    """The loaded library.  Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def dptr(a) -> POINTER(c_double):
    return a.ctypes.data_as(_PD) if a is not None else None
