"""MI355X backend for LangevInMedium.jl's per-particle Langevin step.

The step runs in liblmhip.so (hand-written FP64 HIP kernels for gfx950, C ABI in
include/lmhip.h).  This package is the host side: the backend types, `simulate_ensemble_bulk`
with the reference's arguments and return shapes, the host-built transport tables and a
`Context` for direct use of the library.
"""
from .backends import AbstractBackend, HIP_GCBackend, HIPBackend
from .context import Context, LMHError
from .media import Grid2p1Medium
from .simulate import simulate_ensemble_bulk
from .transport import (GevInvTofm, build_juttner_invcdf, build_tau_drag_spline,
                        build_taun_current_spline, effective_DsT, eval_tau_n_spline, fmGeV,
                        tau_drag, tau_n_main3)
from .utils import check_momentum_dims

__all__ = [
    "AbstractBackend", "HIPBackend", "HIP_GCBackend", "Context", "LMHError", "Grid2p1Medium",
    "simulate_ensemble_bulk", "GevInvTofm", "fmGeV", "tau_drag", "tau_n_main3", "effective_DsT",
    "build_tau_drag_spline", "build_taun_current_spline", "eval_tau_n_spline",
    "build_juttner_invcdf", "check_momentum_dims",
]
