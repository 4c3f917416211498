"""lmh_observables_noalloc: the observables with their partial sums in the library's static
device array, written to a caller-owned device buffer, against lmh_observables (which allocates
its own partial buffer), through the C ABI.  Two calls in a row, the second on a smaller state and
with fewer bins, so that the array's leftover contents from the first call cannot leak into the
second; the same kernels over the same slot order give the same bits."""
import ctypes

import numpy as np
import pytest

import lmhip
from lmhip import _ffi, media

pytestmark = pytest.mark.gpu


def noalloc(ctx, nbins, pt_min, pt_max, sigma, species):
    import torch
    out = torch.full((5 + nbins,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    spec = _ffi.lmh_obs_spec(nbins, species, pt_min, pt_max, sigma)
    L = _ffi.lib()
    rc = L.lmh_observables_noalloc(ctx._h, ctypes.byref(spec), ctypes.c_void_p(out.data_ptr()))
    assert rc == 0, L.lmh_last_error(ctx._h)
    ctx.synchronize()
    return out.cpu().numpy()


def test_observables_noalloc_matches_observables(gpu_ok):
    rng = np.random.default_rng(12)
    cases = [(300_001, [200_000, 100_001], 40, 0.25), (5_000, [3_000, 2_000], 12, 0.0)]
    for N, counts, nbins, sigma in cases:
        pos = np.zeros((2, N), order="F")
        mom = np.asfortranarray(rng.normal(0, 2.0, (3, N)))
        with lmhip.Context(0) as ctx:
            ctx.set_medium_radial(*media.uniform_bath(0.3))
            ctx.set_particles(pos, mom, [1.5, 4.75], counts)
            for sp in (-1, 1):
                for sg in (0.0, sigma):
                    want = ctx.observables(nbins, 0.0, 8.0, sg, sp)
                    got = noalloc(ctx, nbins, 0.0, 8.0, sg, sp)
                    print(f"N={N} nbins={nbins} sigma={sg} species={sp}: N_sel {got[0]:.0f}", flush=True)
                    assert np.array_equal(got, want), np.max(np.abs(got - want))
                    assert got[0] == (N if sp < 0 else counts[1])


def test_observables_noalloc_rejects_oversized_spec(gpu_ok):
    with lmhip.Context(0) as ctx:
        ctx.set_medium_radial(*media.uniform_bath(0.3))
        ctx.set_particles(np.zeros((2, 10), order="F"), np.ones((3, 10), order="F"), [1.5])
        spec = _ffi.lmh_obs_spec(121, -1, 0.0, 8.0, 0.0)
        rc = _ffi.lib().lmh_observables_noalloc(ctx._h, ctypes.byref(spec), ctypes.c_void_p(1))
        assert rc == _ffi.LMH_E_INVALID
