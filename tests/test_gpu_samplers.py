"""The initial-state samplers on the device against the oracle, draw by draw: the FONLL sampler
(utils.jl:33-172, with the antithetic pairing of simulate_cpu.jl:66-89) and the thermal p_z
completion (utils.jl:499-558).  Both consume Philox words of the same counters, so every quark's
position and momentum agree with the oracle's to 1e-12 (relative, floored at 1e-6); the draws'
accept/reject decisions agree exactly (no quark takes a different attempt)."""
import math

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import flowing_medium, rel_dev
from lmhip import _ffi
from lmhip import utils as U

pytestmark = pytest.mark.gpu
TOL = 1e-12
RG = np.linspace(0.0, 10.0, 41)
PG = np.linspace(0.0, 20.0, 161)


def _density():
    f = np.exp(-PG[:, None] / 2.0) * np.exp(-RG[None, :] ** 2 / 8.0)
    f[:, 7] = 0.0                                  # one r node with no quarks: |p| = 0 there
    return f


def _close(a, b, floor=1e-6):
    d = rel_dev(a, b, floor).max()
    print(f"max rel dev {d:.2e}", flush=True)
    return d <= TOL


@pytest.mark.parametrize("cartesian", [False, True])
@pytest.mark.parametrize("antithetic,N", [(False, 50_000), (True, 50_001), (True, 1)])
def test_sample_fonll_matches_oracle(gpu_ok, cartesian, antithetic, N):
    t = U.fonll_tables(RG, PG, _density(), cartesian_spatial_sampling=cartesian)
    with lmhip.Context(0) as ctx:
        gx, gp = ctx.sample_fonll(t, N, antithetic=antithetic, seed=19, stream=4)
    ox, op = O.sample_fonll(t, N, antithetic=antithetic, seed=19, stream=4)
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gp))
    assert _close(gx, ox) and _close(gp, op)
    if antithetic and N > 1:
        assert np.array_equal(gx[:, 0:N - 1:2], gx[:, 1:N:2])
        assert np.array_equal(gp[:, 0:N - 1:2], -gp[:, 1:N:2])


def test_sample_fonll_empty_and_errors(gpu_ok):
    t = U.fonll_tables(RG, PG, _density(), cartesian_spatial_sampling=True)
    with lmhip.Context(0) as ctx:
        x, p = ctx.sample_fonll(t, 0)
        assert x.shape == (2, 0) and p.shape == (2, 0)
        bad = dict(t, cdf_p=t["cdf_p"][:-1])
        with pytest.raises(ValueError):
            ctx.sample_fonll(bad, 10)
        with pytest.raises(lmhip.LMHError, match="vanishes everywhere"):
            ctx.sample_fonll(dict(t, n_r=np.zeros_like(t["n_r"])), 10)


@pytest.mark.parametrize("dims,antithetic", [(2, False), (2, True), (1, False)])
def test_append_thermal_pz_matches_oracle(gpu_ok, dims, antithetic):
    """T(r, t0) from a flowing radial medium (both rejection envelopes occur: T spans 0.05-0.45)"""
    rng = np.random.default_rng(8)
    N = 40_001
    xg, tg, Tf, Vf, _ = flowing_medium()
    pos = np.asfortranarray(rng.uniform(0.0, 9.0, (dims, N)))
    mom = np.asfortranarray(rng.normal(0.0, 1.5, (2, N)))
    keep = O.Keep()
    om = O.radial_medium(keep, xg, tg, Tf, Vf)
    with lmhip.Context(0) as ctx:
        ctx.set_medium_radial(xg, tg, Tf, Vf)
        got = ctx.append_thermal_pz(mom, pos, 1.5, 0.6, antithetic=antithetic, seed=23, stream=2)
    want = O.append_thermal_pz(mom, pos, 1.5, 0.6, om, antithetic=antithetic, seed=23, stream=2)
    assert np.array_equal(got[:2], mom)
    assert _close(got, want)


def test_simulate_samples_on_the_device(gpu_ok):
    """simulate_ensemble_bulk with no x_init/p_init: FONLL draws (Cartesian, antithetic), the
    thermal p_z (mirrored across the pairs) and the initial boost, as the oracle composes them"""
    xg, tg, Tf, Vf, _ = flowing_medium()
    N, t0, dt, seed = 3001, 0.5, 0.01, 41
    f = _density()
    tp, mh, ph = lmhip.simulate_ensemble_bulk(
        lmhip.HIPBackend(), RG, PG, f, Tf, Vf, (xg, tg), N_particles=N, Δt=dt, initial_time=t0,
        final_time=t0 + 2 * dt, save_interval=dt, m=1.5, DsT=0.6366, dimensions=2, momentum_dimensions=3,
        antithetic_momenta=True, seed=seed)
    t = U.fonll_tables(RG, PG, f, cartesian_spatial_sampling=True)
    x, p2 = O.sample_fonll(t, N, antithetic=True, seed=seed, stream=0)
    keep = O.Keep()
    om = O.radial_medium(keep, xg, tg, Tf, Vf)
    p3 = O.append_thermal_pz(p2, x, 1.5, t0, om, antithetic=True, seed=seed, stream=0)
    d = lmhip.build_tau_drag_spline(1.5, 0.6366, Tmin=float(Tf.min()), Tmax=float(Tf.max()), nT=1024)
    sp = O.species(keep, [1.5], [N], drag=[d])
    prm = _ffi.make_params(dt=dt, t0=t0, DsT=0.6366, seed=seed)
    P, M = x.copy(order="F"), p3.copy(order="F")
    assert O.boost_to_lab_all(prm, om, sp, P, M, 0) == 0
    assert _close(ph[0], P) and _close(mh[0], M)
    assert np.all(np.isfinite(mh[-1])) and math.isclose(tp[-1], t0 + 2 * dt)
