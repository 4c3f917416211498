"""The Bjorken-expanding workload in Milne coordinates through the C ABI, against the oracle.

The workload: T(tau) = T0 (tau0/tau)^(1/3) with T0 = 0.4 GeV, tau0 = 0.6 fm/c, uniform in r, the
fluid at rest in the comoving frame, stopped where T falls to 0.155 GeV (tau = 10.3118 fm/c, 1942
steps of dt = 0.005); charm quarks (m = 1.5) with the post-point kick and the drag
gamma(p, T) = gamma0 T^2/(1 + p^2/Lambda^2), Lambda = 5 GeV; p_T from dN/dp_T ~ p_T/(1 + p_T^2/4)^3
on [0, 20] GeV and rapidity uniform in [-1, 1] at eta_s = 0.  Bar: 1e-12 per particle (the
Philox record regenerated by the oracle from the same counters), and the reference's gate R
(p_z tau conserved, test_momentum_dims3.jl:173-188) at 1e-9.
"""
import math
import zlib

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import flowing_medium, load_tables, rel_dev, run_radial_pair, tables
from lmhip import _ffi, media

pytestmark = pytest.mark.gpu
TOL = 1e-12
GAMMA0 = (1 / 0.197327) / (1.5 * media.DsT_from_2piT(4.0))     # gamma(0, T) = the D_s(2 pi T) = 4 drag


def check(g, o, tol=TOL):
    (gp, gm), (op, om) = g, o
    assert np.all(np.isfinite(gm)) and np.all(np.isfinite(gp))
    dm = rel_dev(gm, om, 1e-6).max()
    dx = rel_dev(gp, op, 1e-6).max()
    print(f"max rel dev: mom {dm:.2e} pos {dx:.2e}", flush=True)
    assert dm <= tol and dx <= tol, (dm, dx)


MILNE_CASES = {
    "langevin": dict(),
    "langevin_rta": dict(collision=_ffi.LMH_LANGEVIN_RTA),
    "rta": dict(collision=_ffi.LMH_RTA),
    "fluid": dict(collision=_ffi.LMH_FLUID),
    "postpoint_gamma_p": dict(integrator=_ffi.LMH_POSTPOINT, drag_model=_ffi.LMH_DRAG_GAMMA_P),
    "midpoint": dict(integrator=_ffi.LMH_MIDPOINT),
    "nonrel": dict(relativistic=False),
}


@pytest.mark.parametrize("name", list(MILNE_CASES))
@pytest.mark.parametrize("noise", ["fixed", "philox"])
def test_milne_parity(gpu_ok, name, noise):
    """Milne coordinates on a flowing radial medium (the transverse boosts active), the driver's
    initial boost included: rows (x, y, eta_s), (p_x, p_y, p_z)"""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    N = 3000
    xg, tg, Tf, Vf, _ = flowing_medium()
    pos = np.asfortranarray(np.vstack([rng.normal(0, 4.0, (2, N)), 0.3 * rng.standard_normal(N)]))
    mom = np.asfortranarray(rng.normal(0, 1.5, (3, N)))
    nm = _ffi.LMH_NOISE_FIXED if noise == "fixed" else _ffi.LMH_NOISE_PHILOX
    p = _ffi.make_params(dt=0.01, t0=0.5, DsT=0.6366, noise=nm, seed=77, stream=3, gamma0=GAMMA0, Lambda=5.0,
                         coords=_ffi.LMH_MILNE, **MILNE_CASES[name])
    rta = p.collision in (_ffi.LMH_RTA, _ffi.LMH_LANGEVIN_RTA)
    g, o, _ = run_radial_pair(p, (xg, tg, Tf, Vf, None), [1.5], [N], pos, mom, 40, seed=5, rta=rta,
                              initial_boost=True)
    check(g, o)
    assert np.mean(g[0][2] != pos[2]) > 0.99                 # eta_s moved


def config2_state(N, rng):
    xg, tg, Tf, Vf, tau_end = media.bjorken_bath(T0=0.4, tau0=0.6, Tstop=0.155)
    mom = media.pt_rapidity_momenta(N, 1.5, rng, ymax=1.0, ptmax=20.0)
    pos = np.zeros((3, N), order="F")
    steps = int(math.floor((tau_end - 0.6) / 0.005))
    return (xg, tg, Tf, Vf), pos, mom, steps


def test_config2_bjorken_workload(gpu_ok):
    """the Bjorken workload in full length (1942 steps, T from 0.4 down to 0.155 GeV) at 2e4
    charm quarks, Philox noise, post-point kick with gamma(p, T): every quark against the oracle"""
    rng = np.random.default_rng(2)
    N = 20_000
    (xg, tg, Tf, Vf), pos, mom, steps = config2_state(N, rng)
    assert steps == 1942
    T_end = float(np.interp(0.6 + steps * 0.005, tg, Tf[0]))
    assert 0.155 <= T_end < 0.1552
    p = _ffi.make_params(dt=0.005, t0=0.6, DsT=media.DsT_from_2piT(4.0), noise=_ffi.LMH_NOISE_PHILOX,
                         seed=2026, stream=1, coords=_ffi.LMH_MILNE, integrator=_ffi.LMH_POSTPOINT,
                         drag_model=_ffi.LMH_DRAG_GAMMA_P, gamma0=GAMMA0, Lambda=5.0)
    drag, _, _ = tables([1.5], Tf, p.DsT, 3)
    keep = O.Keep()
    om = O.radial_medium(keep, xg, tg, Tf, Vf)
    sp = O.species(keep, [1.5], [N], drag=drag)
    opos, omom = pos.copy(order="F"), mom.copy(order="F")
    chunk = 250
    with lmhip.Context(0) as ctx:
        ctx.set_medium_radial(xg, tg, Tf, Vf)
        load_tables(ctx, drag, None, None)
        ctx.set_particles(pos, mom, [1.5])
        ctx.boost_to_lab(p, 0)
        assert O.boost_to_lab_all(p, om, sp, opos, omom, 0) == 0
        for s0 in range(1, steps + 1, chunk):
            n = min(chunk, steps + 1 - s0)
            ctx.run(p, s0, n, None, 0)
            rec = O.noise_record(p, 3, 3, N, s0, n)
            rc, bad = O.run_cartesian(p, om, sp, opos, omom, s0, n, np.ascontiguousarray(rec))
            assert rc == 0, (rc, bad)
        gpos, gmom = ctx.get_particles()
    check((gpos, gmom), (opos, omom))
    pt0 = np.hypot(mom[0], mom[1]).mean()
    pt1 = np.hypot(gmom[0], gmom[1]).mean()
    print(f"<p_T> {pt0:.4f} -> {pt1:.4f} GeV; <|eta_s|> {np.abs(gpos[2]).mean():.4f}", flush=True)
    assert pt1 < pt0                                    # the drag cools the hard tail


def test_config2_driver_gate_R_and_eta(gpu_ok):
    """through simulate_ensemble_bulk(milne=True) with momentum_langevin=false (the reference's
    gate R set-up; quarks at the transverse origin without p_T, so no boost applies):
    <p_z^2> scales as (tau0/tau)^2 to 1e-9, and every quark's eta_s is that of its straight free
    path"""
    rng = np.random.default_rng(4)
    N = 4000
    (xg, tg, Tf, Vf), _, _, _ = config2_state(N, rng)
    y = rng.uniform(-1.0, 1.0, N)
    mom = np.asfortranarray(np.vstack([np.zeros(N), np.zeros(N), 1.5 * np.sinh(y)]))
    t0, dt, steps = 0.6, 0.005, 1942
    tp, mh, ph = lmhip.simulate_ensemble_bulk(
        lmhip.HIPBackend(), xg, np.linspace(0, 10, 10), None, Tf, Vf, (xg, tg),
        N_particles=N, Δt=dt, initial_time=t0, final_time=t0 + (steps + 0.5) * dt, save_interval=steps * dt,
        m=1.5, DsT=0.6366, dimensions=2, momentum_dimensions=3, x_init=np.zeros((2, N)), p_init=mom,
        momentum_langevin=False, milne=True, seed=1)
    assert ph[-1].shape == (3, N) and mh[-1].shape == (3, N)
    tau = t0 + steps * dt
    assert np.mean(mh[-1][2] ** 2) == pytest.approx(np.mean(mom[2] ** 2) * (t0 / tau) ** 2, rel=1e-9)
    # straight line z = tanh(y)(t - t0): t from t^2 - z^2 = tau^2, eta_s = atanh(z/t)
    v = np.tanh(y)
    a, b, c = 1 - v * v, 2 * v * v * t0, -(v * v * t0 * t0 + tau * tau)
    t = (-b + np.sqrt(b * b - 4 * a * c)) / (2 * a)
    eta = np.arctanh(v * (t - t0) / t)
    assert np.allclose(ph[-1][2], eta, rtol=1e-11, atol=1e-13)
