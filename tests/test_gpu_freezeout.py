"""On-the-fly freeze-out capture (lmh_freezeout_enable / lmh_freezeout_get) against the oracle's
restatement of kernel_capture_freezeout_gpu! (kernels_gpu.jl:583-635), which keeps its own
last-above cache and runs as a separate pass after every step; the kernels latch inside the
step from the step's input state.  Fixed noise, runs split over several lmh_run calls so that
the per-particle last temperature is carried between launches.  Bar: 1e-12 relative."""
from __future__ import annotations

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import (flowing_medium, load_tables, rel_dev, rta_noise_fixup, small_grid, tables)
from lmhip import _ffi, media

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _compare(g, o, N):
    (gp, gm, gt, gf), fo = g, o
    assert np.array_equal(gf, fo.fo_flag)
    sel = gf == 1.0
    assert sel.sum() > 0.1 * N, "too few crossings to test"
    assert np.all(rel_dev(gm[:, sel], fo.fo_mom[:, sel], 1e-6) <= TOL)
    assert np.all(rel_dev(gp[:, sel], fo.fo_pos[:, sel], 1e-6) <= TOL)
    assert np.all(np.abs(gt[sel] - fo.fo_tau[sel]) <= TOL * np.abs(fo.fo_tau[sel]))


@pytest.mark.parametrize("dims,interp,spl", [(3, True, 1), (3, False, 4), (1, True, 3), (2, True, 1)])
def test_freezeout_radial_medium(gpu_ok, dims, interp, spl):
    xg, tg, Tf, Vf, V2 = flowing_medium()
    N, steps, chunks = 3000, 60, (25, 20, 15)
    rng = np.random.default_rng(dims * 10 + spl)
    pos = np.asfortranarray(rng.normal(0, 4, (dims, N)))
    pdim = dims
    mom = np.asfortranarray(rng.normal(0, 1.5, (pdim, N)))
    p = _ffi.make_params(dt=0.1, t0=0.0, DsT=0.6366, noise=_ffi.LMH_NOISE_FIXED)
    drag, rt, jt = tables([1.5], Tf, 0.6366, pdim)
    C = lmhip.Context(0)
    noise = rng.standard_normal((steps, pdim, N))
    Tfo = 0.25
    with C as ctx:
        ctx.set_medium_radial(xg, tg, Tf, Vf, None)
        load_tables(ctx, drag, rt, jt)
        ctx.set_particles(pos, mom, [1.5])
        ctx.freezeout_enable(Tfo, interp)
        s = 1
        for n in chunks:
            ctx.run(p, s, n, np.ascontiguousarray(noise[s - 1:s - 1 + n]), spl)
            s += n
        got = ctx.freezeout_get()
    keep = O.Keep()
    med = O.radial_medium(keep, xg, tg, Tf, Vf)
    sp = O.species(keep, [1.5], [N], drag=drag)
    op, om = pos.copy(order="F"), mom.copy(order="F")
    fo = O.Freezeout(dims, pdim, N)
    for s in range(1, steps + 1):
        rc, _ = O.run_cartesian(p, med, sp, op, om, s, 1, np.ascontiguousarray(noise[s - 1:s]))
        assert rc == 0
        fo.capture(med, None, op, om, s, p.dt, p.t0, Tfo, interp)
    _compare(got, fo, N)


@pytest.mark.parametrize("collision", [_ffi.LMH_LANGEVIN, _ffi.LMH_LANGEVIN_RTA])
def test_freezeout_grid_medium(gpu_ok, collision):
    g_ = small_grid()
    N, steps, chunks = 4000, 50, (1, 29, 20)
    rng = np.random.default_rng(collision + 5)
    pos = media.hydro_positions(N, rng) * 2.0
    mom = media.pt_rapidity_momenta(N, 1.5, rng)
    p = _ffi.make_params(dt=0.05, t0=0.6, DsT=0.6366, noise=_ffi.LMH_NOISE_FIXED, collision=collision)
    rta = collision == _ffi.LMH_LANGEVIN_RTA
    drag, rt, jt = tables([1.5], g_.T, 0.6366, 3, rta, media.tau_R_law(0.16))
    Tfo = 0.2
    with lmhip.Context(0) as ctx:
        ctx.set_tuning(block=256, steps_per_launch=1, slice=1, defer=1)
        ctx.set_medium_grid2p1(g_.x0, g_.dx, g_.nx, g_.y0, g_.dy, g_.ny, g_.tgrid, g_.T, g_.vx, g_.vy)
        load_tables(ctx, drag, rt, jt)
        ctx.set_particles(pos, mom, [1.5])
        ctx.freezeout_enable(Tfo, True)
        noise = rta_noise_fixup(p, rng.standard_normal((steps, ctx.noise_channels(p), N)), 2, 3)
        s = 1
        for n in chunks:
            ctx.run(p, s, n, np.ascontiguousarray(noise[s - 1:s - 1 + n]), 0)
            s += n
        got = ctx.freezeout_get()
        # the capture state is reset with the particles
        ctx.set_particles(pos, mom, [1.5])
        assert np.all(ctx.freezeout_get()[3] == 0.0)
    keep = O.Keep()
    gm = O.grid_medium(keep, g_)
    sp = O.species(keep, [1.5], [N], drag=drag, rta=rt, jt=jt)
    op, om = pos.copy(order="F"), mom.copy(order="F")
    fo = O.Freezeout(2, 3, N)
    for s in range(1, steps + 1):
        rc, _ = O.run_grid(p, gm, sp, op, om, s, 1, np.ascontiguousarray(noise[s - 1:s]))
        assert rc == 0
        fo.capture(None, gm, op, om, s, p.dt, p.t0, Tfo, True)
    _compare(got, fo, N)


def test_freezeout_errors(gpu_ok):
    with lmhip.Context(0) as ctx:
        with pytest.raises(lmhip.LMHError, match="set the particles"):
            ctx.freezeout_enable(0.15)
        xg, tg, Tf, Vf, _ = flowing_medium()
        ctx.set_medium_radial(xg, tg, Tf, Vf, None)
        ctx.set_particles(np.zeros((2, 10), order="F"), np.zeros((2, 10), order="F"), [1.5])
        with pytest.raises(lmhip.LMHError, match="not enabled"):
            ctx.freezeout_get()
        ctx.freezeout_enable(0.15)
        p = _ffi.make_params(dt=0.01, coords=_ffi.LMH_GENERAL_COORDS, noise=_ffi.LMH_NOISE_PHILOX, DsT=0.6)
        with pytest.raises(lmhip.LMHError, match="Cartesian"):
            ctx.run(p, 1, 1)


def test_freezeout_driver_bjorken(gpu_ok):
    """simulate_ensemble_bulk(..., freezeout_capture=true) returns the latched states
    (simulate_gpu.jl:394-402); in a Bjorken bath every quark latches at tau0 (T0/Tfo)^3"""
    xg, tg, Tf, Vf, _ = media.bjorken_bath(T0=0.4, tau0=0.6, Tstop=0.15, ntau=4096)
    N = 2000
    rng = np.random.default_rng(12)
    x0 = np.asfortranarray(rng.normal(0, 2, (3, N)))
    p0 = media.isotropic_momenta(N, 5.0, 3, rng)
    Tfo = 0.2
    fo = lmhip.simulate_ensemble_bulk(lmhip.HIPBackend(), xg, None, None, Tf, Vf, (xg, tg),
                                      x_init=x0, p_init=p0, N_particles=N, Δt=0.01, initial_time=0.6,
                                      final_time=5.2, m=1.5, DsT=0.6366, dimensions=3, Tfo=Tfo,
                                      freezeout_capture=True, freezeout_interp=True)
    assert set(fo) == {"pos", "mom", "tau", "flag"}
    assert fo["pos"].shape == (3, N) and fo["mom"].shape == (3, N)
    assert np.all(fo["flag"] == 1.0)
    assert np.max(np.abs(fo["tau"] - 0.6 * (0.4 / Tfo) ** 3)) < 1e-5
