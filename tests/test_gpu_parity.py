"""Per-particle parity of the HIP step with the CPU oracle, through the C ABI.

Bar: every particle's momentum and position agree with the oracle to
||delta|| / max(||ref||, 1e-6) <= 1e-12 (FP64; the kernels and the oracle share the fixed-noise
record, or the Philox record the oracle regenerates from the same counters).
"""
import math
import zlib

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import flowing_medium, rel_dev, run_grid_pair, run_radial_pair, small_grid
from lmhip import _ffi, media

pytestmark = pytest.mark.gpu
TOL = 1e-12
P_FLOOR = 1e-6
X_FLOOR = 1e-6


def check(g, o, tol=TOL):
    (gp, gm), (op, om) = g, o
    assert np.all(np.isfinite(gm)) and np.all(np.isfinite(gp))
    dm = rel_dev(gm, om, P_FLOOR).max()
    dx = rel_dev(gp, op, X_FLOOR).max()
    print(f"max rel dev: mom {dm:.2e} pos {dx:.2e}", flush=True)
    assert dm <= tol and dx <= tol, (dm, dx)


def test_config1_uniform_bath_full(gpu_ok):
    """The CPU-runnable workload in full: 1e4 charm quarks (m=1.5), 3 momentum dims, T=0.3 at
    rest, D_s(2 pi T)=4, Cartesian, dt=0.01, 500 steps, |p| ~ U[0,10] isotropic, x=0, fixed
    noise from a seed; the driver's initial boost included."""
    rng = np.random.default_rng(1)
    N, m = 10_000, 1.5
    pos = np.zeros((3, N), order="F")
    mom = media.isotropic_momenta(N, 10.0, 3, rng)
    xg, tg, Tf, Vf = media.uniform_bath(0.3)
    p = _ffi.make_params(dt=0.01, t0=0.0, DsT=media.DsT_from_2piT(4.0), noise=_ffi.LMH_NOISE_FIXED)
    noise = np.random.default_rng(2).standard_normal((500, 3, N))
    g, o, _ = run_radial_pair(p, (xg, tg, Tf, Vf, None), [m], [N], pos, mom, 500, noise=noise,
                              initial_boost=True)
    check(g, o)


CASES = {
    "cart2": dict(dims=2, pdim=2),
    "cart3": dict(dims=3, pdim=3),
    "plane_pz": dict(dims=2, pdim=3),
    "plane_pz_bjorken": dict(dims=2, pdim=3, bjorken_redshift=True),
    "radial_mode": dict(dims=1, pdim=1),
    "nonrel": dict(dims=2, pdim=3, relativistic=False),
    "nonrel_radial": dict(dims=1, pdim=1, relativistic=False),
    "v2_modulation": dict(dims=2, pdim=2, v2=True, psi2=0.3),
    "fluid": dict(dims=2, pdim=3, collision=_ffi.LMH_FLUID, bjorken_redshift=True),
    "fluid_radial": dict(dims=1, pdim=1, collision=_ffi.LMH_FLUID),
    "fluid_nonrel": dict(dims=3, pdim=3, collision=_ffi.LMH_FLUID, relativistic=False),
    "rta": dict(dims=2, pdim=3, collision=_ffi.LMH_RTA),
    "rta_radial": dict(dims=1, pdim=1, collision=_ffi.LMH_RTA),
    "langevin_rta": dict(dims=2, pdim=3, collision=_ffi.LMH_LANGEVIN_RTA),
    "midpoint": dict(dims=2, pdim=3, integrator=_ffi.LMH_MIDPOINT),
    "postpoint": dict(dims=2, pdim=3, integrator=_ffi.LMH_POSTPOINT),
    "gamma_p_postpoint": dict(dims=2, pdim=3, integrator=_ffi.LMH_POSTPOINT, drag_model=_ffi.LMH_DRAG_GAMMA_P),
    "diffusion": dict(dims=2, pdim=2, position_diffusion=True),
    "diffusion_radial": dict(dims=1, pdim=1, position_diffusion=True),
    "reflecting": dict(dims=2, pdim=2, reflecting_boundary=True),
    "reflecting_radial": dict(dims=1, pdim=1, reflecting_boundary=True),
    "dst_zero_table": dict(dims=2, pdim=2, DsT=-1.0),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("noise", ["fixed", "philox"])
def test_radial_medium_parity(gpu_ok, name, noise):
    c = dict(CASES[name])
    dims, pdim = c.pop("dims"), c.pop("pdim")
    use_v2 = c.pop("v2", False)
    DsT = c.pop("DsT", media.DsT_from_2piT(4.0))
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    N = 3000
    xg, tg, Tf, Vf, V2 = flowing_medium()
    pos = np.asfortranarray(rng.normal(0, 4.0, (dims, N)))
    if dims == 1:
        pos = np.abs(pos)
    if c.get("reflecting_boundary"):
        pos *= 4.0                                          # many particles beyond r_max = 20
    mom = np.asfortranarray(rng.normal(0, 1.5, (pdim, N)))
    nm = _ffi.LMH_NOISE_FIXED if noise == "fixed" else _ffi.LMH_NOISE_PHILOX
    p = _ffi.make_params(dt=0.01, t0=0.5, DsT=DsT, noise=nm, seed=1234, stream=7, gamma0=5.3, Lambda=5.0, **c)
    rta = p.collision in (_ffi.LMH_RTA, _ffi.LMH_LANGEVIN_RTA)
    g, o, _ = run_radial_pair(p, (xg, tg, Tf, Vf, V2 if use_v2 else None), [1.5], [N], pos, mom, 40,
                              seed=3, rta=rta)
    check(g, o)


def test_species_segments(gpu_ok):
    """70/30 charm/bottom in one run equals the oracle's two segments (each its own mass and
    tables)"""
    rng = np.random.default_rng(11)
    N = 5000
    counts = [3500, 1500]
    xg, tg, Tf, Vf, V2 = flowing_medium()
    pos = np.asfortranarray(rng.normal(0, 4.0, (2, N)))
    mom = np.asfortranarray(rng.normal(0, 1.5, (3, N)))
    p = _ffi.make_params(dt=0.01, t0=0.5, DsT=0.6366, noise=_ffi.LMH_NOISE_PHILOX, seed=5,
                         collision=_ffi.LMH_LANGEVIN_RTA)
    g, o, _ = run_radial_pair(p, (xg, tg, Tf, Vf, None), [1.5, 4.75], counts, pos, mom, 30, rta=True)
    check(g, o)


@pytest.mark.parametrize("spl", [1, 3, 16])
def test_steps_per_launch_invariance(gpu_ok, spl):
    """keeping particles in registers for k steps per launch changes nothing"""
    rng = np.random.default_rng(5)
    N = 4000
    xg, tg, Tf, Vf, V2 = flowing_medium()
    pos = np.asfortranarray(rng.normal(0, 4.0, (2, N)))
    mom = np.asfortranarray(rng.normal(0, 1.5, (3, N)))
    p = _ffi.make_params(dt=0.01, t0=0.5, DsT=0.6366, noise=_ffi.LMH_NOISE_PHILOX, seed=9)
    res = []
    for k in (1, spl):
        with lmhip.Context(0) as ctx:
            ctx.set_medium_radial(xg, tg, Tf, Vf)
            d = lmhip.build_tau_drag_spline(1.5, 0.6366, Tmin=float(Tf.min()), Tmax=float(Tf.max()), nT=1024)
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, 0, *d)
            ctx.set_particles(pos, mom, [1.5])
            ctx.run(p, 1, 37, None, k)
            res.append(ctx.get_particles())
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("noise", ["fixed", "philox"])
def test_general_coords_parity(gpu_ok, noise):
    rng = np.random.default_rng(21)
    N, m = 4000, 1.5
    xg, tg, Tf, Vf, _ = flowing_medium()
    r = np.abs(rng.normal(3.0, 2.0, N))
    pr = rng.normal(0, 1.0, N)
    pos = np.asfortranarray(np.vstack([np.full(N, 0.6), r]))
    mom = np.asfortranarray(np.vstack([np.sqrt(m * m + pr * pr), np.abs(pr)]))
    nm = _ffi.LMH_NOISE_FIXED if noise == "fixed" else _ffi.LMH_NOISE_PHILOX
    p = _ffi.make_params(dt=0.005, t0=0.6, DsT=0.6366, coords=_ffi.LMH_GENERAL_COORDS, noise=nm, seed=77)
    g, o, _ = run_radial_pair(p, (xg, tg, Tf, Vf, None), [m], [N], pos, mom, 60)
    check(g, o)


GRID_CASES = {
    "langevin": dict(),
    "langevin_rta": dict(collision=_ffi.LMH_LANGEVIN_RTA),
    "rta": dict(collision=_ffi.LMH_RTA),
    "fluid": dict(collision=_ffi.LMH_FLUID),
    "nonrel": dict(relativistic=False),
    "midpoint": dict(integrator=_ffi.LMH_MIDPOINT),
    "gamma_p_post": dict(integrator=_ffi.LMH_POSTPOINT, drag_model=_ffi.LMH_DRAG_GAMMA_P),
    "diffusion": dict(position_diffusion=True),
    "bjorken": dict(bjorken_redshift=True),
}


@pytest.mark.parametrize("name", list(GRID_CASES))
@pytest.mark.parametrize("noise", ["fixed", "philox"])
@pytest.mark.parametrize("spl", [1, 5])
def test_grid_parity(gpu_ok, name, noise, spl):
    """2+1D grid medium with two species; spl=1 reads the medium slice interpolated in tau once
    per step, spl=5 interpolates every corner per particle: both round identically."""
    g_ = small_grid()
    rng = np.random.default_rng(31)
    N = 6000
    counts = [4200, 1800]
    pos = media.hydro_positions(N, rng) * 2.0
    mom = np.asfortranarray(np.hstack([media.pt_rapidity_momenta(4200, 1.5, rng),
                                       media.pt_rapidity_momenta(1800, 4.75, rng)]))
    nm = _ffi.LMH_NOISE_FIXED if noise == "fixed" else _ffi.LMH_NOISE_PHILOX
    p = _ffi.make_params(dt=0.02, t0=0.6, DsT=0.6366, noise=nm, seed=2024, gamma0=5.3, Lambda=5.0,
                         **GRID_CASES[name])
    rta = p.collision in (_ffi.LMH_RTA, _ffi.LMH_LANGEVIN_RTA)
    g, o = run_grid_pair(p, g_, [1.5, 4.75], counts, pos, mom, 30, spl=spl, rta=rta,
                         rta_time=media.tau_R_law(0.16))
    # position diffusion: D = (D_s T / T) / hbar c, and this medium's cold corona (T falling to
    # 1e-10 GeV) makes a quark's jump there hundreds of fm and T at its landing point steep in
    # the position, so the dynamics amplifies ulp-level differences step after step; the kernels
    # therefore push with the reference form x + dt p/E and draw the diffusion normals with the
    # host's correctly rounded operations when position diffusion is on, and the bar is 1e-12
    check(g, o)


@pytest.mark.parametrize("noise", ["fixed", "philox"])
@pytest.mark.parametrize("spl", [1, 5])
def test_grid_parity_diffusion_warm_medium(gpu_ok, noise, spl):
    """position diffusion at the 1e-12 bar on the same medium with T floored at 0.1 GeV: the
    looser bar of the diffusion case above comes from the cold corona alone"""
    import dataclasses
    g0 = small_grid()
    g_ = dataclasses.replace(g0, T=np.maximum(g0.T, 0.1))
    rng = np.random.default_rng(31)
    N = 6000
    counts = [4200, 1800]
    pos = media.hydro_positions(N, rng) * 2.0
    mom = np.asfortranarray(np.hstack([media.pt_rapidity_momenta(4200, 1.5, rng),
                                       media.pt_rapidity_momenta(1800, 4.75, rng)]))
    nm = _ffi.LMH_NOISE_FIXED if noise == "fixed" else _ffi.LMH_NOISE_PHILOX
    p = _ffi.make_params(dt=0.02, t0=0.6, DsT=0.6366, noise=nm, seed=2024, position_diffusion=True)
    g, o = run_grid_pair(p, g_, [1.5, 4.75], counts, pos, mom, 30, spl=spl)
    check(g, o)


@pytest.mark.parametrize("slice_", [0, 1])
def test_grid_deferred_collisions_identical(gpu_ok, slice_):
    """queueing the colliding particles for a second kernel gives the fused kernel's bits"""
    g_ = small_grid()
    rng = np.random.default_rng(41)
    N = 20000
    pos = media.hydro_positions(N, rng) * 2.0
    mom = media.pt_rapidity_momenta(N, 1.5, rng)
    p = _ffi.make_params(dt=0.05, t0=0.6, DsT=0.6366, noise=_ffi.LMH_NOISE_PHILOX, seed=8,
                         collision=_ffi.LMH_LANGEVIN_RTA)
    from helpers import tables, load_tables
    drag, rt, jt = tables([1.5], g_.T, 0.6366, 3, True, media.tau_R_law(0.16))
    res = []
    for defer in (0, 1):
        with lmhip.Context(0) as ctx:
            ctx.set_tuning(block=256, steps_per_launch=1, slice=slice_, defer=defer)
            ctx.set_medium_grid2p1(g_.x0, g_.dx, g_.nx, g_.y0, g_.dy, g_.ny, g_.tgrid, g_.T, g_.vx, g_.vy)
            load_tables(ctx, drag, rt, jt)
            ctx.set_particles(pos, mom, [1.5])
            ctx.run(p, 1, 25, None, 0)
            res.append(ctx.get_particles())
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    # and collisions did happen
    assert np.mean(np.any(res[1][1] != mom, axis=0)) > 0.5


def test_cell_sort_identical(gpu_ok):
    """ordering the quarks by grid cell changes no result: two species with RTA, re-sorts inside
    a run and the caller's order restored between runs give the unsorted run's bits"""
    g_ = small_grid()
    rng = np.random.default_rng(43)
    counts = [12000, 8000]
    N = sum(counts)
    pos = media.hydro_positions(N, rng) * 2.0
    mom = np.asfortranarray(np.hstack([media.pt_rapidity_momenta(counts[0], 1.5, rng),
                                       media.pt_rapidity_momenta(counts[1], 4.75, rng)]))
    p = _ffi.make_params(dt=0.05, t0=0.6, DsT=0.6366, noise=_ffi.LMH_NOISE_PHILOX, seed=9,
                         collision=_ffi.LMH_LANGEVIN_RTA)
    from helpers import tables, load_tables
    drag, rt, jt = tables([1.5, 4.75], g_.T, 0.6366, 3, True, media.tau_R_law(0.16))
    res = {}
    for every in (0, 4, 1):
        with lmhip.Context(0) as ctx:
            ctx.set_tuning(sort_every=every)
            ctx.set_medium_grid2p1(g_.x0, g_.dx, g_.nx, g_.y0, g_.dy, g_.ny, g_.tgrid, g_.T, g_.vx, g_.vy)
            load_tables(ctx, drag, rt, jt)
            ctx.set_particles(pos, mom, [1.5, 4.75], counts)
            ctx.profile_start()
            ctx.run(p, 1, 10, None, 0)
            obs = ctx.observables(20, 0.0, 10.0, 0.0, -1)
            mid = ctx.get_particles()
            ctx.run(p, 11, 15, None, 0)
            prof = ctx.profile_stop()
            res[every] = (obs, mid, ctx.get_particles(), prof["sort"]["launches"])
    for every in (4, 1):
        assert res[every][3] > 0, "no sort ran"
        assert np.array_equal(res[every][0], res[0][0])
        for k in (1, 2):
            assert np.array_equal(res[every][k][0], res[0][k][0]) and np.array_equal(res[every][k][1], res[0][k][1])
    assert res[0][3] == 0


def test_nonfinite_reported_at_first_bad_step(gpu_ok):
    rng = np.random.default_rng(0)
    N = 1000
    xg, tg, Tf, Vf = media.uniform_bath(0.3)
    pos = np.asfortranarray(rng.normal(0, 1, (3, N)))
    mom = np.asfortranarray(rng.normal(0, 1, (3, N)))
    p = _ffi.make_params(dt=0.01, DsT=0.6366, noise=_ffi.LMH_NOISE_FIXED)
    noise = rng.standard_normal((10, 3, N))
    noise[6, 2, 500] = np.inf
    with lmhip.Context(0) as ctx:
        ctx.set_medium_radial(xg, tg, Tf, Vf)
        ctx.set_time_table(_ffi.LMH_TABLE_DRAG, 0, *lmhip.build_tau_drag_spline(1.5, 0.6366, Tmin=0.3, Tmax=0.3 + 1e-15, nT=1024))
        ctx.set_particles(pos, mom, [1.5])
        with pytest.raises(lmhip.LMHError, match="NaN/Inf at step 10"):
            ctx.run(p, 4, 10, noise, 0)
        assert ctx.nonfinite_step() == 10


def test_observables_parity(gpu_ok):
    rng = np.random.default_rng(8)
    N = 300_001
    pos = np.zeros((2, N), order="F")
    mom = np.asfortranarray(rng.normal(0, 2.0, (3, N)))
    with lmhip.Context(0) as ctx:
        ctx.set_medium_radial(*media.uniform_bath(0.3))
        ctx.set_particles(pos, mom, [1.5, 4.75], [200_000, 100_001])
        for sigma in (0.0, 0.25):
            for sp, (i0, i1) in ((-1, (0, N)), (1, (200_000, N))):
                got = ctx.observables(40, 0.0, 8.0, sigma, sp)
                want = O.observables(mom, i0, i1, 40, 0.0, 8.0, sigma)
                assert got[0] == want[0]
                assert np.allclose(got[1:5], want[1:5], rtol=1e-12)
                if sigma == 0.0:
                    assert np.array_equal(got[5:], want[5:])
                else:
                    assert np.allclose(got[5:], want[5:], rtol=1e-11, atol=1e-9)


def test_simulate_ensemble_bulk_matches_oracle_driver(gpu_ok):
    """the host driver (initial boost, chunks between saves, snapshots) in fixed-noise mode"""
    rng = np.random.default_rng(3)
    N = 2000
    xg, tg, Tf, Vf, _ = flowing_medium()
    x0 = np.asfortranarray(rng.normal(0, 3, (2, N)))
    p0 = np.asfortranarray(rng.normal(0, 1, (3, N)))
    DsT, dt = 0.6366, 0.01
    t0, t1 = 0.5, 0.955
    steps, save_every, num_saves = lmhip.utils.step_counts(t0, t1, dt, 0.1)
    assert (steps, save_every, num_saves) == (45, 10, 4)
    noise = rng.standard_normal((steps, 3, N))
    tp, mh, ph = lmhip.simulate_ensemble_bulk(
        lmhip.HIPBackend(), xg, np.linspace(0, 10, 10), None, Tf, Vf, (xg, tg),
        N_particles=N, Δt=dt, initial_time=t0, final_time=t1, save_interval=0.1,
        m=1.5, DsT=DsT, dimensions=2, momentum_dimensions=3, x_init=x0, p_init=p0, noise=noise)
    assert len(mh) == len(ph) == len(tp) == 5
    keep = O.Keep()
    om = O.radial_medium(keep, xg, tg, Tf, Vf)
    d = lmhip.build_tau_drag_spline(1.5, DsT, Tmin=float(Tf.min()), Tmax=float(Tf.max()), nT=1024)
    sp = O.species(keep, [1.5], [N], drag=[d])
    p = _ffi.make_params(dt=dt, t0=0.5, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    P, M = x0.copy(order="F"), p0.copy(order="F")
    assert O.boost_to_lab_all(p, om, sp, P, M, 0) == 0
    check((ph[0], mh[0]), (P, M))
    for k in range(4):
        rc, _ = O.run_cartesian(p, om, sp, P, M, 1 + 10 * k, 10, np.ascontiguousarray(noise[10 * k:10 * k + 10]))
        assert rc == 0
        check((ph[k + 1], mh[k + 1]), (P, M))


def test_general_coords_driver_shapes(gpu_ok):
    xg, tg, Tf, Vf, _ = flowing_medium()
    prof = lambda r, t: 0.3
    zero = lambda r, t: 0.0
    tp, mh, ph = lmhip.simulate_ensemble_bulk(
        lmhip.HIP_GCBackend(), prof, zero, zero, Tf, Vf, (xg, tg), N_particles=5000, Δt=0.01,
        initial_time=0.6, final_time=1.6, save_interval=0.25, m=1.5, DsT=0.6366, seed=3)
    assert len(tp) == len(mh) == len(ph) == 5
    assert mh[0].shape == (5000,) and ph[0].shape == (1, 5000)
    assert np.all(np.isfinite(mh[-1])) and np.all(mh[-1] >= 1.5 * (1 - 1e-6))


@pytest.mark.parametrize("medium,defer", [("grid", 0), ("grid", 1), ("radial", 0)])
def test_specialized_kernels_identical_to_generic(gpu_ok, medium, defer):
    """the kernels compiled for the common configuration give the generic kernels' bits, with the
    collisions redrawn in the step kernel or queued for a second one"""
    N = 5000
    res = []
    for spec in (0, 1):
        with lmhip.Context(0) as ctx:
            ctx.set_tuning(block=256, steps_per_launch=1, slice=1, defer=defer, specialize=spec)
            if medium == "grid":
                g_ = small_grid()
                pos = media.hydro_positions(N, np.random.default_rng(77)) * 2.0
                mom = media.pt_rapidity_momenta(N, 1.5, np.random.default_rng(78))
                from helpers import tables, load_tables
                drag, rt, jt = tables([1.5], g_.T, 0.6366, 3, True, media.tau_R_law(0.16))
                ctx.set_medium_grid2p1(g_.x0, g_.dx, g_.nx, g_.y0, g_.dy, g_.ny, g_.tgrid, g_.T, g_.vx, g_.vy)
                load_tables(ctx, drag, rt, jt)
                coll = _ffi.LMH_LANGEVIN_RTA
            else:
                xg, tg, Tf, Vf, V2 = flowing_medium()
                pos = np.asfortranarray(np.random.default_rng(79).normal(0, 3, (3, N)))
                mom = np.asfortranarray(np.random.default_rng(80).normal(0, 2, (3, N)))
                ctx.set_medium_radial(xg, tg, Tf, Vf, V2)
                from helpers import tables, load_tables
                drag, rt, jt = tables([1.5], Tf, 0.6366, 3)
                load_tables(ctx, drag, rt, jt)
                coll = _ffi.LMH_LANGEVIN
            p = _ffi.make_params(dt=0.02, t0=0.6, DsT=0.6366, noise=_ffi.LMH_NOISE_PHILOX, seed=5,
                                 collision=coll)
            ctx.set_particles(pos, mom, [1.5])
            ctx.run(p, 1, 20, None, 0)
            res.append(ctx.get_particles())
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_snapshots_equal_synchronous_downloads(gpu_ok):
    """lmh_snapshot copies the state in stream order while later steps run: every snapshot equals
    the state lmh_get_particles returns at that point of a second, synchronous run"""
    g_ = small_grid()
    N = 20000
    pos = media.hydro_positions(N, np.random.default_rng(3)) * 2.0
    mom = media.pt_rapidity_momenta(N, 1.5, np.random.default_rng(4))
    from helpers import tables, load_tables
    drag, rt, jt = tables([1.5], g_.T, 0.6366, 3, True, media.tau_R_law(0.16))
    p = _ffi.make_params(dt=0.02, t0=0.6, DsT=0.6366, noise=_ffi.LMH_NOISE_PHILOX, seed=9,
                         collision=_ffi.LMH_LANGEVIN_RTA, check_finite=False)
    res = []
    for async_ in (False, True):
        with lmhip.Context(0) as ctx:
            ctx.set_tuning(block=256, steps_per_launch=1, slice=1, defer=1)
            ctx.set_medium_grid2p1(g_.x0, g_.dx, g_.nx, g_.y0, g_.dy, g_.ny, g_.tgrid, g_.T, g_.vx, g_.vy)
            load_tables(ctx, drag, rt, jt)
            ctx.set_particles(pos, mom, [1.5])
            out, idx = [], []
            for c in range(5):
                ctx.run(p, 1 + 7 * c, 7)
                if async_:
                    idx.append(ctx.snapshot())
                else:
                    out.append(ctx.get_particles())
            if async_:
                out = [ctx.snapshot_get(k) for k in idx]
                # freed indices are reused, and a snapshot is fetched once
                k = ctx.snapshot()
                assert k in idx
                ctx.snapshot_get(k)
                with pytest.raises(lmhip.LMHError):
                    ctx.snapshot_get(k)
            res.append(out)
    for (sp, sm), (gp, gm) in zip(res[0], res[1]):
        assert np.array_equal(sp, gp) and np.array_equal(sm, gm)
