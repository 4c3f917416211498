"""The benchmarked workloads at their full shapes through the C ABI, against the oracle.

* The 2+1D hydro workload exactly as timed (media.hydro_ensemble / media.hydro_tables: 10^7
  quarks, 70% charm and 30% bottom, on the 200 x 200 x 120 grid; Langevin kicks with RTA
  collisions of tau_R = 5 (eta/s)/T; Philox noise; dt = 0.01; the specialised step kernel and
  the launch shape the benchmark uses), the driver's initial boost included, 40 steps with the
  quarks re-sorted by cell at steps 1, 17 and 33: every quark against the oracle at 1e-12.
* 256 batched runs on grids of that shape (D_s(2 pi T), T0, eccentricity and a log-uniform
  particle count drawn per run, the counts cut to 10^3..10^4 so the check runs in seconds):
  every run bit for bit the same run alone, and the largest run against the oracle at 1e-12.
* The differentiation workload's shape (10^6 charm quarks x 1000 steps of dt = 0.01 fm/c on a
  200 x 200 x 120 grid, checkpoints every 50 steps; the <p_T>, smoothed 40-bin spectrum and
  <cos 2 phi> losses; d/dD_sT with the drag table, d/dgamma0 and d/dLambda with gamma(p, T)) on
  a medium whose fields are affine in (x, y, tau): its interpolation has no kinks at cell
  boundaries, so central differences at the stated relative step 1e-4 (and 5e-5, combined by
  Richardson's rule to cancel their O(h^2) error) are the smooth pathwise derivative, and the
  adjoint must agree with them to 1e-4 relative.
"""
import math

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import rel_dev
from lmhip import _ffi, media, transport
from lmhip.batch import Batch

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _report(name, gpos, gmom, opos, omom):
    dm = rel_dev(gmom, omom, 1e-6)
    dx = rel_dev(gpos, opos, 1e-6)
    print(f"{name}: max rel dev mom {dm.max():.2e} pos {dx.max():.2e}; "
          f"quarks above 1e-13: {int(np.sum((dm > 1e-13) | (dx > 1e-13)))}", flush=True)
    assert np.all(np.isfinite(gmom)) and np.all(np.isfinite(gpos))
    assert dm.max() <= TOL and dx.max() <= TOL, (dm.max(), dx.max())


def test_hydro_workload_full_size(gpu_ok):
    N, steps, chunk = 10_000_000, 40, 4
    g, masses, counts, pos, mom = media.hydro_ensemble(N)
    DsT = media.DsT_from_2piT(4.0)
    tabs = media.hydro_tables(g, masses, DsT, 3)
    p = _ffi.make_params(dt=0.01, t0=0.6, DsT=DsT, seed=3, stream=0, collision=_ffi.LMH_LANGEVIN_RTA,
                         noise=_ffi.LMH_NOISE_PHILOX, check_finite=False)
    print(f"hydro workload: N={N} steps={steps} on the GPU", flush=True)
    with lmhip.Context(0) as ctx:
        ctx.set_tuning(block=256, steps_per_launch=1, slice=1, defer=0, sort_every=16)
        ctx.set_medium_grid2p1(g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, g.tgrid, g.T, g.vx, g.vy)
        for k, (drag, rta, jt) in enumerate(tabs):
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, k, *drag)
            ctx.set_time_table(_ffi.LMH_TABLE_RTA, k, *rta)
            ctx.set_juttner_table(k, *jt)
        ctx.set_particles(pos, mom, masses, counts)
        ctx.boost_to_lab(p, 0)
        ctx.profile_start()
        ctx.run(p, 1, steps, None, 0)
        prof = ctx.profile_stop()
        gpos, gmom = ctx.get_particles()
    assert prof["sort"]["launches"] > 0, "no re-sort ran"
    print(f"GPU done ({prof['sort']['launches']} sort launches); oracle", flush=True)
    keep = O.Keep()
    gm = O.grid_medium(keep, g)
    sp = O.species(keep, masses, counts, drag=[t[0] for t in tabs], rta=[t[1] for t in tabs],
                   jt=[t[2] for t in tabs])
    opos, omom = pos.copy(order="F"), mom.copy(order="F")
    assert O.boost_to_lab_grid(p, gm, sp, opos, omom, 0) == 0
    for s0 in range(1, steps + 1, chunk):
        rec = O.noise_record(p, 2, 3, N, s0, chunk)
        rc, bad = O.run_grid(p, gm, sp, opos, omom, s0, chunk, rec)
        assert rc == 0, (rc, bad)
        del rec
    _report("hydro workload", gpos, gmom, opos, omom)
    # the collisions and kicks did act on the whole ensemble
    assert np.mean(np.any(gmom != mom, axis=0)) > 0.999


def test_batch_256_runs_full_grid_shape(gpu_ok):
    R, steps, m, stream0 = 256, 20, 1.5, 0
    rng = np.random.default_rng(4)
    Ds = rng.uniform(1.5, 7.0, R)
    T0s = rng.uniform(0.3, 0.5, R)
    counts = np.round(10.0 ** rng.uniform(3.0, 4.0, R)).astype(np.int64)
    eccs = rng.uniform(0.0, 0.4, R)

    def grid(r):
        return media.hydro_grid(nx=200, ny=200, nt=120, L=15.0, tau0=0.6, tau1=12.0, T0=float(T0s[r]),
                                ecc=float(eccs[r]))

    # the temperature range of all runs, as the batch benchmark builds its shared tables
    Tmin, Tmax = math.inf, -math.inf
    for r in range(R):
        T = grid(r).T
        Tmin, Tmax = min(Tmin, max(float(np.min(T)), 0.0)), max(Tmax, float(np.max(T)))
    rt = transport.build_relaxation_spline(media.tau_R_law(0.16), Tmin=Tmin, Tmax=Tmax, nT=1024)
    jt = transport.build_juttner_invcdf(m, 3, Tmin=Tmin, Tmax=Tmax, nT=256, nU=2048, np_=8192)
    drags = [transport.build_tau_drag_spline(m, media.DsT_from_2piT(float(Ds[r])), Tmin=Tmin, Tmax=Tmax, nT=1024)
             for r in range(R)]
    pos = [media.hydro_positions(int(counts[r]), rng, ecc=float(eccs[r])) for r in range(R)]
    mom = [media.pt_rapidity_momenta(int(counts[r]), m, rng) for r in range(R)]
    p = _ffi.make_params(dt=0.01, t0=0.6, DsT=0.0, seed=3, stream=stream0, collision=_ffi.LMH_LANGEVIN_RTA,
                         noise=_ffi.LMH_NOISE_PHILOX)
    print(f"batch: {R} runs, {int(counts.sum())} quarks, {steps} steps on the GPU", flush=True)
    with Batch(0) as b:
        rta_t = b.add_time_table(*rt)
        jt_t = b.add_juttner_table(*jt)
        desc = []
        for r in range(R):
            med = b.add_medium(grid(r))
            desc.append(dict(count=int(counts[r]), mass=m, DsT=media.DsT_from_2piT(float(Ds[r])), medium=med,
                             drag_table=b.add_time_table(*drags[r]), rta_table=rta_t, juttner_table=jt_t))
        b.set_runs(desc, np.hstack(pos), np.hstack(mom))
        b.boost_to_lab(p, 0)
        tune, _ = b.autotune(p, trial_steps=3)
        # the tuned launch shape with a cell sort every 10 steps (several sorts in the run)
        b.set_tuning(block=tune["block"], steps_per_launch=tune["steps_per_launch"],
                     persistent=tune["persistent"], graph_steps=10, sort_every=10)
        b.run(p, 1, steps)
        bpos, bmom = b.get_particles()
    print(f"batch done (tuning {tune}); the runs alone", flush=True)
    begin = np.concatenate([[0], np.cumsum(counts)])
    largest = int(np.argmax(counts))
    with lmhip.Context(0) as ctx:
        ctx.set_tuning(slice=0)
        for r in range(R):
            g = grid(r)
            pr = _ffi.make_params(dt=0.01, t0=0.6, DsT=media.DsT_from_2piT(float(Ds[r])), seed=3,
                                  stream=stream0 + r, collision=_ffi.LMH_LANGEVIN_RTA, noise=_ffi.LMH_NOISE_PHILOX)
            ctx.set_medium_grid2p1(g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, g.tgrid, g.T, g.vx, g.vy)
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, 0, *drags[r])
            ctx.set_time_table(_ffi.LMH_TABLE_RTA, 0, *rt)
            ctx.set_juttner_table(0, *jt)
            ctx.set_particles(pos[r], mom[r], [m])
            ctx.boost_to_lab(pr, 0)
            ctx.run(pr, 1, steps, None, 0)
            apos, amom = ctx.get_particles()
            sl = slice(begin[r], begin[r + 1])
            assert np.array_equal(bpos[:, sl], apos) and np.array_equal(bmom[:, sl], amom), f"run {r}"
    # the largest run against the oracle
    r = largest
    g = grid(r)
    pr = _ffi.make_params(dt=0.01, t0=0.6, DsT=media.DsT_from_2piT(float(Ds[r])), seed=3, stream=stream0 + r,
                          collision=_ffi.LMH_LANGEVIN_RTA, noise=_ffi.LMH_NOISE_PHILOX)
    keep = O.Keep()
    gm = O.grid_medium(keep, g)
    sp = O.species(keep, [m], [int(counts[r])], drag=[drags[r]], rta=[rt], jt=[jt])
    opos, omom = pos[r].copy(order="F"), mom[r].copy(order="F")
    assert O.boost_to_lab_grid(pr, gm, sp, opos, omom, 0) == 0
    rec = O.noise_record(pr, 2, 3, int(counts[r]), 1, steps)
    rc, bad = O.run_grid(pr, gm, sp, opos, omom, 1, steps, rec)
    assert rc == 0, (rc, bad)
    sl = slice(begin[r], begin[r + 1])
    _report(f"batch run {r} ({counts[r]} quarks)", bpos[:, sl], bmom[:, sl], opos, omom)


def _affine_grid():
    """T = 0.30 + 0.004 x - 0.003 y - 0.008 (tau - 0.6) GeV (0.105 .. 0.405 over the grid), flow
    (0.2 + 0.004 x, 0.1 + 0.003 y): trilinear interpolation reproduces it exactly, |v| is nowhere
    near 0 (where the boost direction would be singular), T nowhere near 0"""
    x = np.linspace(-15.0, 15.0, 200)
    y = np.linspace(-15.0, 15.0, 200)
    tg = np.linspace(0.6, 12.0, 120)
    X, Y, TAU = np.meshgrid(x, y, tg, indexing="ij")
    T = np.asfortranarray(0.30 + 0.004 * X - 0.003 * Y - 0.008 * (TAU - 0.6))
    vx = np.asfortranarray(0.2 + 0.004 * X + 0.0 * TAU)
    vy = np.asfortranarray(0.1 + 0.003 * Y + 0.0 * TAU)
    return media.Grid2p1Medium(float(x[0]), float(x[1] - x[0]), 200, float(y[0]), float(y[1] - y[0]), 200,
                               tg, T, vx, vy)


def test_adjoint_full_shape_fd_h1e4(gpu_ok):
    from lmhip.adjoint import GridEnsemble
    N, steps, K = 1_000_000, 1000, 50
    g = _affine_grid()
    rng = np.random.default_rng(12)
    # |x|, |y| <= 3 fm: in 10 fm/c no quark reaches the grid's edge (no position clamp)
    pos = np.asfortranarray(rng.uniform(-3.0, 3.0, (2, N)))
    mom = media.pt_rapidity_momenta(N, 1.5, rng)
    obs_kw = dict(nbins=40, pt_min=0.0, pt_max=10.0, sigma=0.25)
    w = {"mean_pT": np.zeros(45), "cos2phi": np.zeros(45), "spectrum": np.zeros(45)}
    w["mean_pT"][1] = 1.0 / N
    w["cos2phi"][3] = 1.0 / N
    w["spectrum"][5:] = np.random.default_rng(6).normal(size=40) / N
    print(f"adjoint full shape: N={N} steps={steps} K={K} on the GPU", flush=True)
    with lmhip.Context(0) as ctx:
        for drag_model, theta, ks in ((_ffi.LMH_DRAG_TABLE, np.array([media.DsT_from_2piT(4.0), 0.0, 5.0]), (0,)),
                                      (_ffi.LMH_DRAG_GAMMA_P, np.array([media.DsT_from_2piT(4.0), 5.3, 5.0]), (1, 2))):
            ens = GridEnsemble(ctx, g, [1.5], [N], pos, mom, dt=0.01, t0=0.6, nsteps=steps,
                               drag_model=drag_model, collision=_ffi.LMH_LANGEVIN, seed=3,
                               checkpoint_every=K, **obs_kw)
            ens.forward(theta)
            grads = {name: ens.pullback(wl) for name, wl in w.items()}
            for k in ks:
                # central differences at h = 1e-4 theta_k and h/2, and their Richardson combination
                # (4 D(h/2) - D(h))/3, which cancels the O(h^2) term: <cos 2 phi> has a large
                # third derivative from quarks at small p_T, where the angle turns fastest
                fds = []
                for hr in (1e-4, 5e-5):
                    h = hr * theta[k]
                    tp, tm = theta.copy(), theta.copy()
                    tp[k] += h
                    tm[k] -= h
                    fds.append((ens.run_plain(tp) - ens.run_plain(tm)) / (2 * h))
                for name, wl in w.items():
                    d1, d2 = float(fds[0] @ wl), float(fds[1] @ wl)
                    fd = (4.0 * d2 - d1) / 3.0
                    adj = float(grads[name][k])
                    err = abs(adj - fd) / abs(fd)
                    print(f"  model {drag_model} theta[{k}] {name}: adjoint {adj:.9e} central difference "
                          f"h=1e-4 {d1:.9e} (rel {abs(adj - d1) / abs(d1):.1e}), Richardson {fd:.9e} "
                          f"(rel {err:.1e})", flush=True)
                    assert err <= 1e-4, (drag_model, k, name, adj, fd)
