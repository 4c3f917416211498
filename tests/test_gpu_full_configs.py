"""The benchmarked workloads at their full shapes through the C ABI, against the oracle.

* The 2+1D hydro workload exactly as timed (media.hydro_ensemble / media.hydro_tables: 10^7
  quarks, 70% charm and 30% bottom, on the 200 x 200 x 120 grid; Langevin kicks with RTA
  collisions of tau_R = 5 (eta/s)/T; Philox noise; dt = 0.01; the specialised step kernel and
  the launch shape the benchmark uses), the driver's initial boost included, 40 steps with the
  quarks re-sorted by cell at steps 1, 17 and 33: every quark against the oracle at 1e-12.
* 256 batched runs on grids of that shape (D_s(2 pi T), T0, eccentricity and a log-uniform
  particle count drawn per run, the counts cut to 10^3..10^4 so the check runs in seconds):
  every run bit for bit the same run alone, and the largest run against the oracle at 1e-12.
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
        b.set_tuning(block=tune["block"], steps_per_launch=tune["steps_per_launch"],
                     persistent=tune["persistent"], graph_steps=10)
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
