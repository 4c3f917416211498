"""Batches of independent runs (lmh_batch_*): every run of a batch must reproduce the same run
alone (Context.run with Philox stream + r) bit for bit, for every launch shape, with and without
the captured graph; the single-run path is itself checked against the oracle elsewhere, and run 0
is checked against the oracle here as well."""
from __future__ import annotations

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import rel_dev, small_grid
from lmhip import _ffi, media, transport
from lmhip.batch import Batch

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
DT, T0 = 0.02, 0.6
STREAM0 = 40


def _runs(seed=7, counts=(1, 300, 0, 2500, 777), rta=True):
    rng = np.random.default_rng(seed)
    runs = []
    for k, n in enumerate(counts):
        ecc = 0.1 * k
        g = media.hydro_grid(nx=40 + 4 * k, ny=36, nt=10, L=15.0, tau0=0.6, tau1=12.0,
                             T0=0.3 + 0.05 * k, ecc=ecc)
        m = 1.5 if k % 2 == 0 else 4.75
        DsT = media.DsT_from_2piT(1.5 + 1.3 * k)
        pos = media.hydro_positions(n, rng, ecc=ecc) * 1.5
        mom = media.pt_rapidity_momenta(n, m, rng)
        Tmin = max(float(np.min(g.T)), 0.0)
        Tmax = max(float(np.max(g.T)), Tmin + EPS)
        drag = transport.build_tau_drag_spline(m, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024)
        rt = jt = None
        if rta:
            rt = transport.build_relaxation_spline(media.tau_R_law(0.16), Tmin=Tmin, Tmax=Tmax, nT=1024)
            jt = lmhip.build_juttner_invcdf(m, 3, Tmin=Tmin, Tmax=Tmax, nT=32, nU=256, np_=1024)
        runs.append(dict(g=g, m=m, DsT=DsT, pos=pos, mom=mom, drag=drag, rt=rt, jt=jt, n=n))
    return runs


def _params(collision, DsT=0.0, stream=STREAM0, **kw):
    return _ffi.make_params(dt=DT, t0=T0, DsT=DsT, seed=99, stream=stream, collision=collision,
                            noise=_ffi.LMH_NOISE_PHILOX, gamma0=5.3, Lambda=5.0, **kw)


def _load_batch(b, runs):
    desc = []
    for r in runs:
        med = b.add_medium(r["g"])
        d = dict(count=r["n"], mass=r["m"], DsT=r["DsT"], medium=med, drag_table=b.add_time_table(*r["drag"]))
        if r["rt"] is not None:
            d["rta_table"] = b.add_time_table(*r["rt"])
            d["juttner_table"] = b.add_juttner_table(*r["jt"])
        desc.append(d)
    pos = np.asfortranarray(np.hstack([r["pos"] for r in runs]))
    mom = np.asfortranarray(np.hstack([r["mom"] for r in runs]))
    b.set_runs(desc, pos, mom)


def _alone(runs, collision, steps, boost=True, slice_min=0, **kw):
    out = []
    for k, r in enumerate(runs):
        if r["n"] == 0:
            out.append((np.zeros((2, 0)), np.zeros((3, 0))))
            continue
        p = _params(collision, DsT=r["DsT"], stream=STREAM0 + k, **kw)
        g = r["g"]
        with lmhip.Context(0) as ctx:
            # the batch kernels interpolate each corner in tau (the unsliced form), except for
            # the runs of at least slice_min quarks, which read per-step slices as lmh_run's
            # one-step launches do
            ctx.set_tuning(slice=1 if 0 < slice_min <= r["n"] else 0)
            ctx.set_medium_grid2p1(g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, g.tgrid, g.T, g.vx, g.vy)
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, 0, *r["drag"])
            if r["rt"] is not None:
                ctx.set_time_table(_ffi.LMH_TABLE_RTA, 0, *r["rt"])
                ctx.set_juttner_table(0, *r["jt"])
            ctx.set_particles(r["pos"], r["mom"], [r["m"]])
            if boost:
                ctx.boost_to_lab(p, 0)
            ctx.run(p, 1, steps, None, 0)
            out.append(ctx.get_particles())
    return out


def _split(pos, mom, runs):
    res, b0 = [], 0
    for r in runs:
        res.append((pos[:, b0:b0 + r["n"]], mom[:, b0:b0 + r["n"]]))
        b0 += r["n"]
    return res


@pytest.mark.parametrize("collision", [_ffi.LMH_LANGEVIN, _ffi.LMH_LANGEVIN_RTA, _ffi.LMH_FLUID])
def test_batch_equals_runs_alone(gpu_ok, collision):
    runs = _runs()
    steps = 25
    ref = _alone(runs, collision, steps)
    with Batch(0) as b:
        _load_batch(b, runs)
        p = _params(collision)
        b.boost_to_lab(p, 0)
        b.run(p, 1, steps)
        got = _split(*b.get_particles(), runs)
    for k, ((gp, gm), (rp, rm)) in enumerate(zip(got, ref)):
        assert np.array_equal(gp, rp) and np.array_equal(gm, rm), f"run {k}"


def test_batch_run_against_oracle(gpu_ok):
    """run 3 of a batch (Langevin + RTA) against the oracle fed the Philox record of stream + 3"""
    runs = _runs(rta=True)
    steps = 20
    with Batch(0) as b:
        _load_batch(b, runs)
        p = _params(_ffi.LMH_LANGEVIN_RTA)
        b.run(p, 1, steps)
        got = _split(*b.get_particles(), runs)
    r = runs[3]
    p3 = _params(_ffi.LMH_LANGEVIN_RTA, DsT=r["DsT"], stream=STREAM0 + 3)
    noise = np.ascontiguousarray(O.noise_record(p3, 2, 3, r["n"], 1, steps))
    keep = O.Keep()
    gm = O.grid_medium(keep, r["g"])
    inv, nU, nT, Tt, idt = r["jt"]
    sp = O.species(keep, [r["m"]], [r["n"]], drag=[r["drag"]], rta=[r["rt"]], jt=[(inv, nU, nT, Tt, idt)])
    opos, omom = r["pos"].copy(order="F"), r["mom"].copy(order="F")
    rc, bad = O.run_grid(p3, gm, sp, opos, omom, 1, steps, noise)
    assert rc == 0, (rc, bad)
    assert np.all(rel_dev(got[3][1], omom, 1e-6) <= 1e-12)
    assert np.all(rel_dev(got[3][0], opos, 1e-6) <= 1e-12)


@pytest.mark.parametrize("spl,graph_steps,sort_every", [(1, 0, 0), (4, 0, 5), (3, 4, 0), (2, 3, 4)])
def test_batch_medium_slices_equal_runs_alone(gpu_ok, spl, graph_steps, sort_every):
    """slice_min 500: runs 3 (2500 quarks) and 4 (777) read per-step node slices (22 direct steps:
    slices written for 16 steps, then 6; or graph replays, with sorts that alternate the state
    between its two buffers and so between two captured graphs), runs 0 and 1 the corners; every
    run bit for bit the same run alone with the record slices, and the largest sliced run against
    the oracle at 1e-12"""
    runs = _runs(seed=17)
    steps = 22
    ref = _alone(runs, _ffi.LMH_LANGEVIN_RTA, steps, slice_min=500)
    with Batch(0) as b:
        b.set_tuning(block=128, steps_per_launch=spl, graph_steps=graph_steps, sort_every=sort_every,
                     slice_min=500)
        assert b.get_tuning()["slice_min"] == 500
        _load_batch(b, runs)
        p = _params(_ffi.LMH_LANGEVIN_RTA)
        b.boost_to_lab(p, 0)
        b.run(p, 1, steps)
        got = _split(*b.get_particles(), runs)
    for k, ((gp, gm), (rp, rm)) in enumerate(zip(got, ref)):
        assert np.array_equal(gp, rp) and np.array_equal(gm, rm), f"run {k}"
    r = runs[3]
    p3 = _params(_ffi.LMH_LANGEVIN_RTA, DsT=r["DsT"], stream=STREAM0 + 3)
    keep = O.Keep()
    gm_ = O.grid_medium(keep, r["g"])
    inv, nU, nT, Tt, idt = r["jt"]
    sp = O.species(keep, [r["m"]], [r["n"]], drag=[r["drag"]], rta=[r["rt"]], jt=[(inv, nU, nT, Tt, idt)])
    opos, omom = r["pos"].copy(order="F"), r["mom"].copy(order="F")
    assert O.boost_to_lab_grid(p3, gm_, sp, opos, omom, 0) == 0
    rc, bad = O.run_grid(p3, gm_, sp, opos, omom, 1, steps,
                         np.ascontiguousarray(O.noise_record(p3, 2, 3, r["n"], 1, steps)))
    assert rc == 0, (rc, bad)
    assert np.all(rel_dev(got[3][1], omom, 1e-6) <= 1e-12)
    assert np.all(rel_dev(got[3][0], opos, 1e-6) <= 1e-12)


@pytest.mark.parametrize("block,persistent,spl", [(64, 0, 1), (128, 1, 3), (256, 1, 1), (256, 0, 4)])
def test_batch_launch_shape_invariance(gpu_ok, block, persistent, spl):
    runs = _runs(seed=3)
    p = _params(_ffi.LMH_LANGEVIN_RTA)
    res = []
    for tune in ((256, 0, 1), (block, persistent, spl)):
        with Batch(0) as b:
            b.set_tuning(block=tune[0], persistent=tune[1], steps_per_launch=tune[2])
            _load_batch(b, runs)
            b.run(p, 1, 13)
            res.append(b.get_particles())
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_batch_graph_replays_match_direct_launches(gpu_ok):
    """a captured graph of 5 steps replayed 4 times, with downloads and other launches between
    the replays: every replay's state equals the directly launched steps"""
    runs = _runs(seed=11)
    p = _params(_ffi.LMH_LANGEVIN_RTA)
    direct = []
    with Batch(0) as b:
        _load_batch(b, runs)
        for c in range(4):
            b.run(p, 1 + 5 * c, 5)
            direct.append(b.get_particles())
    with Batch(0) as b:
        b.set_tuning(block=128, steps_per_launch=1, persistent=1, graph_steps=5)
        _load_batch(b, runs)
        for c in range(4):
            b.run(p, 1 + 5 * c, 5)
            got = b.get_particles()
            b.observables(nbins=4)
            assert np.array_equal(got[0], direct[c][0]) and np.array_equal(got[1], direct[c][1]), c
    # 12 steps with graph_steps 5: two replays and two direct steps, then 8: one replay and 3
    with Batch(0) as b:
        b.set_tuning(graph_steps=5)
        _load_batch(b, runs)
        b.run(p, 1, 12)
        b.run(p, 13, 8)
        got = b.get_particles()
    assert np.array_equal(got[0], direct[3][0]) and np.array_equal(got[1], direct[3][1])


@pytest.mark.parametrize("graph_steps", [0, 3])
def test_batch_cell_sort_identical(gpu_ok, graph_steps):
    """each run's quarks ordered by grid cell every 7 steps (rounded up to whole graph replays of
    3): particles and observables bit for bit the unsorted batch's, also between calls and after
    the caller's order is restored for a download in the middle"""
    runs = _runs(seed=13)
    p = _params(_ffi.LMH_LANGEVIN_RTA)
    with Batch(0) as b:
        _load_batch(b, runs)
        b.run(p, 1, 11)
        mid = b.get_particles()
        b.run(p, 12, 19)
        ref = b.get_particles()
        ref_obs = b.observables(nbins=12, pt_max=6.0, sigma=0.3)
    with Batch(0) as b:
        b.set_tuning(block=128, steps_per_launch=2, graph_steps=graph_steps, sort_every=7)
        assert b.get_tuning()["sort_every"] == 7
        _load_batch(b, runs)
        b.run(p, 1, 11)
        got_mid = b.get_particles()
        b.run(p, 12, 19)
        obs = b.observables(nbins=12, pt_max=6.0, sigma=0.3)
        got = b.get_particles()
    assert np.array_equal(got_mid[0], mid[0]) and np.array_equal(got_mid[1], mid[1])
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(obs, ref_obs)


def test_batch_autotune_keeps_state_and_results(gpu_ok):
    runs = _runs(seed=5)
    p = _params(_ffi.LMH_LANGEVIN_RTA)
    with Batch(0) as b:
        _load_batch(b, runs)
        ref_before = b.get_particles()
        b.run(p, 1, 10)
        ref = b.get_particles()
    with Batch(0) as b:
        _load_batch(b, runs)
        tune, ms = b.autotune(p, trial_steps=2)
        assert tune["block"] in (64, 128, 256) and tune["persistent"] in (0, 1)
        assert tune["steps_per_launch"] in (1, 2, 4) and ms > 0.0
        assert tune["sort_every"] in (0, 50, 100, 200)
        # no run reaches 2^14 quarks here: every threshold slices none
        assert tune["slice_min"] in (0, 1 << 14, 1 << 16, 1 << 18)
        assert b.get_tuning() == tune
        now = b.get_particles()
        assert np.array_equal(now[0], ref_before[0]) and np.array_equal(now[1], ref_before[1])
        assert b.nonfinite_step() == -1
        b.run(p, 1, 10)
        got = b.get_particles()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_batch_observables_per_run(gpu_ok):
    runs = _runs(seed=2)
    p = _params(_ffi.LMH_LANGEVIN)
    with Batch(0) as b:
        _load_batch(b, runs)
        b.run(p, 1, 5)
        pos, mom = b.get_particles()
        for sigma in (0.0, 0.4):
            obs = b.observables(nbins=20, pt_min=0.0, pt_max=8.0, sigma=sigma)
            b0 = 0
            for k, r in enumerate(runs):
                ref = O.observables(mom, b0, b0 + r["n"], 20, 0.0, 8.0, sigma)
                assert np.allclose(obs[k], ref, rtol=1e-12, atol=1e-9), k
                b0 += r["n"]


def test_batch_errors(gpu_ok):
    runs = _runs(seed=1, rta=False)
    with Batch(0) as b:
        with pytest.raises(lmhip.LMHError, match="no runs"):
            b.run(_params(_ffi.LMH_LANGEVIN), 1, 1)
        _load_batch(b, runs)
        with pytest.raises(lmhip.LMHError, match="RTA needs"):
            b.run(_params(_ffi.LMH_LANGEVIN_RTA), 1, 1)
        with pytest.raises(lmhip.LMHError, match="Philox"):
            b.run(_ffi.make_params(dt=DT, noise=_ffi.LMH_NOISE_FIXED), 1, 1)
        with pytest.raises(lmhip.LMHError):
            b.set_tuning(block=96)
