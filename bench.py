#!/usr/bin/env python3
"""Throughput of the fused FP64 Langevin step on one MI355X.

Workload (the headline configuration): a 2+1D hydro-like medium on a 200 x 200 grid over
[-15, 15]^2 fm with 120 tau slices from 0.6 to 12 fm/c (T Gaussian, sigma_x = 3, sigma_y = 4 fm,
Bjorken-cooled peak T0 = 0.45 GeV, radial flow v_r = tanh(0.08 r (tau - 0.6))), 10^7 quarks
(70% charm m = 1.5, 30% bottom m = 4.75 GeV) with positions ~ T^4 at tau0, p_T ~ p_T/(1+p_T^2/4)^3
on [0, 20] GeV and rapidity U[-1, 1]; D_s(2 pi T) = 4, RTA with tau_R = 5 (eta/s)/T, eta/s = 0.16,
dt = 0.01 fm/c, Philox noise.  One step = one pass of the fused step over all 10^7 quarks.

The state is resident in HBM before the timed region; the timed region is bracketed by device
synchronisations and timed with HIP events on the library's stream.  The quarks are re-sorted by
grid cell every --sort-every steps: ms_per_step carries that sort amortised over its cadence (the
time of a re-sort, measured, divided by the cadence, in place of the sorts the window happens to
hold), so the figure does not depend on where the timed window falls in the cadence.

Other workloads (--workload): bjorken (the Bjorken-expanding medium in Milne coordinates, 10^6
charm quarks, post-point kick with gamma(p, T)), batch (256 independent runs in one launch) and
adjoint (gradient / forward time of the pathwise adjoint).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "langevinmedium.jl_amd")
for _p in (PKG, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
BYTES_PER_PARTICLE_STEP = 80   # read + write x, y, px, py, pz (FP64)
# HBM bytes per launch of the step kernel on the default workload, from rocprofv3 FETCH_SIZE
# (x2.00) + WRITE_SIZE (x1.00), factors from a 1 GiB device copy: profiles/k_cart_grid_1e7_grid2p1_pmc.md
# (the 80 algorithmic bytes, the 4-byte identity read of the cell-sorted order and the colliding
# quarks' Juttner-table reads)
MEASURED_TRAFFIC_PER_QUARK = 89.17


def build_workload(N, seed=3):
    from lmhip import media
    return media.hydro_ensemble(N, seed)


def tables_for(g, masses, DsT, pdim):
    from lmhip import media
    return media.hydro_tables(g, masses, DsT, pdim)


COLLISIONS = {"langevin_rta": 3, "langevin": 0, "rta": 1, "fluid": 2}


def make_params(DsT, noise_fixed=False, collision="langevin_rta"):
    from lmhip import _ffi
    return _ffi.make_params(dt=0.01, t0=0.6, DsT=DsT, seed=3, stream=0,
                            collision=COLLISIONS[collision],
                            noise=_ffi.LMH_NOISE_FIXED if noise_fixed else _ffi.LMH_NOISE_PHILOX,
                            check_finite=False)


def cpu_baseline(g, masses, tabs, DsT, budget_s=12.0, n1=20000, nmt=200000):
    """The oracle (C restatement of the reference CPU kernels, one pass per kernel) on a bounded
    sample of the same workload: n quarks of the same mix, the Philox noise record it reads
    generated inside the timed loop, as the reference draws randn every step.  Timed on one
    thread (n1 quarks) and on all the host threads OpenMP is given (nmt quarks, the per-particle
    loops split between threads), half the budget each; `value` is the all-thread figure."""
    import ctypes
    import numpy as np
    import _oracle as O
    from lmhip import media

    L = O.lib()

    def timed(n, threads, budget):
        rng = np.random.default_rng(99)
        nc = int(round(0.7 * n))
        counts = [nc, n - nc]
        pos = media.hydro_positions(n, rng)
        mom = np.asfortranarray(np.hstack([media.pt_rapidity_momenta(counts[0], masses[0], rng),
                                           media.pt_rapidity_momenta(counts[1], masses[1], rng)]))
        keep = O.Keep()
        gm = O.grid_medium(keep, g)
        sp = O.species(keep, masses, counts, drag=[t[0] for t in tabs], rta=[t[1] for t in tabs],
                       jt=[(t[2][0], t[2][1], t[2][2], t[2][3], t[2][4]) for t in tabs])
        p = make_params(DsT, noise_fixed=True)
        C = L.lmo_noise_channels(ctypes.byref(p), 2, 3)
        rec = np.zeros((1, C, n))
        L.lmo_set_threads(threads)
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget:
            L.lmo_noise_record(ctypes.byref(p), 2, 3, n, 1 + steps, O.dp(rec[0]))
            rc, _ = O.run_grid(p, gm, sp, pos, mom, 1 + steps, 1, rec)
            assert rc == 0
            steps += 1
        return n * steps / (time.perf_counter() - t0), steps

    threads = max(1, int(L.lmo_get_threads()))
    v1, s1 = timed(n1, 1, budget_s / 2)
    vm, sm = timed(nmt, threads, budget_s / 2)
    L.lmo_set_threads(threads)
    return {"value": vm, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "one_core_value": v1,
            "sample": f"oracle (C restatement of kernels_cpu.jl, one pass per kernel) on the same workload: "
                      f"{nmt} quarks x {sm} steps on {threads} threads (OpenMP over particles) and "
                      f"{n1} quarks x {s1} steps on 1 thread, {budget_s / 2:.0f} s each"}


def bjorken_setup(N, seed=2):
    """the Bjorken workload's medium, initial state, parameters and step count"""
    import numpy as np
    from lmhip import _ffi, media

    rng = np.random.default_rng(seed)
    xg, tg, Tf, Vf, tau_end = media.bjorken_bath(T0=0.4, tau0=0.6, Tstop=0.155)
    mom = media.pt_rapidity_momenta(N, 1.5, rng, ymax=1.0, ptmax=20.0)
    pos = np.zeros((3, N), order="F")
    dt = 0.005
    steps = int(math.floor((tau_end - 0.6) / dt))
    DsT = media.DsT_from_2piT(4.0)
    gamma0 = media.fmGeV / (1.5 * DsT)          # gamma(0, T) = T^2/(m D_s T): D_s(2 pi T) = 4 at p = 0
    params = _ffi.make_params(dt=dt, t0=0.6, DsT=DsT, seed=2026, stream=1, coords=_ffi.LMH_MILNE,
                              collision=_ffi.LMH_LANGEVIN, integrator=_ffi.LMH_POSTPOINT,
                              drag_model=_ffi.LMH_DRAG_GAMMA_P, gamma0=gamma0, Lambda=5.0,
                              noise=_ffi.LMH_NOISE_PHILOX, check_finite=False)
    return (xg, tg, Tf, Vf), pos, mom, params, steps


def bench_bjorken(args):
    """The Bjorken workload in Milne coordinates: T(tau) = 0.4 (0.6/tau)^(1/3) GeV, uniform in r,
    the fluid at rest in the comoving frame; 10^6 charm quarks (m = 1.5 GeV) with
    p_T ~ p_T/(1+p_T^2/4)^3 on [0, 20] GeV and rapidity U[-1, 1] at eta_s = 0; post-point kick with
    gamma(p, T) = gamma0 T^2/(1 + p^2/Lambda^2), gamma0 = the D_s(2 pi T) = 4 drag at p = 0,
    Lambda = 5 GeV; dt = 0.005 fm/c from tau = 0.6 fm/c until T = 0.155 GeV (1942 steps); Philox
    noise.  The timed region is the whole run (the initial boost and all 1942 steps) after
    --warmup steps of the same workload; one step = one pass over all quarks.  Algorithmic bytes:
    96 per particle-step (x, y, eta_s, p_x, p_y, p_z read and written)."""
    import ctypes
    import numpy as np
    import torch
    import lmhip
    import _oracle as O
    from lmhip import _ffi

    N = args.bj_n
    (xg, tg, Tf, Vf), pos, mom, params, steps = bjorken_setup(N)
    ctx = lmhip.Context(0)
    ctx.set_medium_radial(xg, tg, Tf, Vf)
    ctx.set_particles(pos, mom, [1.5])
    ctx.boost_to_lab(params, 0)
    ctx.run(params, 1, args.warmup, None, 1)
    ctx.set_particles(pos, mom, [1.5])
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    ctx.boost_to_lab(params, 0)
    ctx.run(params, 1, steps, None, 1)
    ev1.record(stream)
    ev1.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1)
    obs = ctx.observables(40, 0.0, 10.0, 0.0, -1)
    gpos, gmom = ctx.get_particles()
    ctx.close()
    if not (np.all(np.isfinite(gpos)) and np.all(np.isfinite(gmom))):
        raise RuntimeError("non-finite state")
    bytes_per_launch = 96 * N
    t_launch = ms / 1e3 / (steps + 1)
    # CPU baseline: the oracle's Milne run on a bounded sample of the same workload
    L = O.lib()

    def timed(n, threads, budget):
        (xg_, tg_, Tf_, Vf_), p_, m_, prm, _ = bjorken_setup(n, seed=99)
        keep = O.Keep()
        om = O.radial_medium(keep, xg_, tg_, Tf_, Vf_)
        sp = O.species(keep, [1.5], [n])
        prm.noise = _ffi.LMH_NOISE_FIXED
        C = L.lmo_noise_channels(ctypes.byref(prm), 3, 3)
        rec = np.zeros((1, C, n))
        L.lmo_set_threads(threads)
        k = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget:
            L.lmo_noise_record(ctypes.byref(prm), 3, 3, n, 1 + k, O.dp(rec[0]))
            rc, _ = O.run_cartesian(prm, om, sp, p_, m_, 1 + k, 1, rec)
            assert rc == 0
            k += 1
        return n * k / (time.perf_counter() - t0), k

    cpu = None
    if not args.no_cpu:
        threads = max(1, int(L.lmo_get_threads()))
        v1, s1 = timed(20000, 1, args.cpu_budget / 2)
        vm, sm = timed(200000, threads, args.cpu_budget / 2)
        L.lmo_set_threads(threads)
        cpu = {"value": vm, "unit": "particle-steps/s", "cores": threads, "kind": "port", "one_core_value": v1,
               "sample": f"oracle (C restatement of kernels_cpu.jl + the Milne eta_s push) on the same workload: "
                         f"200000 quarks x {sm} steps on {threads} threads and 20000 quarks x {s1} steps on 1 "
                         f"thread, {args.cpu_budget / 2:.0f} s each"}
    out = {
        "metric": "particle-steps/s, Bjorken medium in Milne coordinates, 1e6 charm",
        "value": round(N * steps / (ms / 1e3), 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(ms / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Bjorken T=0.4(0.6/tau)^(1/3) GeV to T=0.155 ({steps} steps of dt=0.005 fm/c), "
                               f"Milne (x, y, eta_s); {N} charm m=1.5; post-point kick, "
                               f"gamma(p,T)=gamma0 T^2/(1+p^2/25), gamma0={params.gamma0:.4f}; Philox; FP64",
                   "quarks": N, "wall_s_timed": round(wall, 4), "mean_pT": round(obs[1] / obs[0], 6),
                   "mean_abs_eta_s": round(float(np.mean(np.abs(gpos[2]))), 6)},
        "roofline": {"bound": "hbm", "achieved": round(bytes_per_launch / t_launch / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(bytes_per_launch / t_launch / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_milne (post-point, GAMMA_P, Philox)",
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(t_launch * 1e3, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """256 independent runs in one launch: each run draws D_s(2 pi T) ~ U[1.5, 7], T0 ~ U[0.3, 0.5]
    GeV, a charm count 10^U[3, 6] and a profile eccentricity ~ U[0, 0.4] (seed 4), on its own
    200 x 200 x 120 grid; Langevin + RTA as in the headline workload.  The drag tables are per run,
    the RTA and Juttner tables are shared over the common temperature range.  The launch shape is
    chosen by the library's autotuner (block, steps per launch, persistent launch and the cadence of
    the runs' device cell sorts), then `--steps` steps run as replays of a captured graph."""
    import numpy as np
    import torch
    from lmhip import _ffi, media, transport
    from lmhip.batch import Batch

    rng = np.random.default_rng(4)
    R = args.runs
    Ds = rng.uniform(1.5, 7.0, R)
    T0s = rng.uniform(0.3, 0.5, R)
    counts = np.round(10.0 ** rng.uniform(3.0, 6.0, R)).astype(np.int64)
    eccs = rng.uniform(0.0, 0.4, R)
    m = 1.5
    t_setup = time.perf_counter()
    b = Batch(0)
    Tmin, Tmax = None, None
    meds = []
    for r in range(R):
        g = media.hydro_grid(nx=200, ny=200, nt=120, L=15.0, tau0=0.6, tau1=12.0, T0=float(T0s[r]),
                             ecc=float(eccs[r]))
        lo, hi = max(float(np.min(g.T)), 0.0), float(np.max(g.T))
        Tmin = lo if Tmin is None else min(Tmin, lo)
        Tmax = hi if Tmax is None else max(Tmax, hi)
        meds.append(b.add_medium(g))
        del g
    rt = b.add_time_table(*transport.build_relaxation_spline(media.tau_R_law(0.16), Tmin=Tmin, Tmax=Tmax, nT=1024))
    jt = b.add_juttner_table(*transport.build_juttner_invcdf(m, 3, Tmin=Tmin, Tmax=Tmax, nT=256, nU=2048, np_=8192))
    desc, pos, mom = [], [], []
    for r in range(R):
        DsT = media.DsT_from_2piT(float(Ds[r]))
        drag = b.add_time_table(*transport.build_tau_drag_spline(m, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024))
        desc.append(dict(count=int(counts[r]), mass=m, DsT=DsT, medium=meds[r], drag_table=drag,
                         rta_table=rt, juttner_table=jt))
        pos.append(media.hydro_positions(int(counts[r]), rng, ecc=float(eccs[r])))
        mom.append(media.pt_rapidity_momenta(int(counts[r]), m, rng))
    b.set_runs(desc, np.hstack(pos), np.hstack(mom))
    del pos, mom
    N = int(counts.sum())
    params = make_params(0.0, False, "langevin_rta")
    b.boost_to_lab(params, 0)
    tune, tuned_ms = b.autotune(params, trial_steps=3)
    if args.slice_min >= 0:
        tune["slice_min"] = args.slice_min
    b.set_tuning(block=tune["block"], steps_per_launch=tune["steps_per_launch"],
                 persistent=tune["persistent"], graph_steps=10, sort_every=tune["sort_every"],
                 slice_min=tune["slice_min"])
    setup = time.perf_counter() - t_setup
    print(f"batch setup {setup:.1f} s; {R} runs, {N} quarks, tuning {tune} ({tuned_ms:.3f} ms/step)",
          file=sys.stderr, flush=True)
    steps = max(10, (args.steps // 10) * 10)
    b.run(params, 1, args.warmup)
    b.nonfinite_step()
    stream = torch.cuda.ExternalStream(b.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    b.run(params, 1 + args.warmup, steps)
    ev1.record(stream)
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1)
    bad = b.nonfinite_step()
    if bad >= 0:
        raise RuntimeError(f"non-finite state at step {bad}")
    obs = b.observables(nbins=40, pt_min=0.0, pt_max=10.0)
    b.close()
    out = {
        "metric": "particle-steps/s, 256 batched runs (one launch per step, graph-captured)",
        "value": round(N * steps / (ms / 1e3), 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(ms / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{R} runs: D_s(2piT)~U[1.5,7], T0~U[0.3,0.5] GeV, N~10^U[3,6] charm, "
                               "ecc~U[0,0.4], each on its own 200x200x120 grid; Langevin + RTA; "
                               "dt=0.01; Philox; FP64",
                   "runs": R, "quarks": N, "tuning": tune, "autotune_ms_per_step": round(tuned_ms, 4),
                   "graph_steps": 10, "setup_s": round(setup, 1),
                   "mean_pT_run0": round(float(obs[0, 1] / max(obs[0, 0], 1.0)), 6)},
        "roofline": {"bound": "hbm", "achieved": round(BYTES_PER_PARTICLE_STEP * N / (ms / steps / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(BYTES_PER_PARTICLE_STEP * N / (ms / steps / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_batch_grid (whole step, all runs)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """The pathwise adjoint on the headline medium: 10^6 charm quarks x 1000 steps, checkpoints
    every 50 steps, Langevin with D_s(2 pi T) = 4 (drag table; d/dDsT) and with the GAMMA_P drag
    gamma0 = 5.3, Lambda = 5 GeV (d/dgamma0, d/dLambda).  Losses: <p_T>, a Gaussian-smoothed
    (sigma = 0.25 GeV) 40-bin p_T spectrum on [0, 10] GeV weighted by fixed random weights, and
    <cos 2 phi>.  Times: the plain forward run (lmh_run), the checkpointed forward (after one
    untimed one that allocates the checkpoints), one pullback per loss (after one untimed pullback
    that allocates the backward workspace), and the same with the
    whole tape kept in device memory (keep_tape: 64 GB at this size; the pullback runs the reverse
    sweeps only, with bit-identical gradients); gradients checked against central differences of
    the plain run at relative steps 1e-4, 1e-5 and 1e-6."""
    import numpy as np
    import torch
    import lmhip
    from lmhip import _ffi, media
    from lmhip.adjoint import GridEnsemble

    N, steps, K = args.adj_n, args.adj_steps, args.adj_k
    rng = np.random.default_rng(5)
    g = media.hydro_grid(nx=200, ny=200, nt=120, L=15.0, tau0=0.6, tau1=12.0, T0=0.45)
    pos = media.hydro_positions(N, rng)
    mom = media.pt_rapidity_momenta(N, 1.5, rng)
    obs_kw = dict(nbins=40, pt_min=0.0, pt_max=10.0, sigma=0.25)
    w = {}
    w["mean_pT"] = np.zeros(45); w["mean_pT"][1] = 1.0 / N
    w["cos2phi"] = np.zeros(45); w["cos2phi"][3] = 1.0 / N
    w["spectrum"] = np.zeros(45); w["spectrum"][5:] = np.random.default_rng(6).normal(size=40) / N
    ctx = lmhip.Context(0)
    ctx.set_tuning(block=256, steps_per_launch=1, slice=1, defer=1)
    stream = torch.cuda.ExternalStream(ctx.stream)

    def timed(fn):
        ctx.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        r = fn()
        e1.record(stream)
        e1.synchronize()
        ctx.synchronize()
        return r, e0.elapsed_time(e1), (time.perf_counter() - t0) * 1e3

    results = {}
    for name, drag_model, theta, ks in (
            ("table", _ffi.LMH_DRAG_TABLE, np.array([media.DsT_from_2piT(4.0), 0.0, 5.0]), (0,)),
            ("gamma_p", _ffi.LMH_DRAG_GAMMA_P, np.array([media.DsT_from_2piT(4.0), 5.3, 5.0]), (1, 2))):
        ens = GridEnsemble(ctx, g, [1.5], [N], pos, mom, dt=0.01, t0=0.6, nsteps=steps,
                           drag_model=drag_model, collision=_ffi.LMH_LANGEVIN, seed=3,
                           checkpoint_every=K, **obs_kw)
        ens.reset(theta)
        ctx.run(ens.params(theta), 1, 20, None, 0)                  # warm-up (code objects)
        p = ens.reset(theta)
        _, t_fwd, _ = timed(lambda: ctx.run(p, 1, steps, None, 0))
        obs_plain = ctx.observables(**obs_kw)
        p = ens.reset(theta)
        ens.run_forward(p, theta, keep_tape=False)                  # checkpoint buffers, untimed
        p = ens.reset(theta)
        _, t_cfwd, _ = timed(lambda: ens.run_forward(p, theta, keep_tape=False))
        obs_ck = ctx.observables(**obs_kw)
        ens.pullback(w["mean_pT"])                                  # workspace allocation, untimed
        grads, t_bwd = {}, {}
        for lname, wl in w.items():
            gr, t_b, _ = timed(lambda: ens.pullback(wl))
            grads[lname], t_bwd[lname] = gr, t_b
        # the whole tape kept: forward pass writes it, the pullbacks only reverse
        p = ens.reset(theta)
        _, t_kfwd, _ = timed(lambda: ens.run_forward(p, theta, keep_tape=True))
        obs_keep = ctx.observables(**obs_kw)
        ens.pullback(w["mean_pT"])                                  # untimed, as above
        t_kbwd, same_grads = {}, True
        for lname, wl in w.items():
            gr, t_b, _ = timed(lambda: ens.pullback(wl))
            t_kbwd[lname] = t_b
            same_grads = same_grads and bool(np.array_equal(gr, grads[lname]))
        # central differences of the plain run at relative steps 1e-4 (the stated check), 1e-5
        # and 1e-6: trajectories that cross a grid cell or a clamp between the two runs make the
        # difference quotient differ from the pathwise derivative by O(h) per crossing
        fd = {}
        for k in ks:
            for hr in (1e-4, 1e-5, 1e-6):
                h = hr * theta[k]
                tp, tm = theta.copy(), theta.copy()
                tp[k] += h
                tm[k] -= h
                op, om = ens.run_plain(tp), ens.run_plain(tm)
                for lname, wl in w.items():
                    fd[(lname, k, hr)] = float((op - om) @ wl / (2 * h))
        checks = {}
        for lname in w:
            for k in ks:
                a = float(grads[lname][k])
                checks[f"{lname}/d{['DsT', 'gamma0', 'Lambda'][k]}"] = {
                    "adjoint": a,
                    **{f"fd_h{hr:g}": fd[(lname, k, hr)] for hr in (1e-4, 1e-5, 1e-6)},
                    **{f"rel_err_h{hr:g}": abs(a - fd[(lname, k, hr)]) / max(abs(fd[(lname, k, hr)]), 1e-300)
                       for hr in (1e-4, 1e-5, 1e-6)}}
        results[name] = {"forward_ms": round(t_fwd, 2), "checkpointed_forward_ms": round(t_cfwd, 2),
                         "backward_ms": {k: round(v, 2) for k, v in t_bwd.items()},
                         "ratio_one_loss": round((t_cfwd + t_bwd["mean_pT"]) / t_fwd, 3),
                         "ratio_three_losses": round((t_cfwd + sum(t_bwd.values())) / t_fwd, 3),
                         "same_final_state": bool(np.array_equal(obs_plain, obs_ck)),
                         "keep_tape": {"forward_ms": round(t_kfwd, 2),
                                       "backward_ms": {k: round(v, 2) for k, v in t_kbwd.items()},
                                       "ratio_one_loss": round((t_kfwd + t_kbwd["mean_pT"]) / t_fwd, 3),
                                       "ratio_three_losses": round((t_kfwd + sum(t_kbwd.values())) / t_fwd, 3),
                                       "same_final_state": bool(np.array_equal(obs_plain, obs_keep)),
                                       "same_gradients": same_grads},
                         "checks": checks}
        print(f"{name}: {json.dumps(results[name])}", file=sys.stderr, flush=True)
    ctx.close()
    # the headline: checkpoints every K steps with the segments recomputed (the larger ratio of the
    # two drag models); the whole tape kept in HBM (64 GB here) beside it
    ratio_keep = max(r["keep_tape"]["ratio_one_loss"] for r in results.values())
    ratio_ck = max(r["ratio_one_loss"] for r in results.values())
    out = {"metric": "adjoint/forward time ratio (gradient of one loss), 1e6 quarks x 1000 steps",
           "value": ratio_ck, "unit": "x", "n_gpus": 1, "steps": steps, "warmup": 20,
           "ms_per_step": round(results["gamma_p"]["forward_ms"] / steps, 5), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"pathwise adjoint, 2+1D grid 200x200x120, {N} charm x {steps} steps, "
                                  f"checkpoints every {K} steps, segments recomputed; Langevin; Philox; FP64",
                      "ratio_per_drag_model": {k: r["ratio_one_loss"] for k, r in results.items()},
                      "keep_tape_ratio": ratio_keep, "results": results},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def bench_observables(args):
    """The on-device observables of 10^7 quarks: N, sum p_T, sum p_T^2, sum cos 2phi,
    sum p_T cos 2phi and a 40-bin p_T histogram on [0, 10] GeV (lmh_observables_noalloc, the result
    in a device buffer), on the headline workload's initial state (70 % charm, 30 % bottom, the
    p_T spectrum of media.pt_rapidity_momenta, rapidity in [-1, 1]); `--steps` calls timed with HIP
    events on the library's stream after `--warmup` untimed ones.  Algorithmic bytes: px and py,
    16 B per quark."""
    import ctypes
    import numpy as np
    import torch
    import lmhip
    from lmhip import _ffi, media

    N = args.n
    g, masses, counts, pos, mom = media.hydro_ensemble(N)
    ctx = lmhip.Context(0)
    ctx.set_particles(pos, mom, masses, counts)
    del pos, mom
    out = torch.zeros(45, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    spec = _ffi.lmh_obs_spec(40, -1, 0.0, 10.0, 0.0)
    L = _ffi.lib()

    def call():
        rc = L.lmh_observables_noalloc(ctx._h, ctypes.byref(spec), ctypes.c_void_p(out.data_ptr()))
        if rc != 0:
            raise RuntimeError(L.lmh_last_error(ctx._h))

    for _ in range(max(1, args.warmup // 10)):
        call()
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        call()
    ev1.record(stream)
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    res = out.cpu().numpy()
    ctx.close()
    if res[0] != N:
        raise RuntimeError(f"observables counted {res[0]} quarks, not {N}")
    achieved = 16.0 * N / (ms / 1e3) / 1e9
    print(json.dumps({
        "metric": "on-device observables of 1e7 quarks (<p_T>, <cos 2 phi>, 40-bin spectrum) per call",
        "value": round(ms, 5), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": max(1, args.warmup // 10),
        "ms_per_step": round(ms, 5), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{N} quarks of the headline's initial state, 40 bins on [0, 10] GeV, "
                               "lmh_observables_noalloc (partials and final sums, result on the device)",
                   "mean_pT": round(res[1] / res[0], 6)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_obs_partial + k_obs_final", "bytes_per_launch": 16 * N},
        "cpu_baseline": None}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="grid", choices=["grid", "bjorken", "batch", "adjoint", "observables"],
                    help="grid: the headline 10^7-quark step (default); bjorken: the Bjorken medium in "
                         "Milne coordinates, 10^6 charm, 1942 steps; batch: 256 independent runs "
                         "in one launch; adjoint: gradient / forward time of the pathwise adjoint; "
                         "observables: the on-device observables of 10^7 quarks")
    ap.add_argument("--bj-n", type=int, default=1_000_000, help="bjorken: quarks")
    ap.add_argument("--runs", type=int, default=256, help="batch: number of runs")
    ap.add_argument("--slice-min", type=int, default=-1,
                    help="batch: runs of at least this many quarks read per-step medium slices "
                         "(-1: the autotuner's choice)")
    ap.add_argument("--adj-n", type=int, default=1_000_000, help="adjoint: quarks")
    ap.add_argument("--adj-steps", type=int, default=1000, help="adjoint: steps")
    ap.add_argument("--adj-k", type=int, default=50, help="adjoint: checkpoint cadence")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100,
                    help="untimed steps first: the GPU's power controller boosts the clock of an idle "
                         "GPU, overshoots and settles within ~15 ms (~60 headline steps) of load")
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--spl", type=int, default=1, help="steps per kernel launch")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--slice", type=int, default=1)
    ap.add_argument("--defer", type=int, default=0)
    ap.add_argument("--sort-every", type=int, default=64,
                    help="steps between device sorts of the quarks by grid cell (0: never)")
    ap.add_argument("--collision", default="langevin_rta", choices=sorted(COLLISIONS),
                    help="diagnostic: another momentum update than the workload's")
    ap.add_argument("--noise", default="philox", choices=["philox", "fixed"],
                    help="diagnostic: read a device-resident fixed-noise record instead of Philox")
    ap.add_argument("--host-sort", action="store_true",
                    help="diagnostic: order each species' quarks by grid cell before the upload")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the step kernel from rocprofv3 counters "
                         "(default: the measured figure when the workload is the default one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(particles shard with no per-step exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "bjorken":
        return bench_bjorken(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "observables":
        return bench_observables(args)

    import numpy as np
    import torch
    import lmhip
    from lmhip import _ffi, media

    N = args.n
    DsT = media.DsT_from_2piT(4.0)
    t_setup = time.perf_counter()
    g, masses, counts, pos, mom = build_workload(N)
    if args.host_sort:
        ix = np.clip(np.floor((pos[0] - g.x0) / g.dx), 0, g.nx - 2).astype(np.int64)
        iy = np.clip(np.floor((pos[1] - g.y0) / g.dy), 0, g.ny - 2).astype(np.int64)
        key = iy * g.nx + ix
        order = np.concatenate([c0 + np.argsort(key[c0:c0 + cn], kind="stable")
                                for c0, cn in zip(np.cumsum([0] + counts[:-1]), counts)])
        pos = np.asfortranarray(pos[:, order])
        mom = np.asfortranarray(mom[:, order])
    tabs = tables_for(g, masses, DsT, 3)
    params = make_params(DsT, args.noise == "fixed", args.collision)
    ctx = lmhip.Context(0)
    ctx.set_tuning(block=args.block, steps_per_launch=max(1, args.spl), slice=args.slice, defer=args.defer,
                   sort_every=args.sort_every)
    ctx.set_medium_grid2p1(g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, g.tgrid, g.T, g.vx, g.vy)
    for k, (drag, rta, jt) in enumerate(tabs):
        ctx.set_time_table(_ffi.LMH_TABLE_DRAG, k, *drag)
        ctx.set_time_table(_ffi.LMH_TABLE_RTA, k, *rta)
        ctx.set_juttner_table(k, jt[0], jt[1], jt[2], jt[3], jt[4])
    ctx.set_particles(pos, mom, masses, counts)
    ctx.boost_to_lab(params, 0)
    ctx.synchronize()
    print(f"setup {time.perf_counter() - t_setup:.1f} s; N={N} medium {g.nbytes() / 1e6:.0f} MB", file=sys.stderr, flush=True)

    record = None
    if args.noise == "fixed":
        # one record for every step run below, (steps, C, N) on the device
        total = args.warmup + args.steps
        C = ctx.noise_channels(params)
        record = torch.randn((total, C, N), dtype=torch.float64, device="cuda")
        orta = 3 if args.collision == "langevin_rta" else 0
        if args.collision in ("langevin_rta", "rta"):
            record[:, orta:orta + 2, :].uniform_(0.0, 1.0)
        torch.cuda.synchronize()

    def run(step0, n):
        if record is None:
            ctx.run(params, step0, n, None, args.spl)
        else:
            ctx.run_device_noise(params, step0, n, record[step0 - 1].data_ptr(), args.spl)

    run(1, args.warmup)
    ctx.synchronize()
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)
    print(f"warm-up done; timing {args.steps} steps", file=sys.stderr, flush=True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    run(1 + args.warmup, args.steps)
    ev1.record(stream)
    ev1.synchronize()
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1)
    timed_obs = ctx.observables(40, 0.0, 10.0, 0.0, -1)
    # the same steps again from the same initial state, every launch bracketed by events on the
    # library's stream: the device time of each kernel class over exactly the timed steps (the
    # step kernel's cost changes as the medium cools and expands)
    ctx.set_particles(pos, mom, masses, counts)
    del pos, mom
    ctx.boost_to_lab(params, 0)
    run(1, args.warmup)
    ctx.synchronize()
    ctx.profile_start()
    run(1 + args.warmup, args.steps)
    prof = ctx.profile_stop()
    if not np.array_equal(ctx.observables(40, 0.0, 10.0, 0.0, -1), timed_obs):
        raise RuntimeError("the profiled pass did not reproduce the timed pass")
    bad = ctx.nonfinite_step()
    if bad >= 0:
        raise RuntimeError(f"non-finite state at step {bad}")
    obs = ctx.observables(40, 0.0, 10.0, 0.0, -1)
    # the re-sort of the cadence's steady state: the mean of the window's own sorts (each a sort
    # of the order the previous one left, sort_every steps earlier); a cadence longer than the
    # window is timed on two more steps with a sort before each, the second sorting the order the
    # first left (observables() has restored the caller's order, so the first is a full sort)
    sort_ms = 0.0
    if args.sort_every > 0 and prof["sort"]["launches"] > 0:
        sort_ms = prof["sort"]["ms"] / prof["sort"]["launches"]
    elif args.sort_every > 0:
        ctx.set_tuning(block=args.block, steps_per_launch=max(1, args.spl), slice=args.slice,
                       defer=args.defer, sort_every=1)
        run(1 + args.warmup + args.steps, 1)
        ctx.profile_start()
        run(2 + args.warmup + args.steps, 1)
        p_sort = ctx.profile_stop()
        sort_ms = p_sort["sort"]["ms"] / max(1, p_sort["sort"]["launches"])
    ctx.close()

    # the window's own sorts (if any) replaced by the re-sort amortised over the cadence
    window_sort_ms = prof["sort"]["ms"]
    amortised_sort_ms = sort_ms * args.steps / args.sort_every if args.sort_every > 0 else 0.0
    ms_total = ms - window_sort_ms + amortised_sort_ms
    value = N * args.steps / (ms_total / 1e3)
    # the step kernel's mean launch, from HIP events on the library's stream over the timed steps:
    # the unbracketed window less the (bracketed) other kernels in it, and the mean of the same
    # launches each bracketed by events in the second pass.  Both include some dispatch time
    # beside the kernel's own (the first the gaps between launches, the second the events'), so
    # both bound the kernel's duration from above: the smaller is the figure used
    kernel_launches = prof["step"]["launches"]
    other_ms = sum(v["ms"] for k, v in prof.items() if k != "step")
    t_window = (ms - other_ms) / 1e3 / kernel_launches
    t_bracketed = prof["step"]["ms"] / 1e3 / kernel_launches
    t_launch = min(t_window, t_bracketed)
    bytes_per_launch = BYTES_PER_PARTICLE_STEP * N
    achieved = bytes_per_launch / t_launch / 1e9
    default_shape = (args.spl <= 1 and args.collision == "langevin_rta" and args.noise == "philox" and args.sort_every == 64
                     and args.defer == 0 and args.slice == 1 and not args.host_sort)
    traffic = args.traffic
    if traffic is None and default_shape:
        traffic = round(MEASURED_TRAFFIC_PER_QUARK * N)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_cart_grid (PDIM=3, Philox, Langevin+RTA)", "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(t_launch * 1e3, 4),
            "avg_launch_ms_window": round(t_window * 1e3, 4),
            "avg_launch_ms_event_bracketed": round(t_bracketed * 1e3, 4),
            "steps_per_launch": max(1, args.spl),
            "other_kernels_ms_per_step": {**{k: round(v["ms"] / args.steps, 4) for k, v in prof.items()
                                             if k not in ("step", "sort")},
                                          "sort": round(sort_ms / args.sort_every, 4) if args.sort_every > 0 else 0.0},
            "sort_ms_per_sort": round(sort_ms, 4),
            "sort_note": "a sort: its keys and radix sort, and the step launch that applies the "
                         "permutation by reading through it (that step's work included); the "
                         "step kernel's launches are the other steps"}
    out = {
        "metric": "particle-steps/s at 1e7 quarks, 2+1D medium",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_total / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2+1D hydro grid 200x200x120 over [-15,15]^2 fm, tau 0.6-12 fm/c; "
                               f"{N} quarks (70% charm m=1.5, 30% bottom m=4.75); Langevin D_s(2piT)=4 + "
                               "RTA tau_R=5(eta/s)/T; dt=0.01 fm/c; Philox noise; FP64",
                   "quarks": N, "grid": [200, 200, 120], "block": args.block,
                   "steps_per_launch": max(1, args.spl), "slice": args.slice, "defer": args.defer,
                   "sort_every": args.sort_every,
                   "collision": args.collision, "noise": args.noise, "host_sort": args.host_sort,
                   "wall_s_timed": round(wall, 4), "mean_pT": round(obs[1] / obs[0], 6),
                   "ms_per_step_window": round(ms / args.steps, 5),
                   "sorts_in_window": prof["sort"]["launches"],
                   "ms_per_step_note": "the window's step, slice and sort launches, with its sorts "
                                       "replaced by one re-sort's time x steps / sort_every"},
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, masses, tabs, DsT, budget_s=args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
