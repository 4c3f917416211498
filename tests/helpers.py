"""Shared set-up of the parity tests: media, tables and paired GPU / oracle runs."""
from __future__ import annotations

import numpy as np

import _oracle as O
import lmhip
from lmhip import _ffi, media

EPS = np.finfo(float).eps


def rel_dev(a, b, floor):
    """per-particle ||a - b|| / max(||b||, floor) over the rows of (rows, N) matrices"""
    return np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), floor)


def flowing_medium():
    """a radial medium with non-uniform r nodes, inhomogeneous T and radial flow"""
    xg = np.sqrt(np.linspace(0.0, 400.0, 81))
    tg = np.arange(0.0, 10.01, 0.5)
    R, Tt = np.meshgrid(xg, tg, indexing="ij")
    Tf = np.asfortranarray(0.4 * np.exp(-R * R / 50.0) * (1 - 0.03 * Tt) + 0.05)
    Vf = np.asfortranarray(0.6 * np.tanh(R / 5.0) * (1 + 0.02 * Tt))
    V2 = np.asfortranarray(0.1 * R / (1 + R) + 0 * Tt)
    return xg, tg, Tf, Vf, V2


def tables(masses, Tf, DsT, pdim, rta=False, rta_time=None):
    Tmin = max(float(np.min(Tf)), 0.0)
    Tmax = max(float(np.max(Tf)), Tmin + EPS)
    drag, rt, jt = [], [], []
    for m in masses:
        drag.append(lmhip.build_tau_drag_spline(m, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024))
        if rta:
            if rta_time is None:
                rt.append(lmhip.build_taun_current_spline(m, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024))
            else:
                from lmhip.transport import build_relaxation_spline
                rt.append(build_relaxation_spline(rta_time, Tmin=Tmin, Tmax=Tmax, nT=1024))
            inv, nU, nT, Tt, idt = lmhip.build_juttner_invcdf(m, pdim, Tmin=Tmin, Tmax=Tmax, nT=64, nU=512, np_=2048)
            jt.append((inv, nU, nT, Tt, idt))
    return drag, (rt if rta else None), (jt if rta else None)


def load_tables(ctx, drag, rt, jt):
    for k, d in enumerate(drag):
        ctx.set_time_table(_ffi.LMH_TABLE_DRAG, k, *d)
        if rt is not None:
            ctx.set_time_table(_ffi.LMH_TABLE_RTA, k, *rt[k])
            inv, nU, nT, Tt, idt = jt[k]
            ctx.set_juttner_table(k, inv, nU, nT, Tt, idt)


def rta_noise_fixup(params, noise, dims, pdim):
    """uniform channels of a fixed record must lie in [0, 1)"""
    C = noise.shape[1]
    orta = -1
    c = 0
    if params.collision in (_ffi.LMH_LANGEVIN, _ffi.LMH_LANGEVIN_RTA):
        c += pdim
    if params.collision in (_ffi.LMH_RTA, _ffi.LMH_LANGEVIN_RTA):
        orta = c
    if orta >= 0:
        noise[:, orta:orta + 2, :] = np.abs(np.tanh(noise[:, orta:orta + 2, :]))
        noise[:, orta, :] *= 0.5          # a collision probability well above dt/tau
    assert C >= 0
    return noise


def run_radial_pair(params, med, masses, counts, pos, mom, steps, seed=0, step_begin=1, spl=0,
                    rta=False, rta_time=None, noise=None, initial_boost=False):
    xg, tg, Tf, Vf, V2 = med
    dims, pdim, N = pos.shape[0], mom.shape[0], pos.shape[1]
    drag, rt, jt = tables(masses, Tf, params.DsT if params.DsT > 0 else 0.2, pdim, rta, rta_time)
    with lmhip.Context(0) as ctx:
        ctx.set_medium_radial(xg, tg, Tf, Vf, V2)
        load_tables(ctx, drag, rt, jt)
        ctx.set_particles(pos, mom, masses, counts)
        if params.noise == _ffi.LMH_NOISE_FIXED and noise is None:
            rng = np.random.default_rng(seed)
            noise = rta_noise_fixup(params, rng.standard_normal((steps, ctx.noise_channels(params), N)), dims, pdim)
        if initial_boost:
            ctx.boost_to_lab(params, 0)
        ctx.run(params, step_begin, steps, noise, spl)
        gpos, gmom = ctx.get_particles()
    if params.noise == _ffi.LMH_NOISE_PHILOX:
        noise = O.noise_record(params, dims, pdim, N, step_begin, steps)
    keep = O.Keep()
    om = O.radial_medium(keep, xg, tg, Tf, Vf, V2)
    sp = O.species(keep, masses, counts, drag=drag, rta=rt, jt=jt)
    opos, omom = pos.copy(order="F"), mom.copy(order="F")
    if initial_boost:
        assert O.boost_to_lab_all(params, om, sp, opos, omom, 0) == 0
    if params.coords == _ffi.LMH_GENERAL_COORDS:
        rc, bad = O.run_gc(params, om, sp, opos, omom, step_begin, steps, np.ascontiguousarray(noise))
    else:
        rc, bad = O.run_cartesian(params, om, sp, opos, omom, step_begin, steps, np.ascontiguousarray(noise))
    assert rc == 0, (rc, bad)
    return (gpos, gmom), (opos, omom), noise


def small_grid(nx=48, ny=40, nt=12, ecc=0.2):
    return media.hydro_grid(nx=nx, ny=ny, nt=nt, L=15.0, tau0=0.6, tau1=12.0, T0=0.45, ecc=ecc)


def run_grid_pair(params, g, masses, counts, pos, mom, steps, seed=0, step_begin=1, spl=0,
                  rta=False, rta_time=None, initial_boost=False):
    pdim, N = mom.shape[0], pos.shape[1]
    drag, rt, jt = tables(masses, g.T, params.DsT if params.DsT > 0 else 0.2, pdim, rta, rta_time)
    noise = None
    with lmhip.Context(0) as ctx:
        ctx.set_medium_grid2p1(g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, g.tgrid, g.T, g.vx, g.vy)
        load_tables(ctx, drag, rt, jt)
        ctx.set_particles(pos, mom, masses, counts)
        if params.noise == _ffi.LMH_NOISE_FIXED:
            rng = np.random.default_rng(seed)
            noise = rta_noise_fixup(params, rng.standard_normal((steps, ctx.noise_channels(params), N)), 2, pdim)
        if initial_boost:
            ctx.boost_to_lab(params, 0)
        ctx.run(params, step_begin, steps, noise, spl)
        gpos, gmom = ctx.get_particles()
    if params.noise == _ffi.LMH_NOISE_PHILOX:
        noise = O.noise_record(params, 2, pdim, N, step_begin, steps)
    keep = O.Keep()
    gm = O.grid_medium(keep, g)
    sp = O.species(keep, masses, counts, drag=drag, rta=rt, jt=jt)
    opos, omom = pos.copy(order="F"), mom.copy(order="F")
    assert not initial_boost
    rc, bad = O.run_grid(params, gm, sp, opos, omom, step_begin, steps, np.ascontiguousarray(noise))
    assert rc == 0, (rc, bad)
    return (gpos, gmom), (opos, omom)
