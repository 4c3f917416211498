"""`simulate_ensemble_bulk` for the MI355X backends: the mirror of the reference's frontend
(src/simulate.jl:21-141, GPU superset src/simulate_gpu_wrapper.jl:15-108) and of its CPU drivers
(src/simulate_cpu.jl:12-323, src/simulate_cpu_general_coords.jl:18-133).

The host part (step counts, initial state handling, table builds, snapshots) runs here, as it does
in the reference's drivers; every step runs in liblmhip.so.  Arguments keep the reference's names,
meanings and defaults; the extensions are keyword arguments the reference does not have:

  noise            "philox" (default) or a fixed-noise record of shape (steps, C, N), the layout
                   of include/lmhip.h, read instead of the driver's randn draws
  seed             Philox key
  steps_per_launch steps a particle stays in registers per kernel launch (0: library default)
  species_counts   contiguous species segments when m is a list of masses
  drag_model       "table" (the reference) or ("gamma_p", gamma0, Lambda)
  rta_time         relaxation-time law tau(T) [fm] for the RTA table instead of tau_n_main3
  collision_mode   also "langevin_rta": the Langevin kick followed by the BGK redraw
  milne            Milne coordinates: with dimensions=2, momentum_dimensions=3 (the reference's
                   Bjorken mode, bjorken_redshift implied) the positions gain a third row, the
                   space-time rapidity eta_s of each quark's free path (x_init may hold it as a
                   third row, else eta_s = 0); p_z is the momentum in the frame comoving with the
                   Bjorken flow at eta_s
"""
from __future__ import annotations

import math
import sys

import numpy as np

from . import _ffi
from .backends import HIP_GCBackend, HIPBackend
from .context import Context, LMHError
from .media import Grid2p1Medium
from .transport import (build_juttner_invcdf, build_relaxation_spline, build_tau_drag_spline,
                        build_taun_current_spline)
from .utils import (check_momentum_dims, fonll_tables, sample_initial_particles_milne, step_counts,
                    step_counts_gc)

_EPS = sys.float_info.epsilon


def simulate_ensemble_bulk(backend, *args, **kwargs):
    """Dispatch on the backend type as simulate.jl does."""
    if isinstance(backend, HIP_GCBackend):
        return _simulate_gc(backend, *args, **kwargs)
    if isinstance(backend, HIPBackend):
        if len(args) == 1 and isinstance(args[0], Grid2p1Medium):
            return _simulate_grid(backend, args[0], **kwargs)
        return _simulate_cartesian(backend, *args, **kwargs)
    raise TypeError(f"no method matching simulate_ensemble_bulk({type(backend).__name__}, ...)")


# ------------------------------------------------------------------------------------------------
def _collision_code(collision_mode, momentum_langevin, DsT):
    cm = str(collision_mode).lstrip(":")
    if cm not in ("langevin", "rta", "langevin_rta"):
        raise ValueError(f"simulate_ensemble_bulk(::HIPBackend): collision_mode={collision_mode} is not "
                         "supported on GPU (only :langevin and :rta).")
    if not momentum_langevin or DsT == 0.0:          # simulate_cpu.jl:229
        return _ffi.LMH_FLUID, cm
    return {"langevin": _ffi.LMH_LANGEVIN, "rta": _ffi.LMH_RTA,
            "langevin_rta": _ffi.LMH_LANGEVIN_RTA}[cm], cm


def _drag_params(drag_model):
    if drag_model in (None, "table"):
        return _ffi.LMH_DRAG_TABLE, 0.0, 1.0
    name, g0, lam = drag_model
    if name != "gamma_p":
        raise ValueError(f"unknown drag_model {drag_model!r}")
    return _ffi.LMH_DRAG_GAMMA_P, float(g0), float(lam)


def _species(m, species_counts, N):
    masses = [float(x) for x in np.atleast_1d(m)]
    if species_counts is None:
        if len(masses) != 1:
            raise ValueError("species_counts is needed with several masses")
        return masses, [N]
    counts = [int(c) for c in species_counts]
    if len(counts) != len(masses) or sum(counts) != N:
        raise ValueError("species_counts must match the masses and sum to N")
    return masses, counts


def _load_tables(ctx, masses, Tfield, *, DsT, momentum_langevin, cm, pdim, dst_kw, rta_time,
                 needs_drag_table=True):
    """simulate_cpu.jl:185-211 and simulate_gpu.jl:90-143: the drag table, the RTA time table
    and the Juttner inverse-CDF table, per species."""
    Tmin = max(float(np.min(Tfield)), 0.0)
    Tmax = max(float(np.max(Tfield)), Tmin + _EPS)
    rta = cm in ("rta", "langevin_rta")
    for k, mk in enumerate(masses):
        if momentum_langevin and DsT > 0.0:
            if needs_drag_table:
                t0_, inv_, vals = build_tau_drag_spline(mk, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024, **dst_kw)
                ctx.set_time_table(_ffi.LMH_TABLE_DRAG, k, t0_, inv_, vals)
            if rta:
                if rta_time is None:
                    t0_, inv_, vals = build_taun_current_spline(mk, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024, **dst_kw)
                else:
                    t0_, inv_, vals = build_relaxation_spline(rta_time, Tmin=Tmin, Tmax=Tmax, nT=1024)
                ctx.set_time_table(_ffi.LMH_TABLE_RTA, k, t0_, inv_, vals)
                inv, nU, nT, Tt, invdT = build_juttner_invcdf(mk, pdim, Tmin=Tmin, Tmax=Tmax,
                                                              nT=256, nU=2048, np_=8192)
                ctx.set_juttner_table(k, inv, nU, nT, Tt, invdT)
        else:
            # the reference's placeholders (simulate_cpu.jl:185-189, simulate_gpu.jl:121-125)
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, k, 0.0, 1.0, np.zeros(2))
            if rta:
                ctx.set_time_table(_ffi.LMH_TABLE_RTA, k, 0.0, 1.0, np.zeros(2))
                ctx.set_juttner_table(k, np.zeros(4), 2, 2, 0.0, 1.0)


def _run_chunks(ctx, params, steps, save_every, num_saves, noise, steps_per_launch, on_snapshot):
    """simulate_cpu.jl:215-316: run the steps, snapshot at step % save_every == 0.  A snapshot is
    taken on the device after its chunk and copied to the host while the next chunks run; the
    host side (on_snapshot(pos, mom), in order) trails by up to two snapshots."""
    fixed = params.noise == _ffi.LMH_NOISE_FIXED
    pending = []
    taken = 0
    step = 1
    while step <= steps:
        n = save_every if step + save_every - 1 <= steps else steps - step + 1
        rec = noise[step - 1: step - 1 + n] if fixed else None
        try:
            ctx.run(params, step, n, rec, steps_per_launch)
        except LMHError as e:
            if e.code == _ffi.LMH_E_NONFINITE:
                raise RuntimeError(str(e)) from None
            raise
        step += n
        if on_snapshot is not None and n == save_every and (step - 1) % save_every == 0 and taken < num_saves:
            pending.append(ctx.snapshot())
            taken += 1
            if len(pending) > 2:
                on_snapshot(*ctx.snapshot_get(pending.pop(0)))
    for k in pending:
        on_snapshot(*ctx.snapshot_get(k))


def _noise_mode(noise, steps, C, N):
    if isinstance(noise, str):
        if noise != "philox":
            raise ValueError("noise must be 'philox' or a fixed-noise record")
        return _ffi.LMH_NOISE_PHILOX, None
    rec = np.ascontiguousarray(noise, dtype=np.float64)
    if rec.shape != (steps, C, N):
        raise ValueError(f"fixed-noise record must be (steps, C, N) = {(steps, C, N)}")
    return _ffi.LMH_NOISE_FIXED, rec


# ------------------------------------------------------------------------------------------------
def _simulate_cartesian(backend, r_grid_Langevin, p_grid_Langevin, heavy_quark_density,
                        TemperatureEvolutionn, VelocityEvolutionn, SpaceTimeGrid, *,
                        N_particles: int = 10_000, Δt: float = 0.001, initial_time: float = 0.0,
                        final_time: float = 1.0, save_interval: float = 0.1, m=1.0,
                        DsT: float = 0.2, DsT_linear: bool = False, DsT_slope: float = 1.765,
                        DsT_offset: float = -0.159, Tfo: float = 0.156, DsT_quad: bool = False,
                        DsT_Tref: float = 0.0, dimensions: int = 3,
                        cartesian_spatial_sampling=None, antithetic_momenta: bool = False,
                        position_diffusion: bool = False, momentum_langevin: bool = True,
                        reflecting_boundary: bool = False, collision_mode="langevin",
                        x_init=None, p_init=None, V2Evolutionn=None, psi2: float = 0.0,
                        freezeout_capture: bool = False, freezeout_interp: bool = True,
                        integrator_mode: int = 0, relativistic: bool = True,
                        momentum_dimensions: int = 0, bjorken_redshift: bool = False,
                        noise="philox", seed: int = 0, steps_per_launch: int = 0,
                        species_counts=None, drag_model="table", rta_time=None, milne: bool = False):
    coll, cm = _collision_code(collision_mode, momentum_langevin, DsT)
    if milne:
        if dimensions != 2 or momentum_dimensions != 3:
            raise ValueError("milne needs dimensions=2 and momentum_dimensions=3 (x, y, eta_s; p_x, p_y, p_z)")
        if freezeout_capture or position_diffusion or reflecting_boundary:
            raise ValueError("milne: no freeze-out capture, position diffusion or reflecting boundary")
        bjorken_redshift = True
    steps, save_every, num_saves = step_counts(initial_time, final_time, Δt, save_interval)
    if freezeout_capture:
        # simulate_gpu.jl:68-75: no trajectory history, one chunk of all the steps
        num_saves, save_every = 1, max(steps, 1)
    xgrid, tgrid = (np.asarray(a, dtype=np.float64) for a in SpaceTimeGrid)
    Tf = np.asfortranarray(TemperatureEvolutionn, dtype=np.float64)
    Vf = np.asfortranarray(VelocityEvolutionn, dtype=np.float64)
    ctx = Context(backend.device)
    try:
        return _simulate_cartesian_ctx(
            ctx, r_grid_Langevin, p_grid_Langevin, heavy_quark_density, xgrid, tgrid, Tf, Vf,
            coll=coll, cm=cm, steps=steps, save_every=save_every, num_saves=num_saves,
            N_particles=N_particles, Δt=Δt, initial_time=initial_time, final_time=final_time, m=m,
            DsT=DsT, DsT_linear=DsT_linear, DsT_slope=DsT_slope, DsT_offset=DsT_offset, Tfo=Tfo,
            DsT_quad=DsT_quad, DsT_Tref=DsT_Tref, dimensions=dimensions,
            cartesian_spatial_sampling=cartesian_spatial_sampling, antithetic_momenta=antithetic_momenta,
            position_diffusion=position_diffusion, momentum_langevin=momentum_langevin,
            reflecting_boundary=reflecting_boundary, x_init=x_init, p_init=p_init,
            V2Evolutionn=V2Evolutionn, psi2=psi2, freezeout_capture=freezeout_capture,
            freezeout_interp=freezeout_interp, integrator_mode=integrator_mode, relativistic=relativistic,
            momentum_dimensions=momentum_dimensions, bjorken_redshift=bjorken_redshift, noise=noise,
            seed=seed, steps_per_launch=steps_per_launch, species_counts=species_counts,
            drag_model=drag_model, rta_time=rta_time, milne=milne)
    finally:
        ctx.close()


def _simulate_cartesian_ctx(ctx, r_grid_Langevin, p_grid_Langevin, heavy_quark_density, xgrid, tgrid, Tf, Vf,
                            *, coll, cm, steps, save_every, num_saves, N_particles, Δt, initial_time,
                            final_time, m, DsT, DsT_linear, DsT_slope, DsT_offset, Tfo, DsT_quad, DsT_Tref,
                            dimensions, cartesian_spatial_sampling, antithetic_momenta, position_diffusion,
                            momentum_langevin, reflecting_boundary, x_init, p_init, V2Evolutionn, psi2,
                            freezeout_capture, freezeout_interp, integrator_mode, relativistic,
                            momentum_dimensions, bjorken_redshift, noise, seed, steps_per_launch,
                            species_counts, drag_model, rta_time, milne=False):
    # the medium first: the thermal p_z completion reads T(r, t0) from it
    ctx.set_medium_radial(xgrid, tgrid, Tf, Vf, V2Evolutionn)
    if x_init is None or p_init is None:
        # simulate_cpu.jl:57-94: sample_particles_from_FONLL, Cartesian by default, with the
        # antithetic pairing; the draws run on the device (Philox block 3 under (seed, 0))
        cart = (dimensions == 1 or dimensions >= 2) if cartesian_spatial_sampling is None \
            else bool(cartesian_spatial_sampling)
        tables = fonll_tables(r_grid_Langevin, p_grid_Langevin, heavy_quark_density, cart)
        x_matrix, p_matrix = ctx.sample_fonll(tables, N_particles, antithetic=antithetic_momenta,
                                              seed=seed, stream=0)
    else:
        x_matrix = np.array(x_init, dtype=np.float64, order="F", ndmin=2)
        p_matrix = np.array(p_init, dtype=np.float64, order="F", ndmin=2)
    N = x_matrix.shape[1]
    eta0 = np.zeros(N)
    if milne and x_matrix.shape[0] == 3:
        eta0 = x_matrix[2].copy()
        x_matrix = np.asfortranarray(x_matrix[:2])
    if p_matrix.shape[1] != N or N_particles != N:
        raise ValueError(f"x_init/p_init carry {N} particles, N_particles is {N_particles}")
    radial_mode = dimensions == 1
    rng = np.random.default_rng(seed)
    if radial_mode:
        # radial reduction, simulate_cpu.jl:102-136
        r = np.sqrt(x_matrix[0] ** 2 + x_matrix[1] ** 2)
        dr0 = abs(float(r_grid_Langevin[1] - r_grid_Langevin[0])) if len(r_grid_Langevin) >= 2 else 0.0
        r_axis_eps = max(1e-12, 0.5 * dr0)
        reg = r > r_axis_eps
        ex = np.empty(N)
        ey = np.empty(N)
        ex[reg] = x_matrix[0, reg] / r[reg]
        ey[reg] = x_matrix[1, reg] / r[reg]
        ns = int(np.count_nonzero(~reg))
        if ns:
            rx, ry = rng.standard_normal(ns), rng.standard_normal(ns)
            invn = 1.0 / (np.sqrt(rx ** 2 + ry ** 2) + _EPS)
            ex[~reg], ey[~reg] = rx * invn, ry * invn
        positions = np.asfortranarray(r.reshape(1, N))
        momenta = np.asfortranarray((ex * p_matrix[0] + ey * p_matrix[1]).reshape(1, N))
    else:
        positions, momenta = x_matrix.copy(order="F"), p_matrix.copy(order="F")
    if positions.shape[0] != dimensions:
        raise ValueError(f"x_init has {positions.shape[0]} rows, dimensions is {dimensions}")
    pdim = dimensions if momentum_dimensions <= 0 else momentum_dimensions
    check_momentum_dims(dimensions, pdim, radial_mode, bjorken_redshift, initial_time)
    masses, counts = _species(m, species_counts, N)
    if momenta.shape[0] < pdim:
        if len(masses) != 1:
            raise ValueError("the thermal p_z completion needs a single species")
        # simulate_cpu.jl:147-151: the thermal p_z at T(r, t0), on the device
        momenta = ctx.append_thermal_pz(momenta, positions, masses[0], initial_time,
                                        antithetic=antithetic_momenta and x_init is None, seed=seed,
                                        stream=0)
    if momenta.shape[0] != pdim:
        raise ValueError(f"momentum rows {momenta.shape[0]} != momentum_dimensions {pdim}")

    dmodel, g0, lam = _drag_params(drag_model)
    dst_kw = dict(DsT_linear=DsT_linear, DsT_slope=DsT_slope, DsT_offset=DsT_offset, Tfo=Tfo,
                  DsT_quad=DsT_quad, DsT_Tref=DsT_Tref)
    _load_tables(ctx, masses, Tf, DsT=DsT, momentum_langevin=momentum_langevin, cm=cm, pdim=pdim,
                 dst_kw=dst_kw, rta_time=rta_time, needs_drag_table=dmodel == _ffi.LMH_DRAG_TABLE)
    if milne:
        positions = np.asfortranarray(np.vstack([positions, eta0]))
    ctx.set_particles(positions, momenta, masses, counts)
    params = _ffi.make_params(dt=Δt, t0=initial_time, DsT=DsT, psi2=psi2, gamma0=g0, Lambda=lam,
                              seed=seed, coords=_ffi.LMH_MILNE if milne else _ffi.LMH_CARTESIAN,
                              collision=coll,
                              integrator=int(integrator_mode), drag_model=dmodel,
                              relativistic=relativistic, bjorken_redshift=bjorken_redshift,
                              position_diffusion=position_diffusion,
                              reflecting_boundary=reflecting_boundary, check_finite=True)
    C = ctx.noise_channels(params)
    params.noise, rec = _noise_mode(noise, steps, C, N)
    ctx.boost_to_lab(params, 0)                     # simulate_cpu.jl:160-163
    if freezeout_capture:
        # simulate_gpu.jl:375-402: the latched states are the result
        ctx.freezeout_enable(Tfo, freezeout_interp)
        _run_chunks(ctx, params, steps, save_every, num_saves, rec, steps_per_launch, None)
        fpos, fmom, ftau, fflag = ctx.freezeout_get()
        return {"pos": fpos, "mom": fmom, "tau": ftau, "flag": fflag}
    pos0, mom0 = ctx.get_particles()
    mom_hist, pos_hist = [mom0], [pos0]

    def snap(p_, m_):
        pos_hist.append(p_)
        mom_hist.append(m_)

    _run_chunks(ctx, params, steps, save_every, num_saves, rec, steps_per_launch, snap)
    time_points = np.linspace(initial_time, final_time, num_saves + 1)
    return time_points, mom_hist, pos_hist


# ------------------------------------------------------------------------------------------------
def _simulate_gc(backend, T_profile_MIS, ur_profile_MIS, mu_profile_MIS, TemperatureEvolutionn,
                 VelocityEvolutionn, SpaceTimeGrid, *, N_particles: int = 10_000,
                 Δt: float = 0.001, initial_time: float = 0.0, final_time: float = 1.0,
                 save_interval: float = 0.1, m: float = 1.0, DsT: float = 0.2,
                 dimensions: int = 2, x_init=None, p_init=None, noise="philox", seed: int = 0,
                 steps_per_launch: int = 0):
    steps, save_every, num_saves = step_counts_gc(initial_time, final_time, Δt, save_interval)
    xgrid, tgrid = (np.asarray(a, dtype=np.float64) for a in SpaceTimeGrid)
    Tf = np.asfortranarray(TemperatureEvolutionn, dtype=np.float64)
    Vf = np.asfortranarray(VelocityEvolutionn, dtype=np.float64)
    rng = np.random.default_rng(seed)
    if x_init is not None and p_init is not None:
        positions = np.array(x_init, dtype=np.float64, order="F", ndmin=2)
        momenta = np.array(p_init, dtype=np.float64, order="F", ndmin=2)
        if positions.shape[0] != dimensions:
            raise AssertionError("Milne sampling requires dim = 2 (tau, r)")
    else:
        positions, momenta = sample_initial_particles_milne(
            m, dimensions, N_particles, initial_time, T_profile_MIS, ur_profile_MIS, mu_profile_MIS,
            (0.0, float(np.max(xgrid))), 200, rng)
    N = positions.shape[1]
    ctx = Context(backend.device)
    try:
        ctx.set_medium_radial(xgrid, tgrid, Tf, Vf)
        if DsT > 0.0:
            Tmin = max(float(np.min(Tf)), 0.0)
            Tmax = max(float(np.max(Tf)), Tmin + _EPS)
            t0_, inv_, vals = build_tau_drag_spline(m, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024)
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, 0, t0_, inv_, vals)
        else:
            ctx.set_time_table(_ffi.LMH_TABLE_DRAG, 0, 0.0, 1.0, np.zeros(2))
        ctx.set_particles(positions, momenta, [float(m)], [N])
        params = _ffi.make_params(dt=Δt, t0=initial_time, DsT=DsT, seed=seed,
                                  coords=_ffi.LMH_GENERAL_COORDS, collision=_ffi.LMH_LANGEVIN,
                                  check_finite=True)
        params.noise, rec = _noise_mode(noise, steps, 1, N)
        mom_hist = [np.abs(momenta[1, :]).copy()]          # simulate_cpu_general_coords.jl:53
        pos_hist = [positions[1:2, :].copy()]

        def snap(p_, m_):
            mom_hist.append(np.sqrt(m_[0] ** 2 + m_[1] ** 2))   # :120-122
            pos_hist.append(p_[1:2, :].copy())

        _run_chunks(ctx, params, steps, save_every, num_saves, rec, steps_per_launch, snap)
    finally:
        ctx.close()
    time_points = np.linspace(initial_time, final_time, num_saves + 1)
    return time_points, mom_hist, pos_hist


# ------------------------------------------------------------------------------------------------
def _simulate_grid(backend, medium: Grid2p1Medium, *, Δt: float = 0.01, initial_time: float = None,
                   final_time: float = None, save_interval: float = 0.1, m=1.5, DsT: float = 0.2,
                   DsT_linear: bool = False, DsT_slope: float = 1.765, DsT_offset: float = -0.159,
                   Tfo: float = 0.156, DsT_quad: bool = False, DsT_Tref: float = 0.0,
                   species_counts=None, x_init=None, p_init=None, momentum_langevin: bool = True,
                   collision_mode="langevin", integrator_mode: int = 0, relativistic: bool = True,
                   bjorken_redshift: bool = False, position_diffusion: bool = False,
                   noise="philox", seed: int = 0, steps_per_launch: int = 0,
                   drag_model="table", rta_time=None, freezeout_capture: bool = False,
                   freezeout_interp: bool = True):
    """The 2+1D grid medium (extension): positions (2, N), momenta (2 or 3, N) in the LRF."""
    coll, cm = _collision_code(collision_mode, momentum_langevin, DsT)
    tg = np.asarray(medium.tgrid, dtype=np.float64)
    t0 = float(tg[0]) if initial_time is None else float(initial_time)
    t1 = float(tg[-1]) if final_time is None else float(final_time)
    steps, save_every, num_saves = step_counts(t0, t1, Δt, save_interval)
    if freezeout_capture:
        num_saves, save_every = 1, max(steps, 1)
    positions = np.array(x_init, dtype=np.float64, order="F", ndmin=2)
    momenta = np.array(p_init, dtype=np.float64, order="F", ndmin=2)
    N = positions.shape[1]
    if positions.shape[0] != 2 or momenta.shape[0] not in (2, 3) or momenta.shape[1] != N:
        raise ValueError("the 2+1D grid needs positions (2, N) and momenta (2 or 3, N)")
    pdim = momenta.shape[0]
    check_momentum_dims(2, pdim, False, bjorken_redshift, t0)
    masses, counts = _species(m, species_counts, N)
    dmodel, g0, lam = _drag_params(drag_model)
    ctx = Context(backend.device)
    try:
        ctx.set_medium_grid2p1(medium.x0, medium.dx, medium.nx, medium.y0, medium.dy, medium.ny,
                               tg, medium.T, medium.vx, medium.vy)
        dst_kw = dict(DsT_linear=DsT_linear, DsT_slope=DsT_slope, DsT_offset=DsT_offset, Tfo=Tfo,
                      DsT_quad=DsT_quad, DsT_Tref=DsT_Tref)
        _load_tables(ctx, masses, medium.T, DsT=DsT, momentum_langevin=momentum_langevin, cm=cm,
                     pdim=pdim, dst_kw=dst_kw, rta_time=rta_time,
                     needs_drag_table=dmodel == _ffi.LMH_DRAG_TABLE)
        ctx.set_particles(positions, momenta, masses, counts)
        params = _ffi.make_params(dt=Δt, t0=t0, DsT=DsT, gamma0=g0, Lambda=lam, seed=seed,
                                  coords=_ffi.LMH_CARTESIAN, collision=coll,
                                  integrator=int(integrator_mode), drag_model=dmodel,
                                  relativistic=relativistic, bjorken_redshift=bjorken_redshift,
                                  position_diffusion=position_diffusion, check_finite=True)
        C = ctx.noise_channels(params)
        params.noise, rec = _noise_mode(noise, steps, C, N)
        ctx.boost_to_lab(params, 0)
        if freezeout_capture:
            ctx.freezeout_enable(Tfo, freezeout_interp)
            _run_chunks(ctx, params, steps, save_every, num_saves, rec, steps_per_launch, None)
            fpos, fmom, ftau, fflag = ctx.freezeout_get()
            return {"pos": fpos, "mom": fmom, "tau": ftau, "flag": fflag}
        pos0, mom0 = ctx.get_particles()
        mom_hist, pos_hist = [mom0], [pos0]

        def snap(p_, m_):
            pos_hist.append(p_)
            mom_hist.append(m_)

        _run_chunks(ctx, params, steps, save_every, num_saves, rec, steps_per_launch, snap)
    finally:
        ctx.close()
    time_points = np.linspace(t0, t1, num_saves + 1)
    return time_points, mom_hist, pos_hist
