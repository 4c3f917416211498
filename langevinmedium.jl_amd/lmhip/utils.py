"""Host utilities the drivers use before the step: the momentum-dimension contract, the FONLL
sampler's inverse-CDF tables (the draws run on the device, Context.sample_fonll), the thermal p_z
completion and the Milne initial-state sampler (mirrors of src/utils.jl), and the driver's
step-count formulas (src/simulate_cpu.jl:47-50, src/simulate_cpu_general_coords.jl:32-35).
The host samplers draw from numpy generators, so they agree with the reference statistically
only."""
from __future__ import annotations

import math

import numpy as np

_EPS = float(np.finfo(np.float64).eps)      # Julia eps(Float64)


def check_momentum_dims(dimensions: int, pdim: int, radial_mode: bool, bjorken_redshift: bool,
                        initial_time: float) -> None:
    """utils.jl:562-574"""
    if pdim != dimensions:
        if not (dimensions == 2 and pdim == 3 and not radial_mode):
            raise ValueError(f"momentum_dimensions={pdim} with dimensions={dimensions} is not supported: "
                             "only (dimensions=2, momentum_dimensions=3) differs from the coupled default.")
    if bjorken_redshift:
        if not (dimensions == 2 and pdim == 3):
            raise ValueError("bjorken_redshift requires dimensions=2 and momentum_dimensions=3 (p_z is row 3).")
        if not initial_time > 0:
            raise ValueError("bjorken_redshift needs initial_time > 0 (Milne tau; dp_z/dtau = -p_z/tau).")


def step_counts(initial_time: float, final_time: float, dt: float, save_interval: float):
    """simulate_cpu.jl:47-50: (steps, save_every, num_saves)."""
    total_time = final_time - initial_time
    steps = math.floor(total_time / dt)
    save_every = round(save_interval / dt)          # Julia round: nearest, ties to even
    if save_every == 0:
        raise ZeroDivisionError("save_interval/dt rounds to 0 (DivideError in the reference)")
    return steps, save_every, steps // save_every


def step_counts_gc(initial_time: float, final_time: float, dt: float, save_interval: float):
    """simulate_cpu_general_coords.jl:32-35: save_every = Int(save_interval/dt), which throws
    unless the ratio is integral."""
    total_time = final_time - initial_time
    steps = math.floor(total_time / dt)
    r = save_interval / dt
    if r != math.floor(r):
        raise ValueError(f"InexactError: Int64({r})")
    save_every = int(r)
    if save_every == 0:
        raise ZeroDivisionError("save_every == 0")
    return steps, save_every, steps // save_every


def sample_pz_conditional_juttner(mT: float, T: float, rng: np.random.Generator) -> float:
    """utils.jl:499-530: one draw from f(p_z) ~ exp(-sqrt(mT^2 + p_z^2)/T)."""
    if not (T > 0.0 and mT > 0.0):
        return 0.0
    a = mT / T
    if a > 1.05:
        c = max(2.0, a / (a - 1.0))
        sig = math.sqrt(c / a)
        for _ in range(100_000):
            y = sig * rng.standard_normal()
            ch = math.cosh(y)
            acc = ch * math.exp(-a * (ch - 1.0) + a * y * y / (2.0 * c))
            if rng.random() <= acc:
                return mT * math.sinh(y)
        return 0.0
    b = math.sqrt(2.0) * T
    for _ in range(1_000_000):
        pz = -b * math.log(rng.random()) * (-1.0 if rng.random() < 0.5 else 1.0)
        acc = math.exp(-(math.sqrt(mT * mT + pz * pz) - (mT + abs(pz)) / math.sqrt(2.0)) / T)
        if rng.random() <= acc:
            return pz
    return 0.0


def append_thermal_pz(momenta, positions, m, T_of_r, *, antithetic=False, rng=None):
    """utils.jl:533-558: (2, N) LRF momenta -> (3, N) with a thermal p_z row (host draws; the
    drivers use the device version, Context.append_thermal_pz)."""
    momenta = np.asarray(momenta, dtype=np.float64)
    positions = np.asarray(positions, dtype=np.float64)
    if momenta.shape[0] != 2:
        raise ValueError(f"append_thermal_pz: expected 2 transverse momentum rows, got {momenta.shape[0]}")
    rng = rng if rng is not None else np.random.default_rng()
    N = momenta.shape[1]
    out = np.zeros((3, N), dtype=np.float64, order="F")
    out[0:2, :] = momenta
    r = np.sqrt(np.sum(positions * positions, axis=0))
    for i in range(N):
        if antithetic and (i % 2 == 1):
            out[2, i] = -out[2, i - 1]
            continue
        T = max(float(T_of_r(r[i])), 0.0)
        mT = math.sqrt(float(m) ** 2 + momenta[0, i] ** 2 + momenta[1, i] ** 2)
        out[2, i] = sample_pz_conditional_juttner(mT, T, rng)
    return out


def _monotone_normalised(cdf):
    """utils.jl:60-67 / 134-141: the first node 0, every node at least eps above its predecessor,
    divided by the last, the last exactly 1"""
    cdf[0] = 0.0
    for k in range(1, cdf.size):
        cdf[k] = max(cdf[k], cdf[k - 1] + _EPS)
    cdf /= cdf[-1]
    cdf[-1] = 1.0
    return cdf


def fonll_tables(r_grid, p_grid, f_HQ_init_FONLL, cartesian_spatial_sampling=False):
    """The inverse-CDF tables of sample_particles_from_FONLL (utils.jl:33-146), built as the
    reference builds them: P(p|r) ~ p f(r, p) per r node (utils.jl:52-71), then either the r
    marginal of P(r, p) ~ r p f(r, p) (polar sampling, utils.jl:130-146) or the spatial density
    n(r) = sum_p p f(r, p) dp with its envelope max n + eps (Cartesian sampling, utils.jl:80-90).
    f is (Np, Nr), f[p, r], or its transpose."""
    rg = np.ascontiguousarray(r_grid, dtype=np.float64).ravel()
    pg = np.ascontiguousarray(p_grid, dtype=np.float64).ravel()
    Nr, Np = rg.size, pg.size
    f = np.asarray(f_HQ_init_FONLL, dtype=np.float64)
    if f.shape != (Np, Nr):
        if f.shape == (Nr, Np):
            f = f.T
        else:
            raise ValueError(f"sample_particles_from_FONLL: f_HQ_init_FONLL has size {f.shape}, expected "
                             f"({Np}, {Nr}) (p×r) or ({Nr}, {Np}) (r×p).")
    dr = float(np.mean(np.diff(rg)))
    dp = float(np.mean(np.diff(pg)))
    P = pg[:, None] * f
    cdf_p = np.zeros((Np, Nr), dtype=np.float64, order="F")
    p_ok = np.zeros(Nr, dtype=np.float64)
    for i in range(Nr):
        pdf = P[:, i]
        if np.sum(pdf) > 0:
            cdf_p[:, i] = _monotone_normalised(np.cumsum(pdf) * dp)
            p_ok[i] = 1.0
    out = dict(r_grid=rg, p_grid=pg, cdf_p=cdf_p, p_ok=p_ok, cartesian=bool(cartesian_spatial_sampling),
               cdf_r=None, n_r=None, rmax=0.0, fmax=0.0)
    if cartesian_spatial_sampling:
        n_r = np.sum(P, axis=0) * dp
        out.update(n_r=n_r, rmax=float(rg[-1]), fmax=float(np.max(n_r)) + _EPS)
    else:
        Prp = rg[None, :] * pg[:, None] * f
        Z = np.sum(Prp) * dr * dp
        if Z == 0:
            raise ValueError("Distribution normalization is zero.")
        Prp /= Z
        P_r = np.sum(Prp, axis=0) * dp
        out.update(cdf_r=_monotone_normalised(np.cumsum(P_r) * dr))
    return out


def sample_initial_particles_milne(m, dim, N_particles, tau, T_profile, ur_profile, mu_profile,
                                   x_range, nbins, rng=None):
    """utils.jl:350-400: rejection-sampled r on a binned Boltzmann weight, thermal p^r;
    positions (tau, r), momenta (p^tau, |p^r|)."""
    if dim != 2:
        raise AssertionError("Milne sampling requires dim = 2 (tau, r)")
    rng = rng if rng is not None else np.random.default_rng()
    edges = np.linspace(x_range[0], x_range[1], nbins + 1)
    dx = edges[1] - edges[0]
    centers = (edges[:-1] + edges[1:]) / 2
    Tv = np.array([T_profile(x, tau) for x in centers])
    urv = np.array([ur_profile(x, tau) for x in centers])
    muv = np.array([mu_profile(x, tau) for x in centers])
    gv = np.sqrt(1 + urv ** 2)
    n_boltz = Tv ** 1.5 / gv * np.exp((muv - m * gv) / Tv)
    pdf = n_boltz / np.sum(n_boltz) / dx
    pmax = np.max(pdf)
    sampled = []
    while len(sampled) < N_particles:
        r_try = rng.uniform(x_range[0], x_range[1], size=max(1024, N_particles))
        idx = np.searchsorted(centers, r_try, side="left")
        p = np.where(idx < len(pdf), pdf[np.minimum(idx, len(pdf) - 1)], 0.0)
        acc = rng.random(r_try.size) < p / pmax
        sampled.extend(r_try[acc].tolist())
    r = np.asarray(sampled[:N_particles])
    T = np.array([T_profile(x, tau) for x in r])
    sig = np.sqrt(m * T)
    pr = sig * rng.standard_normal(N_particles)
    ptau = np.sqrt(m * m + pr * pr)
    pos = np.asfortranarray(np.vstack([np.full(N_particles, float(tau)), r]))
    mom = np.asfortranarray(np.vstack([ptau, np.abs(pr)]))
    return pos, mom
