"""The initial-state samplers (utils.jl:33-172 sample_particles_from_FONLL, utils.jl:499-558
append_thermal_pz): the host-built inverse-CDF tables against closed forms, and the oracle's
draws (the device samplers' definition) against the distributions they sample, by
Kolmogorov-Smirnov tests with seeded draws.  The reference draws from Julia's RNG, so the
samplers agree with it statistically, not draw by draw."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate, stats

import _oracle as O
from lmhip import utils as U

RG = np.linspace(0.0, 10.0, 41)
PG = np.linspace(0.0, 20.0, 161)
T_P = 2.0


def _density():
    """f(r, p) = exp(-p/T_P) exp(-r^2/8) as the reference's (Np, Nr) matrix f[p, r]"""
    return np.exp(-PG[:, None] / T_P) * np.exp(-RG[None, :] ** 2 / 8.0)


def _piecewise_cdf(knots, cdf):
    return lambda x: np.interp(x, knots, cdf)


def test_fonll_tables_closed_forms():
    f = _density()
    t = U.fonll_tables(RG, PG, f, cartesian_spatial_sampling=False)
    # every CDF rises from 0 to 1
    for c in [t["cdf_r"]] + [t["cdf_p"][:, j] for j in range(RG.size)]:
        assert c[0] == 0.0 and c[-1] == 1.0 and np.all(np.diff(c) > 0)
    # the reference's cumsum (node k's mass counted at node k, the first node set to 0) is the
    # closed-form CDF half a cell up, normalised at the last node, to second order in the cell:
    # p|r ~ p exp(-p/T) has CDF 1 - exp(-P/T)(1 + P/T) ...
    dr, dp = RG[1] - RG[0], PG[1] - PG[0]
    F_p = lambda q: 1.0 - np.exp(-q / T_P) * (1.0 + q / T_P)
    exact = F_p(PG + dp / 2) / F_p(PG[-1] + dp / 2)
    exact[0] = 0.0
    assert np.max(np.abs(t["cdf_p"][:, 5] - exact)) < 1e-3
    # ... and the r marginal ~ r exp(-r^2/8) has CDF 1 - exp(-r^2/8)
    F_r = lambda q: 1.0 - np.exp(-q ** 2 / 8.0)
    exact_r = F_r(RG + dr / 2) / F_r(RG[-1] + dr / 2)
    exact_r[0] = 0.0
    assert np.max(np.abs(t["cdf_r"] - exact_r)) < 4e-3
    # the transpose is accepted, a wrong shape is the reference's error
    t2 = U.fonll_tables(RG, PG, f.T, cartesian_spatial_sampling=False)
    assert np.array_equal(t2["cdf_r"], t["cdf_r"])
    with pytest.raises(ValueError, match="f_HQ_init_FONLL has size"):
        U.fonll_tables(RG, PG, f[:-1], cartesian_spatial_sampling=False)
    # Cartesian: the spatial density n(r) = sum_p p f dp and its envelope max + eps
    tc = U.fonll_tables(RG, PG, f, cartesian_spatial_sampling=True)
    dp = float(np.mean(np.diff(PG)))
    assert np.allclose(tc["n_r"], np.exp(-RG ** 2 / 8.0) * np.sum(PG * np.exp(-PG / T_P)) * dp, rtol=1e-13)
    assert tc["rmax"] == RG[-1] and tc["fmax"] == np.max(tc["n_r"]) + np.finfo(float).eps
    # a density that vanishes at one r node gives |p| = 0 there
    f0 = f.copy()
    f0[:, 3] = 0.0
    assert U.fonll_tables(RG, PG, f0)["p_ok"][3] == 0.0


@pytest.mark.parametrize("cartesian", [False, True])
def test_fonll_draws_follow_the_tables(cartesian):
    t = U.fonll_tables(RG, PG, _density(), cartesian_spatial_sampling=cartesian)
    N = 100_000
    x, p = O.sample_fonll(t, N, antithetic=False, seed=11, stream=2)
    r = np.hypot(x[0], x[1])
    if not cartesian:
        # r = inverse CDF of u: exactly the piecewise-linear CDF through (r_grid, cdf_r)
        assert stats.kstest(r, _piecewise_cdf(RG, t["cdf_r"])).pvalue > 1e-4
    else:
        # uniform (x, y) in the disk kept with probability n(r)/f_max: r ~ r n(r)
        n_of = lambda q: np.interp(q, RG, t["n_r"])
        grid = np.linspace(0.0, RG[-1], 4001)
        dens = grid * n_of(grid)
        cdf = integrate.cumulative_trapezoid(dens, grid, initial=0.0)
        assert stats.kstest(r, _piecewise_cdf(grid, cdf / cdf[-1])).pvalue > 1e-4
        assert np.max(r) <= RG[-1]
    # the azimuths are uniform and independent of the radius
    phi = np.arctan2(x[1], x[0])
    assert stats.kstest(phi, "uniform", args=(-math.pi, 2 * math.pi)).pvalue > 1e-4
    # |p| for the quarks of one r node follows that node's p|r CDF
    j = np.clip(np.searchsorted(RG, r, side="right"), 1, RG.size) - 1
    sel = j == 4
    assert np.count_nonzero(sel) > 2000
    pm = np.hypot(p[0], p[1])[sel]
    assert stats.kstest(pm, _piecewise_cdf(PG, t["cdf_p"][:, 4])).pvalue > 1e-4
    phip = np.arctan2(p[1], p[0])
    assert stats.kstest(phip, "uniform", args=(-math.pi, 2 * math.pi)).pvalue > 1e-4
    assert abs(np.corrcoef(phi, phip)[0, 1]) < 0.02


def test_fonll_antithetic_pairs():
    """simulate_cpu.jl:66-89: quarks 2k and 2k+1 share the position and have opposite momenta;
    an odd last quark is drawn on its own"""
    t = U.fonll_tables(RG, PG, _density())
    x, p = O.sample_fonll(t, 1001, antithetic=True, seed=5)
    assert np.array_equal(x[:, 0:1000:2], x[:, 1:1000:2])
    assert np.array_equal(p[:, 0:1000:2], -p[:, 1:1000:2])
    xs, ps = O.sample_fonll(t, 501, antithetic=False, seed=5)
    # pair k is draw k of the plain sampler, the odd last one draw 500
    assert np.array_equal(x[:, 0:1000:2], xs[:, :500]) and np.array_equal(x[:, 1000], xs[:, 500])
    assert np.array_equal(p[:, 1000], ps[:, 500])


@pytest.mark.parametrize("mT,T", [(1.6, 0.3), (1.6, 0.15), (0.3, 0.4)])
def test_thermal_pz_distribution(mT, T):
    """sample_pz_conditional_juttner draws p_z ~ exp(-sqrt(mT^2 + pz^2)/T): both envelopes
    (a = mT/T > 1.05 Gaussian in the rapidity, else Laplace in p_z)"""
    N = 40_000
    xg = np.linspace(0.0, 10.0, 5)
    tg = np.array([0.0, 1.0])
    Tf = np.full((xg.size, tg.size), T, order="F")
    keep = O.Keep()
    med = O.radial_medium(keep, xg, tg, Tf, np.zeros_like(Tf))
    pT = math.sqrt(mT * mT - 1.0) if mT > 1.0 else 0.0
    m = math.sqrt(mT * mT - pT * pT)
    mom = np.zeros((2, N), order="F")
    mom[0] = pT
    pos = np.full((2, N), 1.0, order="F")
    out = O.append_thermal_pz(mom, pos, m, 0.5, med, antithetic=False, seed=21, stream=1)
    assert np.array_equal(out[:2], mom)
    pz = out[2]
    zmax = 40.0 * T + 10.0 * mT
    norm = integrate.quad(lambda z: math.exp(-math.sqrt(mT * mT + z * z) / T), -zmax, zmax, limit=200)[0]
    grid = np.linspace(-zmax, zmax, 20001)
    dens = np.exp(-np.sqrt(mT * mT + grid * grid) / T) / norm
    cdf = integrate.cumulative_trapezoid(dens, grid, initial=0.0)
    assert stats.kstest(pz, _piecewise_cdf(grid, cdf / cdf[-1])).pvalue > 1e-4
    # antithetic: quark 2k+1 takes minus the p_z of quark 2k
    anti = O.append_thermal_pz(mom[:, :10], pos[:, :10], m, 0.5, med, antithetic=True, seed=21, stream=1)
    assert np.array_equal(anti[2, 1::2], -anti[2, 0::2]) and np.array_equal(anti[2, 0::2], pz[:10:2])


def _uniform_bath(T):
    keep = O.Keep()
    xg = np.linspace(0.0, 10.0, 5)
    tg = np.array([0.0, 1.0])
    Tf = np.full((xg.size, tg.size), T, order="F")
    return keep, O.radial_medium(keep, xg, tg, Tf, np.zeros_like(Tf))


def _juttner3(rng, M, T, N):
    """3-D Juttner momenta: |p| ~ p^2 exp(-(E - M)/T) by rejection from the Gamma(3, T) density
    p^2 exp(-p/T) (acceptance exp(-(E - p)/T) <= 1), isotropic directions"""
    out = []
    while sum(a.size for a in out) < N:
        p = rng.gamma(3.0, T, 2 * N)
        E = np.sqrt(p * p + M * M)
        out.append(p[rng.random(p.size) < np.exp(-(E - p) / T)])
    p = np.concatenate(out)[:N]
    z = rng.uniform(-1.0, 1.0, N)
    phi = rng.uniform(0.0, 2 * math.pi, N)
    rho = np.sqrt(1.0 - z * z)
    return np.asfortranarray(np.stack([p * rho * np.cos(phi), p * rho * np.sin(phi), p * z]))


def test_thermal_pz_second_moment_and_isotropy():
    """test_momentum_dims3.jl:98-117: <p_z^2> of the conditional draw at fixed m_T against
    quadrature (rtol 0.02), zero mean; the transverse rows of a 3-D Juttner sample completed with
    the conditional p_z are isotropic; antithetic mirroring is exact"""
    for mT, T in ((1.6, 0.3), (2.5, 0.45), (1.5, 0.156)):
        keep, med = _uniform_bath(T)
        N = 200_000
        mom = np.zeros((2, N), order="F")
        mom[0] = math.sqrt(mT * mT - 1.0)
        pos = np.ones((2, N), order="F")
        s = O.append_thermal_pz(mom, pos, 1.0, 0.5, med, seed=7, stream=3)[2]
        w = lambda z: math.exp(-(math.sqrt(mT * mT + z * z) - mT) / T)
        num = integrate.quad(lambda z: z * z * w(z), -np.inf, np.inf, epsrel=1e-10)[0]
        den = integrate.quad(w, -np.inf, np.inf, epsrel=1e-10)[0]
        assert math.isclose(np.mean(s ** 2), num / den, rel_tol=0.02)
        assert abs(np.mean(s)) < 4 * np.std(s) / math.sqrt(s.size)
    M, T = 1.5, 0.30
    keep, med = _uniform_bath(T)
    P3 = _juttner3(np.random.default_rng(7), M, T, 200_000)
    x = np.zeros((2, P3.shape[1]), order="F")
    P = O.append_thermal_pz(np.asfortranarray(P3[:2]), x, M, 0.5, med, seed=8)
    assert math.isclose(np.mean(P[2] ** 2), np.mean(P3[0] ** 2), rel_tol=0.02)
    assert math.isclose(np.mean(P[2] ** 2), np.mean(P[0] ** 2), rel_tol=0.02)
    Pa = O.append_thermal_pz(np.asfortranarray(P3[:2, :1000]), x[:, :1000], M, 0.5, med, antithetic=True, seed=8)
    assert np.array_equal(Pa[2, 1::2], -Pa[2, 0::2])
