"""Media and initial states of the benchmark workloads.

The reference takes its medium as (n_r, n_t) matrices T and v on an (r, t) grid
(simulate.jl:21-27).  The 2+1D (x, y, tau) grid medium is an extension carried by
`Grid2p1Medium`.  The builders below make the synthetic inputs of the workloads: a static
uniform bath, a Bjorken-cooled bath and a 2+1D hydro-like grid; all particle samplers are
seeded numpy generators.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .transport import fmGeV


@dataclass
class Grid2p1Medium:
    """2+1D medium: nodes x0 + k*dx (k < nx), y0 + k*dy (k < ny), sorted tgrid, and fields
    T, vx, vy of shape (nx, ny, nt) (the flow is the fluid 3-velocity)."""
    x0: float
    dx: float
    nx: int
    y0: float
    dy: float
    ny: int
    tgrid: np.ndarray
    T: np.ndarray
    vx: np.ndarray
    vy: np.ndarray

    @property
    def nt(self) -> int:
        return int(np.asarray(self.tgrid).size)

    def nbytes(self) -> int:
        return 4 * 8 * self.nx * self.ny * self.nt       # device layout: (T, vx, vy, pad)


def DsT_from_2piT(Ds2piT: float) -> float:
    """D_s*T from the conventional D_s(2 pi T)."""
    return Ds2piT / (2.0 * math.pi)


# ---- static uniform bath, fluid at rest (Cartesian, 3 momentum dimensions) ---------------------
def uniform_bath(T=0.3, xmax=30.0, tmax=10.0):
    xgrid = np.arange(0.0, xmax + 0.25, 0.5)
    tgrid = np.arange(0.0, tmax + 0.5, 1.0)
    Tf = np.full((xgrid.size, tgrid.size), float(T), order="F")
    Vf = np.zeros((xgrid.size, tgrid.size), order="F")
    return xgrid, tgrid, Tf, Vf


def isotropic_momenta(N, pmax, dims, rng):
    """|p| uniform in [0, pmax] with isotropic directions (normalised Gaussians)."""
    mag = pmax * rng.random(N)
    d = rng.standard_normal((dims, N))
    d /= np.sqrt(np.sum(d * d, axis=0))
    return np.asfortranarray(d * mag)


# ---- Bjorken-cooled bath: T(tau) = T0 (tau0/tau)^(1/3), uniform in r ---------------------------
def bjorken_bath(T0=0.4, tau0=0.6, Tstop=0.155, rmax=20.0, ntau=2048):
    tau_end = tau0 * (T0 / Tstop) ** 3
    xgrid = np.linspace(0.0, rmax, 41)
    tgrid = np.linspace(tau0, tau_end + 0.05, ntau)
    Tcol = T0 * (tau0 / tgrid) ** (1.0 / 3.0)
    Tf = np.asfortranarray(np.tile(Tcol, (xgrid.size, 1)))
    Vf = np.zeros_like(Tf, order="F")
    return xgrid, tgrid, Tf, Vf, tau_end


def power_law_pt(N, rng, ptmax=20.0):
    """p_T with dN/dp_T ~ p_T/(1 + p_T^2/4)^3 on [0, ptmax], by inverse CDF:
    F(p) = (1 - s^-2)/(1 - s_max^-2) with s = 1 + p^2/4."""
    smax = 1.0 + ptmax * ptmax / 4.0
    u = rng.random(N)
    s = (1.0 - u * (1.0 - smax ** -2)) ** -0.5
    return 2.0 * np.sqrt(np.maximum(s - 1.0, 0.0))


def pt_rapidity_momenta(N, m, rng, ymax=1.0, ptmax=20.0):
    """(3, N) momenta: power-law p_T with uniform azimuth, p_z = m_T sinh(y), y ~ U[-ymax, ymax]."""
    pt = power_law_pt(N, rng, ptmax)
    phi = 2.0 * math.pi * rng.random(N)
    y = ymax * (2.0 * rng.random(N) - 1.0)
    mt = np.sqrt(m * m + pt * pt)
    return np.asfortranarray(np.vstack([pt * np.cos(phi), pt * np.sin(phi), mt * np.sinh(y)]))


# ---- 2+1D hydro-like grid ------------------------------------------------------------------------
def hydro_grid(nx=200, ny=200, nt=120, L=15.0, tau0=0.6, tau1=12.0, T0=0.45, sx=3.0, sy=4.0,
               H=0.08, ecc=0.0):
    """T(tau, x, y) = T0 (tau0/tau)^(1/3) exp(-x^2/(2 sx^2) - y^2/(2 sy^2)), radial flow
    v_r = tanh(H r (tau - tau0)) along r_hat.  ecc > 0 stretches the profile:
    sx -> sx (1 - ecc/2), sy -> sy (1 + ecc/2)."""
    x = np.linspace(-L, L, nx)
    y = np.linspace(-L, L, ny)
    tg = np.linspace(tau0, tau1, nt)
    sxe, sye = sx * (1.0 - 0.5 * ecc), sy * (1.0 + 0.5 * ecc)
    X, Y = np.meshgrid(x, y, indexing="ij")
    prof = np.exp(-X * X / (2 * sxe * sxe) - Y * Y / (2 * sye * sye))
    R = np.sqrt(X * X + Y * Y)
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(R > 0, X / R, 0.0)
        cy = np.where(R > 0, Y / R, 0.0)
    T = np.empty((nx, ny, nt), order="F")
    vx = np.empty((nx, ny, nt), order="F")
    vy = np.empty((nx, ny, nt), order="F")
    for k, tau in enumerate(tg):
        T[:, :, k] = T0 * (tau0 / tau) ** (1.0 / 3.0) * prof
        vr = np.tanh(H * R * (tau - tau0))
        vx[:, :, k] = vr * cx
        vy[:, :, k] = vr * cy
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    return Grid2p1Medium(float(x[0]), float(dx), nx, float(y[0]), float(dy), ny, tg, T, vx, vy)


def hydro_positions(N, rng, sx=3.0, sy=4.0, ecc=0.0):
    """positions with density ~ T^4 at tau0: Gaussian with sigma = s/2 along each axis."""
    sxe, sye = sx * (1.0 - 0.5 * ecc), sy * (1.0 + 0.5 * ecc)
    return np.asfortranarray(np.vstack([0.5 * sxe * rng.standard_normal(N),
                                        0.5 * sye * rng.standard_normal(N)]))


def tau_R_law(eta_over_s=0.16, c=5.0):
    """RTA relaxation time tau_R = c (eta/s)/T [GeV^-1] in fm."""
    return lambda T: c * eta_over_s / np.asarray(T, dtype=np.float64) / fmGeV


def hydro_ensemble(N, seed=3, nx=200, ny=200, nt=120):
    """The 2+1D hydro workload: the hydro_grid medium (200 x 200 cells over +-15 fm, 120
    tau-slices from 0.6 to 12 fm/c, T0 = 0.45 GeV), N quarks of which 70% charm (m = 1.5 GeV)
    and 30% bottom (m = 4.75 GeV), positions ~ T^4 at tau0, p_T ~ p_T/(1 + p_T^2/4)^3 on
    [0, 20] GeV with rapidity uniform in [-1, 1].  Returns (grid, masses, counts, pos, mom)."""
    rng = np.random.default_rng(seed)
    g = hydro_grid(nx=nx, ny=ny, nt=nt, L=15.0, tau0=0.6, tau1=12.0, T0=0.45)
    nc = int(round(0.7 * N))
    counts = [nc, N - nc]
    masses = [1.5, 4.75]
    pos = hydro_positions(N, rng)
    mom = np.asfortranarray(np.hstack([pt_rapidity_momenta(counts[0], masses[0], rng),
                                       pt_rapidity_momenta(counts[1], masses[1], rng)]))
    return g, masses, counts, pos, mom


def hydro_tables(g, masses, DsT, pdim, eta_over_s=0.16):
    """per species: (drag table, RTA table of tau_R = 5 (eta/s)/T, Juttner inverse CDF) over the
    medium's temperature range, at the resolutions the hydro workload uses"""
    from . import transport
    Tmin = max(float(np.min(g.T)), 0.0)
    Tmax = max(float(np.max(g.T)), Tmin + np.finfo(float).eps)
    out = []
    for m in masses:
        drag = transport.build_tau_drag_spline(m, DsT, Tmin=Tmin, Tmax=Tmax, nT=1024)
        rta = transport.build_relaxation_spline(tau_R_law(eta_over_s), Tmin=Tmin, Tmax=Tmax, nT=1024)
        jt = transport.build_juttner_invcdf(m, pdim, Tmin=Tmin, Tmax=Tmax, nT=256, nU=2048, np_=8192)
        out.append((drag, rta, jt))
    return out
