"""Host-side transport tables, the mirror of the reference's Transport module (src/transport.jl).

These run on the host before the step, as they do in the reference: the step only evaluates the
tables (device side, lmhip_device.h).  Names, arguments and error behaviour follow transport.jl.
"""
from __future__ import annotations

import math
import os
import warnings

import numpy as np
from scipy.special import kve

GevInvTofm = 0.197327            # constants.jl:13
fmGeV = 1 / GevInvTofm           # constants.jl:23

_TINY = 1e-300


def _read_scale() -> float:
    # transport.jl:9-13: read at load time; a value other than 1 is a diagnostic mode
    s = float(os.environ.get("LV_TAUN_SCALE", "1.0"))
    if s != 1.0:
        warnings.warn(f"LangevInMedium: tau_n scaled by LV_TAUN_SCALE = {s} - DIAGNOSTIC MODE "
                      "(this also rescales eta_D, kappa and hence D_s)")
    return s


LV_TAUN_SCALE = [_read_scale()]


def besselkx(nu, z):
    """exp(z)*K_nu(z), the quantity Bessels.besselkx returns (transport.jl:77-78,87-88)."""
    return kve(nu, z)


def tau_n_main3(T, m, DsT) -> float:
    """transport.jl:58-104: tau_n = D_s z K3/K2 in fm, via (2K1 - 3K3 + K5) = (48/z^2)K3 for
    z <= 50 and the K3/K2 ratio above."""
    Tm, Mm, Ds = float(T), float(m), float(DsT)
    if not (math.isfinite(Tm) and math.isfinite(Mm) and math.isfinite(Ds)) or Tm <= 0 or Mm <= 0 or Ds <= 0:
        return 0.0
    z = Mm / Tm
    if z <= 0.0:
        return 0.0
    K2x = besselkx(2, z)
    K3x = besselkx(3, z)
    denominator = max(abs(K2x), _TINY)
    sgn = math.copysign(1.0, K2x) if K2x != 0 else 1.0
    if z > 50.0:
        ratio_K3K2 = K3x / denominator * sgn
        tau = (Ds / Tm) * z * ratio_K3K2
        return min(tau / fmGeV, 1e20)
    K1x = besselkx(1, z)
    K5x = besselkx(5, z)
    numerator = 2 * K1x - 3 * K3x + K5x
    ratio = numerator / denominator * sgn
    z3 = z * z * z / Tm
    if not math.isfinite(z3) or z3 > 1e50:
        z3 = 1e50
    tau = (Ds / 48.0) * z3 * ratio
    if not math.isfinite(tau) or abs(tau) > 1e50:
        return 1e50 * (math.copysign(1.0, tau) if tau == tau else float("nan"))
    return tau / fmGeV


def tau_drag(T, m, DsT) -> float:
    """transport.jl:132-137: the drag time m*DsT/T^2 [GeV^-1] in fm."""
    Tm, Mm, Ds = float(T), float(m), float(DsT)
    if not (math.isfinite(Tm) and math.isfinite(Mm) and math.isfinite(Ds)) or Tm <= 0 or Mm <= 0 or Ds <= 0:
        return 0.0
    return Mm * Ds / (Tm * Tm) / fmGeV


def _tau_drag_vec(T, m, DsT):
    T = np.asarray(T, dtype=np.float64)
    Ds = np.broadcast_to(np.asarray(DsT, dtype=np.float64), T.shape)
    ok = np.isfinite(T) & np.isfinite(Ds) & (T > 0) & (Ds > 0) & math.isfinite(m) & (m > 0)
    with np.errstate(all="ignore"):
        v = (m * Ds) / (T * T) / fmGeV
    return np.where(ok, v, 0.0)


def effective_DsT(T, DsT, *, DsT_linear=False, DsT_slope=1.765, DsT_offset=-0.159, Tfo=0.156,
                  DsT_quad=False, DsT_Tref=0.0):
    """transport.jl:139-156"""
    if DsT_quad:
        if DsT_linear:
            raise ValueError("effective_DsT: DsT_quad and DsT_linear are mutually exclusive")
        if not DsT_Tref > 0.0:
            raise ValueError("effective_DsT: DsT_quad requires DsT_Tref > 0")
        x = np.asarray(T, dtype=np.float64) / float(DsT_Tref)
        return float(DsT) * x * x
    if not DsT_linear:
        return np.full(np.shape(T), float(DsT)) if np.ndim(T) else float(DsT)
    Tf = np.maximum(np.asarray(T, dtype=np.float64), float(Tfo))
    return float(DsT_slope) * Tf + float(DsT_offset)


def _build_time_spline(timefn, m, DsT, *, Tmin, Tmax, nT=400, DsT_linear=False, DsT_slope=1.765,
                       DsT_offset=-0.159, Tfo=0.156, DsT_quad=False, DsT_Tref=0.0):
    """transport.jl:165-196 -> (Tmin, invdT, vals)"""
    Tmin_f, Tmax_f, n = float(Tmin), float(Tmax), int(nT)
    if n < 2:
        raise ValueError("nT must be >= 2")
    if not Tmin_f > 0.0:
        raise ValueError("Tmin must be > 0")
    if not Tmax_f > Tmin_f:
        raise ValueError("Tmax must be > Tmin")
    dT = (Tmax_f - Tmin_f) / (n - 1)
    invdT = 1.0 / dT
    Ti = Tmin_f + np.arange(n, dtype=np.float64) * dT
    DsT_eff = np.broadcast_to(effective_DsT(Ti, DsT, DsT_linear=DsT_linear, DsT_slope=DsT_slope,
                                            DsT_offset=DsT_offset, Tfo=Tfo, DsT_quad=DsT_quad,
                                            DsT_Tref=DsT_Tref), Ti.shape)
    if timefn is tau_drag:
        vals = _tau_drag_vec(Ti, float(m), DsT_eff)
    else:
        vals = np.array([timefn(t, m, d) for t, d in zip(Ti, DsT_eff)], dtype=np.float64)
    return Tmin_f, invdT, vals * LV_TAUN_SCALE[0]


def build_tau_drag_spline(m, DsT, **kw):
    """transport.jl:208-209: table of the DRAG time tau_drag (fm)."""
    return _build_time_spline(tau_drag, m, DsT, **kw)


def build_taun_current_spline(m, DsT, **kw):
    """transport.jl:221-222: table of the diffusion-current time tau_n_main3 (fm)."""
    return _build_time_spline(tau_n_main3, m, DsT, **kw)


def build_relaxation_spline(tau_of_T, *, Tmin, Tmax, nT=1024):
    """The same uniform-grid table for an arbitrary relaxation-time law tau(T) [fm], e.g. the
    tau_R = 5 (eta/s)/T of an RTA medium (extension; transport.jl:186-195 layout)."""
    Tmin_f, Tmax_f, n = float(Tmin), float(Tmax), int(nT)
    if n < 2 or not Tmin_f > 0.0 or not Tmax_f > Tmin_f:
        raise ValueError("need nT >= 2 and 0 < Tmin < Tmax")
    dT = (Tmax_f - Tmin_f) / (n - 1)
    Ti = Tmin_f + np.arange(n, dtype=np.float64) * dT
    return Tmin_f, 1.0 / dT, np.asarray(tau_of_T(Ti), dtype=np.float64)


def eval_tau_n_spline(T, Tmin, invdT, tau_vals) -> float:
    """transport.jl:231-253 (host scalar version)."""
    n = len(tau_vals)
    if n < 2:
        return float(tau_vals[0])
    x = float(T)
    if not math.isfinite(x):
        return 0.0
    u = (x - Tmin) * invdT
    i = int(math.floor(u)) + 1
    i = min(max(i, 1), n - 1)
    t = u - (i - 1)
    return (1.0 - t) * float(tau_vals[i - 1]) + t * float(tau_vals[i])


def build_juttner_invcdf(m, dimensions, *, Tmin, Tmax, nT=256, nU=2048, np_=8192,
                         quantile_cap=1.0e-6):
    """transport.jl:277-365 -> (invCDF flat (nT*nU,), nU, nT, Ttab_min, Ttab_invdT).

    invCDF[jT*nU + iu] = p*(u_iu; T_jT): the cumulative-trapezoid CDF of p^(d-1) exp(-(E-m)/T) on
    a fine grid capped at pmax = 16 sqrt(mT) + 16 T, cut at the 1 - quantile_cap quantile, and
    inverted onto the uniform u grid (the reference's two-pointer sweep, as a sorted search)."""
    M = float(m)
    d = int(dimensions)
    a = d - 1
    nTi, nUi, npi = int(nT), int(nU), int(np_)
    if nTi < 2 or nUi < 2 or npi < 2:
        raise ValueError("nT, nU, np must be >= 2")
    Tmin_f = max(float(Tmin), 1e-4)
    Tmax_f = max(float(Tmax), Tmin_f * (1.0 + 1e-6))
    span = Tmax_f - Tmin_f
    Tmin_f = max(Tmin_f - 0.02 * span, 1e-4)
    Tmax_f = Tmax_f + 0.02 * span
    dT = (Tmax_f - Tmin_f) / (nTi - 1)
    invdT = 1.0 / dT
    out = np.empty(nUi * nTi, dtype=np.float64)
    k = np.arange(npi, dtype=np.float64)
    u = np.arange(nUi, dtype=np.float64) / float(nUi - 1)
    thr = 1.0 - float(quantile_cap)
    for jT in range(nTi):
        T = Tmin_f + jT * dT
        pmax = 16.0 * math.sqrt(M * T) + 16.0 * T
        dp = pmax / (npi - 1)
        p = k * dp
        E = np.sqrt(M * M + p * p)
        w = np.exp(-(E - M) / T)
        if a == 1:
            w = p * w
        elif a == 2:
            w = (p * p) * w
        w[0] = 1.0 if a == 0 else 0.0
        inc = 0.5 * (w[:-1] + w[1:]) * dp
        cdf = np.empty(npi)
        cdf[0] = 0.0
        np.cumsum(inc, out=cdf[1:])
        norm = cdf[-1]
        invnorm = 1.0 / norm if norm > 0.0 else 0.0
        cdf *= invnorm
        cdf[-1] = 1.0
        p_hi = p[int(np.argmax(cdf >= thr))]
        kk = np.searchsorted(cdf[1:], u, side="left")
        kk = np.minimum(kk, npi - 2)
        c0, c1 = cdf[kk], cdf[kk + 1]
        p0, p1 = p[kk], p[kk + 1]
        with np.errstate(all="ignore"):
            frac = np.where(c1 > c0, (u - c0) / (c1 - c0), 0.0)
        frac = np.clip(frac, 0.0, 1.0)
        pval = p0 + frac * (p1 - p0)
        out[jT * nUi:(jT + 1) * nUi] = np.where(pval < p_hi, pval, p_hi)
    return out, nUi, nTi, Tmin_f, invdT
