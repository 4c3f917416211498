"""Host-side logic (no GPU): the C-ABI library loads and exports every declared entry point,
the host-built tables agree with the oracle's restatement of transport.jl, and the drivers'
contracts (step counts, momentum dimensions, collision modes) behave as the reference's."""
import ctypes
import math
import os

import numpy as np
import pytest

import _oracle as O
import lmhip
from lmhip import _ffi, media, transport, utils

L = O.lib()
EPS = np.finfo(float).eps


def test_library_exports_every_header_symbol():
    names = _ffi.header_symbols()
    assert len(names) >= 15
    lib = ctypes.CDLL(_ffi.lib_path())
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_ffi.SIGNATURES), set(names) ^ set(_ffi.SIGNATURES)


def test_library_is_gfx950_code_object():
    data = open(_ffi.lib_path(), "rb").read()
    assert b"gfx950" in data


def test_params_struct_layout():
    # 6 doubles, 2 uint64, 10 int32
    assert ctypes.sizeof(_ffi.lmh_params) == 6 * 8 + 2 * 8 + 10 * 4
    assert ctypes.sizeof(_ffi.lmh_obs_spec) == 2 * 4 + 3 * 8


@pytest.mark.parametrize("dims,pdim,coll,diff,want", [
    (3, 3, _ffi.LMH_LANGEVIN, False, 3), (2, 3, _ffi.LMH_RTA, False, 5),
    (2, 3, _ffi.LMH_LANGEVIN_RTA, True, 3 + 5 + 2), (1, 1, _ffi.LMH_FLUID, True, 1),
    (2, 2, _ffi.LMH_FLUID, False, 0)])
def test_noise_channels(dims, pdim, coll, diff, want):
    p = _ffi.make_params(dt=0.01, collision=coll, position_diffusion=diff)
    assert _ffi.lib().lmh_noise_channels(ctypes.byref(p), dims, pdim) == want
    assert L.lmo_noise_channels(ctypes.byref(p), dims, pdim) == want
    p.coords = _ffi.LMH_GENERAL_COORDS
    assert _ffi.lib().lmh_noise_channels(ctypes.byref(p), 2, 2) == 1


def test_context_creation_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(lmhip.LMHError):
        lmhip.Context(0)


# ---------------------------------------------------------------- tables vs the oracle ----------
def oracle_table(which, m, DsT, Tmin, Tmax, nT, **kw):
    vals = np.zeros(nT)
    t0, inv = ctypes.c_double(), ctypes.c_double()
    rc = L.lmo_build_time_spline(which, m, DsT, Tmin, Tmax, nT, int(kw.get("lin", 0)), 1.765, -0.159,
                                 0.156, int(kw.get("quad", 0)), kw.get("Tref", 0.0), 1.0,
                                 ctypes.byref(t0), ctypes.byref(inv), O.dp(vals))
    assert rc == 0
    return t0.value, inv.value, vals


@pytest.mark.parametrize("kw", [{}, {"DsT_linear": True}, {"DsT_quad": True, "DsT_Tref": 0.3}])
def test_drag_table_bit_exact_with_oracle(kw):
    t0, inv, v = transport.build_tau_drag_spline(1.5, 0.11634, Tmin=0.08, Tmax=0.55, nT=1024, **kw)
    o = oracle_table(0, 1.5, 0.11634, 0.08, 0.55, 1024, lin=kw.get("DsT_linear", False),
                     quad=kw.get("DsT_quad", False), Tref=kw.get("DsT_Tref", 0.0))
    assert t0 == o[0] and inv == o[1]
    assert np.array_equal(v, o[2])


def test_taun_table_vs_oracle():
    t0, inv, v = transport.build_taun_current_spline(1.5, 0.11634, Tmin=0.08, Tmax=0.55, nT=256)
    o = oracle_table(1, 1.5, 0.11634, 0.08, 0.55, 256)
    assert np.allclose(v, o[2], rtol=1e-12)


def test_tau_n_main3_host_vs_oracle():
    for T in (0.01, 0.03, 0.1, 0.156, 0.3, 0.6):
        for m in (1.5, 4.75):
            assert transport.tau_n_main3(T, m, 0.2) == pytest.approx(L.lmo_tau_n_main3(T, m, 0.2), rel=1e-12)
    assert transport.tau_n_main3(0.0, 1.5, 0.2) == 0.0


def test_juttner_builder_vs_oracle_and_loop():
    for d in (1, 2, 3):
        inv, nU, nT, Tt, idt = transport.build_juttner_invcdf(1.5, d, Tmin=0.15, Tmax=0.45, nT=32, nU=512, np_=2048)
        oi = np.zeros(nU * nT)
        t0, id_ = ctypes.c_double(), ctypes.c_double()
        assert L.lmo_build_juttner_invcdf(1.5, d, 0.15, 0.45, nT, nU, 2048, 1e-6, O.dp(oi), ctypes.byref(t0), ctypes.byref(id_)) == 0
        assert Tt == t0.value and idt == id_.value
        assert np.allclose(inv, oi, rtol=1e-12, atol=1e-14)
    # a literal restatement of the two-pointer sweep (transport.jl:347-361) on a tiny table
    m, d, nT, nU, npts = 1.5, 3, 3, 7, 33
    inv, *_ = transport.build_juttner_invcdf(m, d, Tmin=0.2, Tmax=0.4, nT=nT, nU=nU, np_=npts)
    Tmin_f = max(0.2, 1e-4)
    Tmax_f = max(0.4, Tmin_f * (1 + 1e-6))
    span = Tmax_f - Tmin_f
    Tmin_f, Tmax_f = max(Tmin_f - 0.02 * span, 1e-4), Tmax_f + 0.02 * span
    dT = (Tmax_f - Tmin_f) / (nT - 1)
    for jT in range(nT):
        T = Tmin_f + jT * dT
        pmax = 16.0 * math.sqrt(m * T) + 16.0 * T
        dp = pmax / (npts - 1)
        pg = [k * dp for k in range(npts)]
        cdf = [0.0]
        wprev = 0.0
        for k in range(1, npts):
            E = math.sqrt(m * m + pg[k] * pg[k])
            w = pg[k] ** 2 * math.exp(-(E - m) / T)
            cdf.append(cdf[-1] + 0.5 * (wprev + w) * dp)
            wprev = w
        nrm = cdf[-1]
        cdf = [c * (1.0 / nrm) for c in cdf]
        cdf[-1] = 1.0
        p_hi = next(pg[k] for k in range(npts) if cdf[k] >= 1 - 1e-6)
        k = 0
        for iu in range(nU):
            u = iu / (nU - 1)
            while k < npts - 1 and cdf[k + 1] < u:
                k += 1
            c0, c1 = cdf[k], cdf[k + 1]
            frac = (u - c0) / (c1 - c0) if c1 > c0 else 0.0
            frac = min(max(frac, 0.0), 1.0)
            pv = pg[k] + frac * (pg[k + 1] - pg[k])
            assert inv[jT * nU + iu] == pytest.approx(min(pv, p_hi), rel=1e-13, abs=1e-15)


def test_eval_tau_n_spline_host_vs_oracle():
    t0, inv, v = transport.build_tau_drag_spline(1.5, 0.2, Tmin=0.1, Tmax=0.5, nT=64)
    vv = np.ascontiguousarray(v)
    for T in (0.05, 0.1, 0.1234, 0.3, 0.5, 0.7, float("nan")):
        assert transport.eval_tau_n_spline(T, t0, inv, v) == L.lmo_eval_time_table(T, t0, inv, O.dp(vv), 64)


# ---------------------------------------------------------------- driver contracts --------------
def test_check_momentum_dims_reference_gate():
    # test_momentum_dims3.jl:185-187
    with pytest.raises(ValueError):
        utils.check_momentum_dims(2, 3, False, True, 0.0)
    with pytest.raises(ValueError):
        utils.check_momentum_dims(2, 4, False, False, 0.4)
    with pytest.raises(ValueError):
        utils.check_momentum_dims(3, 2, False, False, 0.4)
    with pytest.raises(ValueError):
        utils.check_momentum_dims(1, 3, True, False, 0.4)
    utils.check_momentum_dims(2, 3, False, True, 0.4)
    utils.check_momentum_dims(3, 3, False, False, 0.0)


def test_step_counts():
    assert utils.step_counts(0.0, 5.0, 0.01, 0.5) == (500, 50, 10)     # floor(5.0/0.01) = 500
    assert utils.step_counts(0.6, 10.3118, 0.005, 0.1) == (1942, 20, 97)
    assert utils.step_counts(0.0, 1.0, 0.001, 0.1) == (1000, 100, 10)
    assert utils.step_counts(0.0, 1.0, 0.3, 0.45)[1] == 2              # round(1.5) ties to even
    with pytest.raises(ZeroDivisionError):
        utils.step_counts(0.0, 1.0, 0.1, 0.01)
    with pytest.raises(ValueError):
        utils.step_counts_gc(0.0, 1.0, 0.01, 0.015)
    assert utils.step_counts_gc(0.0, 1.0, 0.01, 0.25) == (100, 25, 4)


def test_unsupported_collision_mode_errors_like_gpu_wrapper():
    xg, tg, Tf, Vf = media.uniform_bath()
    with pytest.raises(ValueError, match="not supported on GPU"):
        lmhip.simulate_ensemble_bulk(lmhip.HIPBackend(), xg, xg, None, Tf, Vf, (xg, tg),
                                     collision_mode="bgk", x_init=np.zeros((3, 2)), p_init=np.zeros((3, 2)),
                                     N_particles=2)


def test_dispatch_on_backend_type():
    class Other(lmhip.AbstractBackend):
        pass
    with pytest.raises(TypeError):
        lmhip.simulate_ensemble_bulk(Other(), 1, 2, 3)


def test_append_thermal_pz_statistics():
    # test_momentum_dims3.jl:98-117 (S): the p_z draw is the thermal conditional
    from scipy.integrate import quad
    rng = np.random.default_rng(7)
    mT, T = 1.6, 0.3
    s = np.array([utils.sample_pz_conditional_juttner(mT, T, rng) for _ in range(40_000)])
    num = quad(lambda z: z * z * math.exp(-(math.sqrt(mT * mT + z * z) - mT) / T), -np.inf, np.inf)[0]
    den = quad(lambda z: math.exp(-(math.sqrt(mT * mT + z * z) - mT) / T), -np.inf, np.inf)[0]
    assert np.mean(s ** 2) == pytest.approx(num / den, rel=0.03)
    pz = utils.append_thermal_pz(np.zeros((2, 100)), np.zeros((2, 100)), 1.5, lambda r: 0.3,
                                 antithetic=True, rng=rng)
    assert np.array_equal(pz[2, 1::2], -pz[2, 0::2])


def test_media_builders():
    g = media.hydro_grid(nx=20, ny=20, nt=6)
    assert g.T.shape == (20, 20, 6) and g.T.flags.f_contiguous
    assert np.all(g.vx[:, :, 0] == 0)                              # no flow at tau0
    assert g.T.max() == pytest.approx(0.45 * math.exp(-(g.x0 + g.dx * 9.5) ** 2 * 0), rel=0.1)
    pt = media.power_law_pt(200_000, np.random.default_rng(0))
    assert pt.max() <= 20.0
    # mean p_T of dN/dp ~ p/(1+p^2/4)^3 on [0, 20]
    from scipy.integrate import quad
    f = lambda p: p / (1 + p * p / 4) ** 3
    mean = quad(lambda p: p * f(p), 0, 20)[0] / quad(f, 0, 20)[0]
    assert pt.mean() == pytest.approx(mean, rel=0.01)
