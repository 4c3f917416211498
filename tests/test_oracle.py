"""Pins of the CPU oracle (oracle/lm_oracle.c), before it is trusted as the checker.

Each check comes from one of: the reference's own tests (runtests.jl, test_momentum_dims3.jl,
test_relativistic_switch.jl), a closed form computed independently (mpmath at 50 digits), a
library routine (numpy's Philox, scipy's Bessel K) or an invariant, chosen so that a dropped
term, a wrong sign or a swapped index fails one of them.
"""
import ctypes
import math

import mpmath as mp
import numpy as np
import pytest
from scipy.special import kve

import _oracle as O
import lmhip
from lmhip import _ffi, media

mp.mp.dps = 50
L = O.lib()
FMGEV = 1 / 0.197327
EPS = np.finfo(float).eps


def interp(x, y, V, xi, yi, gc=False):
    x, y = np.ascontiguousarray(x, float), np.ascontiguousarray(y, float)
    V = np.asfortranarray(V, float)
    f = L.lmo_interpolate_2d_gc if gc else L.lmo_interpolate_2d
    return f(O.dp(x), x.size, O.dp(y), y.size, O.dp(V), xi, yi)


# ---------------------------------------------------------------- Philox4x32-10 ----------------
# Random123's known-answer vectors for philox4x32-10 (counter, key -> output)
PHILOX32_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX32_KAT)
def test_philox4x32_known_answers(ctr, key, want):
    out = (ctypes.c_uint32 * 4)()
    L.lmo_philox4x32_10((ctypes.c_uint32 * 4)(*ctr), (ctypes.c_uint32 * 2)(*key), out)
    assert tuple(out) == want


def _philox_words(seed, stream, i, step, block):
    """the four words of a noise block from the KAT-pinned generator and the documented counter
    layout: (i, step, stream_lo, stream_hi mod 2^29 | (2 block + half) << 29), key seed"""
    w = []
    for h in range(2):
        ctr = (i & 0xFFFFFFFF, step & 0xFFFFFFFF, stream & 0xFFFFFFFF,
               ((stream >> 32) & 0x1FFFFFFF) | ((2 * block + h) << 29))
        out = (ctypes.c_uint32 * 4)()
        L.lmo_philox4x32_10((ctypes.c_uint32 * 4)(*ctr),
                            (ctypes.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32), out)
        w += [out[0] | (out[1] << 32), out[2] | (out[3] << 32)]
    return w


def test_noise_block_layout():
    for seed, stream, i, step, block in ((0, 0, 0, 0, 0), (99, 3, 7, 17, 3), (2**64 - 1, 2**61 - 1, 2**32 - 1, 2**32 - 1, 2)):
        w = (ctypes.c_uint64 * 4)()
        L.lmo_noise_block(seed, stream, i, step, block, w)
        assert list(w) == _philox_words(seed, stream, i, step, block)
    # distinct blocks, halves, streams and steps give distinct words
    seen = set()
    for block in (0, 2, 3):
        for stream in (0, 1, 2**32, 2**60):
            for step in (1, 2):
                w = (ctypes.c_uint64 * 4)()
                L.lmo_noise_block(5, stream, 11, step, block, w)
                seen.update(list(w))
    assert len(seen) == 3 * 4 * 4 * 2


def test_noise_record_layout_vs_mpmath():
    """each record channel at 50 digits from the Philox words: the kick normals as the 4-D
    construction of include/lmhip.h (radius^2 = -2 ln(u1 u2), the split v, the angles beta and
    gamma), the collision uniform, the redraw's quantile and Gaussian, the diffusion pair"""
    p = _ffi.make_params(dt=0.01, seed=99, stream=3, noise=_ffi.LMH_NOISE_PHILOX,
                         collision=_ffi.LMH_LANGEVIN_RTA, position_diffusion=True)
    rec = O.noise_record(p, 2, 3, 5, 17, 1)[0]
    assert rec.shape == (3 + 5 + 2, 5)
    two = mp.mpf(2)
    # uniforms: d = 1.m from the top 52 bits, u = d - 1 or 1 - u = d - 1 (so u1 = 2 - d)
    mant = lambda v: 1 + mp.mpf((v >> 12) & (2**52 - 1)) * two ** -52
    u1 = lambda v: 2 - mant(v)
    lo = lambda v: v & 0xFFFFFFFF
    hi = lambda v: v >> 32
    # the diffusion pair's angle: pi (q/2 + d/2 - 3/4), q = top 2 bits, d = 1.m from the next 52 bits
    ang = lambda v: mp.pi * (mp.mpf(v >> 62) / 2 + mant((v << 2) & (2**64 - 1)) / 2 - mp.mpf(0.75))
    for i in range(5):
        w = _philox_words(99, 3, i, 17, 0)
        R2 = -2 * mp.log(u1(w[0]) * (mp.mpf(lo(w[1])) + 1) * two ** -32)
        v = mp.mpf(lo(w[2])) * two ** -32
        beta = mp.pi * (mp.mpf(hi(w[1]) >> 30) / 2 + mp.mpf(hi(w[1]) & 0x3FFFFFFF) * two ** -31 - mp.mpf(0.25))
        gcos = mp.cos(mp.pi * mp.mpf(hi(w[2]) & 0x7FFFFFFF) * two ** -32) * (-1 if hi(w[2]) >> 31 else 1)
        want = (mp.sqrt(v * R2) * mp.cos(beta), mp.sqrt(v * R2) * mp.sin(beta), mp.sqrt((1 - v) * R2) * gcos)
        for c in range(3):
            assert abs(rec[c, i] - float(want[c])) <= 4e-15 * max(1, abs(float(want[c]))), (i, c)
        # the collision uniform is the low 31 bits of w3, the quantile its high half
        assert rec[3, i] == (w[3] & 0x7FFFFFFF) * 2.0 ** -31
        assert rec[4, i] == hi(w[3]) * 2.0 ** -32
        # the redraw's Gaussian: the three kick normals themselves
        assert np.array_equal(rec[5:8, i], rec[0:3, i])
        w2 = _philox_words(99, 3, i, 17, 2)
        x0 = mp.sqrt(-2 * mp.log(u1(w2[0]))) * mp.cos(ang(w2[1]))
        assert abs(rec[8, i] - float(x0)) <= 4e-15 * max(1, abs(float(x0)))
    # the general-coordinate step's normal: the first of the Box-Muller pair of words 0 and 1
    pg = _ffi.make_params(dt=0.01, seed=99, stream=3, noise=_ffi.LMH_NOISE_PHILOX, coords=_ffi.LMH_GENERAL_COORDS)
    recg = O.noise_record(pg, 1, 1, 5, 17, 1)[0]
    for i in range(5):
        w = _philox_words(99, 3, i, 17, 0)
        want = mp.sqrt(-2 * mp.log(u1(w[0]))) * mp.cos(ang(w[1]))
        assert abs(recg[0, i] - float(want)) <= 4e-15 * max(1, abs(float(want)))


def test_kick_normals_jointly_gaussian():
    """the three kick normals are jointly standard normal: |n|^2 ~ chi^2(3), n0^2 + n1^2 ~ chi^2(2),
    n2^2 ~ chi^2(1), and n0 n1, n0 n2, n1 n2, n0^2 n2^2 carry no correlation (4e5 draws; KS
    p > 1e-4, 5 sigma): a radius, split or angle taken from the wrong bits, or v and 1 - v
    swapped into one component, fails here even where each channel alone looks normal"""
    from scipy import stats
    p = _ffi.make_params(dt=0.01, seed=21, stream=0, noise=_ffi.LMH_NOISE_PHILOX, collision=_ffi.LMH_LANGEVIN)
    rec = O.noise_record(p, 2, 3, 200_000, 1, 2)
    n = rec[:, 0:3, :].transpose(1, 0, 2).reshape(3, -1)
    N = n.shape[1]
    assert stats.kstest(np.sum(n ** 2, axis=0), "chi2", args=(3,)).pvalue > 1e-4
    assert stats.kstest(n[0] ** 2 + n[1] ** 2, "chi2", args=(2,)).pvalue > 1e-4
    assert stats.kstest(n[2] ** 2, "chi2", args=(1,)).pvalue > 1e-4
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(np.mean(n[a] * n[b])) < 5 / np.sqrt(N)
    # fourth moments: E[n_a^2 n_b^2] = 1 for independent standard normals (3 for a = b)
    for a, b in ((0, 2), (1, 2), (0, 1)):
        assert abs(np.mean(n[a] ** 2 * n[b] ** 2) - 1) < 5 * np.sqrt(8 / N)
    assert abs(np.mean(n[2] ** 4) - 3) < 5 * np.sqrt(96 / N)


def test_noise_record_statistics():
    """the record's normals are standard normal, its uniforms uniform and the redraw direction
    (the normalised kick normals) isotropic (seeded, 4e5 draws per channel; 5 sigma on the
    moments, Kolmogorov-Smirnov p > 1e-4), and distinct channels are uncorrelated: a wrong word,
    shift or counter in the generator shows up here"""
    from scipy import stats
    p = _ffi.make_params(dt=0.01, seed=7, stream=2, noise=_ffi.LMH_NOISE_PHILOX,
                         collision=_ffi.LMH_LANGEVIN_RTA, position_diffusion=True)
    rec = O.noise_record(p, 2, 3, 200000, 3, 2)
    normals, uniforms, direction = (0, 1, 2, 8, 9), (3, 4), (5, 6, 7)
    assert np.array_equal(rec[:, direction, :], rec[:, 0:3, :])
    # the normalised Gaussian: z = cos(theta) uniform on [-1, 1], the azimuth uniform
    g = rec[:, direction, :] / np.sqrt(np.sum(rec[:, direction, :] ** 2, axis=1, keepdims=True))
    n = g[:, 0, :].size
    for k in range(3):
        x = g[:, k, :].ravel()
        assert abs(x.mean()) < 5 * np.sqrt(1 / (3 * n))
        assert abs(x.var() - 1 / 3) < 5 * np.sqrt(4 / (45 * n))
    assert stats.kstest(g[:, 2, :].ravel(), "uniform", args=(-1.0, 2.0)).pvalue > 1e-4
    phi = np.arctan2(g[:, 1, :], g[:, 0, :]).ravel()
    assert stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > 1e-4
    # the quantile is independent of the normals drawn from the same words' high bits
    for c in normals[:3]:
        assert abs(np.corrcoef(rec[:, 4, :].ravel(), rec[:, c, :].ravel() ** 2)[0, 1]) < 5 / np.sqrt(n)
    for c in normals:
        x = rec[:, c, :].ravel()
        n = x.size
        assert abs(x.mean()) < 5 * np.sqrt(1 / n)
        assert abs(x.var() - 1) < 5 * np.sqrt(2 / n)
        assert stats.kstest(x, "norm").pvalue > 1e-4
    for c in uniforms:
        x = rec[:, c, :].ravel()
        assert 0.0 <= x.min() and x.max() < 1.0
        assert stats.kstest(x, "uniform").pvalue > 1e-4
    flat = rec[:, [0, 1, 2, 3, 4, 8, 9], :].transpose(1, 0, 2).reshape(7, -1)
    cc = np.corrcoef(flat)
    off = cc[~np.eye(cc.shape[0], dtype=bool)]
    assert np.max(np.abs(off)) < 5 / np.sqrt(flat.shape[1])


def test_noise_transforms_accuracy():
    # the Box-Muller log, sincospi and the half-range cospi are built from correctly rounded
    # operations only
    f = L.lmo_noise_transforms
    f.argtypes = [ctypes.c_double, ctypes.c_double, O.PD]
    f.restype = None
    out = np.zeros(4)
    rng = np.random.default_rng(0)
    us = list(rng.random(4000)) + [2.0 ** -53, 1.0, 0.5, 0.7071067811865476, 0.7071067811865475, 1 - 2.0 ** -53]
    xs = list(2 * rng.random(len(us) - 6)) + [0.0, 0.25, 0.5, 1.0, 1.5, 1.75]
    for u, x in zip(us, xs):
        f(u, x, O.dp(out))
        lu = mp.log(u)
        assert abs(out[0] - float(lu)) <= 4e-16 * max(abs(float(lu)), 1e-300) or (u == 1.0 and out[0] == 0.0)
        assert abs(out[1] - float(mp.sinpi(x))) <= 2.3e-16
        assert abs(out[2] - float(mp.cospi(x))) <= 2.3e-16
        # cos(pi x) on [0, 1/2] by one polynomial in x^2: within 4e-16 (the cancellation of its
        # first terms near x = 1/2 costs an ulp or two)
        assert abs(out[3] - float(mp.cospi(mp.mpf(0.25 * x)))) <= 4e-16


def test_noise_record_moments():
    p = _ffi.make_params(dt=0.01, seed=5, noise=_ffi.LMH_NOISE_PHILOX, collision=_ffi.LMH_LANGEVIN_RTA)
    rec = O.noise_record(p, 2, 3, 200_000, 1, 1)[0]
    n = rec.shape[1]
    for c in (0, 1, 2):                                  # the kick normals
        assert abs(rec[c].mean()) < 5 / math.sqrt(n)
        assert abs(rec[c].var() - 1) < 6 * math.sqrt(2 / n)
    assert np.array_equal(rec[5:8], rec[0:3])            # the redraw's Gaussian: the kick normals
    for c in (3, 4):                                     # the collision and quantile uniforms
        assert abs(rec[c].mean() - 0.5) < 5 * math.sqrt(1 / 12 / n)
        assert abs(rec[c].var() - 1 / 12) < 0.003
        assert abs(np.corrcoef(rec[c], rec[0])[0, 1]) < 0.01
    assert abs(np.corrcoef(rec[0], rec[1])[0, 1]) < 0.01
    # u_col is independent of the normals drawn from the same block's high bits
    assert abs(np.corrcoef(rec[3], rec[0] ** 2)[0, 1]) < 0.01


# ---------------------------------------------------------------- interpolation ------------------
def test_interpolation_nodes_affine_and_clamp():
    x = np.array([0.0, 0.3, 1.1, 2.0, 5.0])
    y = np.array([0.0, 1.0, 2.5])
    V = np.add.outer(2.0 + 3.0 * x, -0.5 * y)               # affine in (x, y)
    for i, xi in enumerate(x):
        for j, yj in enumerate(y):
            assert interp(x, y, V, xi, yj) == V[i, j]
    rng = np.random.default_rng(0)
    for _ in range(200):
        xi, yi = rng.uniform(0, 5), rng.uniform(0, 2.5)
        assert interp(x, y, V, xi, yi) == pytest.approx(2 + 3 * xi - 0.5 * yi, rel=1e-14, abs=1e-14)
    # the query is clamped to the grid (kernels_cpu.jl:27-28): no extrapolation
    assert interp(x, y, V, 9.0, 1.0) == pytest.approx(2 + 15 - 0.5)
    assert interp(x, y, V, -3.0, 7.0) == pytest.approx(2 - 1.25)
    # GC variant extrapolates the edge cell (kernels_cpu_GC.jl:37-57)
    assert interp(x, y, V, 9.0, 1.0, gc=True) == pytest.approx(2 + 27 - 0.5, rel=1e-13)
    assert interp(x, y, V, -1.0, 3.5, gc=True) == pytest.approx(2 - 3 - 1.75, rel=1e-13)


def test_interpolation_eps_guard_duplicate_nodes():
    x = np.array([0.0, 1.0, 1.0, 2.0])      # a repeated node: dx = 0 in the cell [1, 1]
    y = np.array([0.0, 1.0])
    V = np.array([[0.0, 0.0], [1.0, 1.0], [5.0, 5.0], [6.0, 6.0]])
    # searchsortedlast picks the last of the duplicates, so x = 1 lands in cell [1, 2]
    assert interp(x, y, V, 1.0, 0.5) == 5.0
    assert L.lmo_searchsortedlast(O.dp(np.ascontiguousarray(x)), 4, 1.0) == 3


def test_searchsortedlast():
    v = np.ascontiguousarray([1.0, 2.0, 2.0, 3.0])
    for xq, want in [(0.5, 0), (1.0, 1), (1.5, 1), (2.0, 3), (2.5, 3), (3.0, 4), (9.0, 4)]:
        assert L.lmo_searchsortedlast(O.dp(v), 4, xq) == want


# ---------------------------------------------------------------- transport ----------------------
def test_tau_drag_reference_gate():
    # runtests.jl:83-93
    for T in (0.156, 0.20, 0.30, 0.40):
        for dst in (0.116, 0.371):
            td = L.lmo_tau_drag(T, 1.5, dst)
            assert td == pytest.approx(1.5 * dst / T**2 / FMGEV, rel=1e-12)
            assert (T / (1.5 * (1 / td))) * T * FMGEV == pytest.approx(dst, rel=1e-12)
    assert L.lmo_tau_drag(-1.0, 1.5, 0.1) == 0.0
    assert L.lmo_tau_drag(0.3, 1.5, 0.0) == 0.0


def test_besselkx_vs_scipy_and_identity():
    for nu in (1, 2, 3, 5):
        for z in (0.5, 1.0, 2.0, 5.0, 9.6, 20.0, 40.0, 49.99, 50.01, 300.0, 1.5e4, 1.5e6):
            assert L.lmo_besselkx(nu, z) == pytest.approx(kve(nu, z), rel=2e-13)
    # runtests.jl:36-40
    for z in (0.5, 1.0, 2.0, 5.0, 9.6, 20.0, 40.0):
        K = lambda n: L.lmo_besselkx(n, z)
        assert 2 * K(1) - 3 * K(3) + K(5) == pytest.approx(48 / z**2 * K(3), rel=1e-9)


def test_tau_n_main3_reference_gates():
    M, DST = 1.5, 0.11634
    ref = lambda T: (DST / T) * (M / T) * (kve(3, M / T) / kve(2, M / T)) / FMGEV
    for T in (0.6, 0.5, 0.4, 0.3, 0.25, 0.2, 0.156, 0.12, 0.1):       # runtests.jl:42-51
        assert L.lmo_tau_n_main3(T, M, DST) == pytest.approx(ref(T), rel=1e-8)
        assert L.lmo_tau_n_main3(T, M, DST) > 0
    assert L.lmo_tau_n_main3(-1.0, M, DST) == 0.0
    assert L.lmo_tau_n_main3(0.3, M, 0.0) == 0.0
    assert L.lmo_tau_n_main3(0.3, M, -1.0) == 0.0
    Tb, Ta = M / 49.99, M / 50.01                                       # :53-59
    assert L.lmo_tau_n_main3(Tb, M, DST) == pytest.approx(ref(Tb), rel=1e-8)
    assert L.lmo_tau_n_main3(Ta, M, DST) == pytest.approx(ref(Ta), rel=1e-6)
    assert L.lmo_tau_n_main3(Tb, M, DST) == pytest.approx(L.lmo_tau_n_main3(Ta, M, DST), rel=1e-3)
    ratio = lambda T: L.lmo_tau_n_main3(T, M, DST) * T**2 * FMGEV / (DST * M)   # :61-66
    assert ratio(0.05) > 1.0 and ratio(0.05) < ratio(0.20)
    assert abs(ratio(0.02) - 1.0) < 0.05
    for T in (0.156, 0.2, 0.3, 0.4):                                    # :92
        z = M / T
        assert L.lmo_tau_n_main3(T, M, DST) / L.lmo_tau_drag(T, M, DST) == pytest.approx(kve(3, z) / kve(2, z), rel=1e-12)


def test_effective_DsT_reference_gate():
    err = ctypes.c_int32(0)
    f = lambda T, lin, quad=0, Tref=0.0: L.lmo_effective_DsT(T, 0.11634, lin, 1.765, -0.159, 0.156, quad, Tref, ctypes.byref(err))
    assert f(0.3, 0) == 0.11634 and f(0.6, 0) == 0.11634                 # runtests.jl:68-77
    assert f(0.30, 1) == pytest.approx(1.765 * 0.30 - 0.159, rel=1e-12)
    assert f(0.10, 1) == pytest.approx(1.765 * 0.156 - 0.159, rel=1e-12)
    assert f(0.2, 0, 1, 0.4) == pytest.approx(0.11634 * 0.25, rel=1e-15)
    f(0.2, 1, 1, 0.4)
    assert err.value == 1                                               # quad and linear exclusive


def build_table(which, m, DsT, Tmin, Tmax, nT, linear=0, quad=0, Tref=0.0, scale=1.0):
    vals = np.zeros(nT)
    t0, inv = ctypes.c_double(), ctypes.c_double()
    rc = L.lmo_build_time_spline(which, m, DsT, Tmin, Tmax, nT, linear, 1.765, -0.159, 0.156, quad,
                                 Tref, scale, ctypes.byref(t0), ctypes.byref(inv), O.dp(vals))
    return rc, t0.value, inv.value, vals


def test_time_tables_and_eval():
    rc, t0, inv, d = build_table(0, 1.5, 0.11634, 0.10, 0.50, 128)
    rc2, _, _, n = build_table(1, 1.5, 0.11634, 0.10, 0.50, 128)
    assert rc == 0 and rc2 == 0
    dT = (0.5 - 0.1) / 127
    for i in range(128):                                                # runtests.jl:96-103
        T = 0.1 + i * dT
        z = 1.5 / T
        assert d[i] == 1.5 * 0.11634 / (T * T) / FMGEV
        assert n[i] / d[i] == pytest.approx(kve(3, z) / kve(2, z), rel=1e-12)
    ev = lambda T: L.lmo_eval_time_table(T, t0, inv, O.dp(d), 128)
    for i in (0, 5, 127):
        assert ev(0.1 + i * dT) == pytest.approx(d[i], rel=1e-13)
    mid = ev(0.1 + 5.5 * dT)
    assert mid == pytest.approx(0.5 * (d[5] + d[6]), rel=1e-12)
    # no clamp on the fraction: linear extrapolation beyond both ends (transport.jl:245-252)
    assert ev(0.1 - dT) == pytest.approx(2 * d[0] - d[1], rel=1e-12)
    assert ev(0.5 + 2 * dT) == pytest.approx(d[127] + 2 * (d[127] - d[126]), rel=1e-12)
    assert ev(float("nan")) == 0.0 and ev(float("inf")) == 0.0
    # degenerate uniform medium (Tmax = Tmin + eps): T = Tmin returns vals[0] exactly
    Tmin = 0.3
    rc, t0, inv, d = build_table(0, 1.5, 0.6366, Tmin, max(Tmin, Tmin + EPS), 1024)
    assert rc == 0 and L.lmo_eval_time_table(0.3, t0, inv, O.dp(d), 1024) == d[0]
    assert build_table(0, 1.5, 0.1, 0.0, 0.5, 16)[0] != 0              # Tmin must be > 0
    assert build_table(0, 1.5, 0.1, 0.5, 0.5, 16)[0] != 0              # Tmax must be > Tmin


# ---------------------------------------------------------------- Juttner tables -----------------
def juttner_quad(m, T, d, g):
    w = lambda p: p ** (d - 1) * mp.exp(-(mp.sqrt(p * p + m * m) - m) / T)
    return mp.quad(lambda p: g(p) * w(p), [0, 10 * T, 40 * T, mp.inf]) / mp.quad(w, [0, 10 * T, 40 * T, mp.inf])


def build_jt(m, d, Tmin, Tmax, nT=256, nU=2048, np_=8192):
    inv = np.zeros(nT * nU)
    t0, idt = ctypes.c_double(), ctypes.c_double()
    rc = L.lmo_build_juttner_invcdf(m, d, Tmin, Tmax, nT, nU, np_, 1e-6, O.dp(inv), ctypes.byref(t0), ctypes.byref(idt))
    assert rc == 0
    return inv, nU, nT, t0.value, idt.value


@pytest.mark.parametrize("d", [1, 2, 3])
def test_juttner_table_moments(d):
    m = 1.5
    inv, nU, nT, t0, idt = build_jt(m, d, 0.2, 0.4)
    u = (np.arange(20_000) + 0.5) / 20_000
    for T in (t0 + 100 / idt, 0.3):          # on a table node and between nodes
        ps = np.array([L.lmo_juttner_invcdf_lookup(O.dp(inv), nU, nT, t0, idt, x, T) for x in u])
        want = float(juttner_quad(m, T, d, lambda p: p * p))
        # linear interpolation of the quantile in u over-weights the tail bin (the reason for
        # the quantile cap, transport.jl:332-345): a few 1e-3 of bias remain by design
        assert np.mean(ps * ps) == pytest.approx(want, rel=5e-3)
    # at a table node the quantile itself matches the exact inverse CDF
    from scipy.integrate import quad
    from scipy.optimize import brentq
    T = t0 + 100 / idt
    w = lambda p: p ** (d - 1) * math.exp(-(math.sqrt(p * p + m * m) - m) / T)
    pmax = 16 * math.sqrt(m * T) + 16 * T
    Z = quad(w, 0, pmax, limit=200)[0]
    for uq in (0.1, 0.5, 0.9, 0.99):
        iu = int(round(uq * (nU - 1)))
        ugrid = iu / (nU - 1)
        exact = brentq(lambda p: quad(w, 0, p, limit=200)[0] / Z - ugrid, 0, pmax, xtol=1e-14)
        assert inv[100 * nU + iu] == pytest.approx(exact, rel=2e-5)


def test_juttner_rejection_sampler_moments():
    m, T = 1.5, 0.3
    for d in (2, 3):
        out = np.zeros(200_000)
        L.lmo_draw_juttner_rejection(m, T, d, 1234 + d, out.size, O.dp(out))
        want = float(juttner_quad(m, T, d, lambda p: p * p))
        assert np.mean(out ** 2) == pytest.approx(want, rel=0.01)


# ---------------------------------------------------------------- boosts -------------------------
def boosts(v, pos, mom, m=1.5, sign=-1, rel=1, dims=2, pdim=2, t=0.0):
    keep = O.Keep()
    xg, tg = np.linspace(0, 10, 11), np.array([0.0, 1.0])
    V = np.full((11, 2), v)
    med = O.radial_medium(keep, xg, tg, V * 0 + 0.3, V)
    P = np.asfortranarray(np.array(pos, float).reshape(dims, -1))
    M = np.asfortranarray(np.array(mom, float).reshape(pdim, -1))
    f = L.lmo_kernel_boost_to_rest if sign < 0 else L.lmo_kernel_boost_to_lab
    f(O.dp(M), O.dp(P), ctypes.byref(med), m, P.shape[1], 0, 0.01, t, dims, pdim, int(dims == 1), 0.0, rel)
    return M


def test_boost_comoving_particle_is_at_rest():
    # a particle moving with the fluid (p = m gamma v r_hat) has ~zero LRF momentum
    m, v = 1.5, 0.6
    g = 1 / math.sqrt(1 - v * v)
    rhat = np.array([0.6, 0.8])
    pos = 3.0 * rhat
    M = boosts(v, pos, m * g * v * rhat, m=m)
    assert np.linalg.norm(M[:, 0]) < 1e-9 * m
    # and the reverse boost of a particle at rest gives the fluid momentum (to the 1e-10 regulator)
    M = boosts(v, pos, [0.0, 0.0], m=m, sign=+1)
    assert np.allclose(M[:, 0], m * g * v * rhat, rtol=1e-9)


def test_boost_perp_unchanged_and_invariant_mass():
    m, v = 1.5, 0.45
    rhat = np.array([1.0, 2.0]) / math.sqrt(5.0)
    perp = np.array([-rhat[1], rhat[0]])
    p = 0.7 * rhat + 1.3 * perp
    M = boosts(v, 4.0 * rhat, p, m=m)[:, 0]
    assert M @ perp == pytest.approx(1.3, rel=1e-14)
    # exact Lorentz boost of the parallel component, up to the 1e-10 regulator
    E = math.sqrt(m * m + p @ p)
    g = 1 / math.sqrt(1 - v * v)
    assert M @ rhat == pytest.approx(g * (0.7 - v * E), rel=1e-9)
    # round trip rest -> lab restores the momentum to the regulator
    P2 = boosts(v, 4.0 * rhat, M, m=m, sign=+1)[:, 0]
    assert np.allclose(P2, p, rtol=1e-9)


def test_boost_at_rest_regulator_and_origin_skip():
    # v = 0: each boost scales p_par by 1/sqrt(1+1e-10) (kernels_cpu.jl:105): must be reproduced
    rhat = np.array([0.6, 0.8])
    p = np.array([0.3, -0.2])
    M = boosts(0.0, 2.0 * rhat, p)[:, 0]
    ppar = p @ rhat
    want = p + (ppar / math.sqrt(1 + 1e-10) - ppar) * rhat
    assert np.allclose(M, want, rtol=0, atol=1e-16)
    assert abs(M @ rhat - ppar) > 1e-12 * abs(ppar)
    # r < eps: untouched (kernels_cpu.jl:75-77)
    M = boosts(0.9, [0.0, 0.0], p)[:, 0]
    assert np.array_equal(M, p)


def test_boost_galilean_and_radial_mode():
    rhat = np.array([0.6, 0.8])
    M = boosts(0.3, 5.0 * rhat, [0.1, 0.2], m=2.0, rel=0)[:, 0]
    assert np.allclose(M, np.array([0.1, 0.2]) - 2.0 * 0.3 * rhat, rtol=0, atol=1e-16)
    M = boosts(0.3, [5.0], [0.4], m=2.0, dims=1, pdim=1)[:, 0]
    g = 1 / math.sqrt(1 - 0.09 + 1e-10)
    assert M[0] == pytest.approx(g * (0.4 - 0.3 * math.sqrt(0.16 + 4.0)), rel=1e-15)


# ---------------------------------------------------------------- full step ----------------------
def uniform_setup(keep, T=0.3, m=1.5, DsT=0.6366197723675814, V=0.0, nT=1024):
    xg = np.arange(0.0, 30.25, 0.5)
    tg = np.arange(0.0, 10.5, 1.0)
    Tf = np.full((xg.size, tg.size), T)
    Vf = np.full((xg.size, tg.size), V)
    med = O.radial_medium(keep, xg, tg, Tf, Vf)
    rc, t0, inv, d = build_table(0, m, DsT, T, T + EPS, nT)
    assert rc == 0
    return med, (t0, inv, d)


def test_ou_mean_step_closed_form():
    # xi = 0, fluid at rest: one step multiplies p by exp(-eta_D m/E dt) (the exact OU mean,
    # kernels_cpu.jl:339-345) and the two regulated boosts, then x += dt p/E.
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.6366197723675814, 0.01
    med, tab = uniform_setup(keep, T, m, DsT)
    pos = np.asfortranarray([[1.0, 2.0, 0.5], [-2.0, 0.3, 1.0]]).T.copy(order="F")
    mom = np.asfortranarray([[0.8, -0.4, 1.1], [2.0, 1.0, -3.0]]).T.copy(order="F")
    sp = O.species(keep, [m], [2], drag=[tab])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    noise = np.zeros((1, 3, 2))
    P, M = pos.copy(order="F"), mom.copy(order="F")
    rc, bad = O.run_cartesian(p, med, sp, P, M, 1, 1, noise)
    assert rc == 0
    eta = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))   # 1/tau_drag
    for i in range(2):
        x = [mp.mpf(v) for v in pos[:, i]]
        q = [mp.mpf(v) for v in mom[:, i]]
        r = mp.sqrt(sum(v * v for v in x))
        rh = [x[0] / r, x[1] / r]
        g = 1 / mp.sqrt(1 + mp.mpf("1e-10"))
        par = q[0] * rh[0] + q[1] * rh[1]
        q = [q[0] + (g * par - par) * rh[0], q[1] + (g * par - par) * rh[1], q[2]]
        E = mp.sqrt(sum(v * v for v in q) + m * m)
        a = mp.exp(-eta * m / E * dt)
        q = [a * v for v in q]
        par = q[0] * rh[0] + q[1] * rh[1]
        q = [q[0] + (g * par - par) * rh[0], q[1] + (g * par - par) * rh[1], q[2]]
        E = mp.sqrt(sum(v * v for v in q) + m * m)
        xn = [x[d] + dt * q[d] / E for d in range(3)]
        assert np.allclose(M[:, i], [float(v) for v in q], rtol=2e-15, atol=0)
        assert np.allclose(P[:, i], [float(v) for v in xn], rtol=2e-15, atol=0)


def test_ou_noise_only_step_closed_form():
    # p = 0 at the origin (no boost), one xi: p = sqrt(dt) kT sqrt((1-a^2)/(2 eta dt)) xi with
    # eta = eta_D m/m (E = m at p = 0): the exact OU variance kappa (1-a^2)/(2 eta)
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.6366197723675814, 0.01
    med, tab = uniform_setup(keep, T, m, DsT)
    sp = O.species(keep, [m], [1], drag=[tab])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    P = np.zeros((3, 1), order="F")
    M = np.zeros((3, 1), order="F")
    xi = np.array([0.7, -1.2, 0.1])
    rc, _ = O.run_cartesian(p, med, sp, P, M, 1, 1, xi.reshape(1, 3, 1))
    assert rc == 0
    eta = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))
    kappa = 2 * m * mp.mpf(T) * eta
    a = mp.exp(-eta * dt)
    sd = mp.sqrt(kappa * (1 - a * a) / (2 * eta))
    # the lab boost at v = 0 scales p_par by 1/sqrt(1+1e-10) only when r > eps; r = 0 here
    assert np.allclose(M[:, 0], [float(sd * v) for v in xi], rtol=1e-14)


def test_ou_stationary_variance_reference_gate():
    # runtests.jl:106-127 through the oracle's step: non-relativistic, exact OU, var -> M T
    keep = O.Keep()
    m, T, DsT = 1.5, 0.3, 0.11634
    med, tab = uniform_setup(keep, T, m, DsT)
    N = 20_000
    tau = 1.5 * DsT / T**2 / FMGEV
    dt = 0.02 * tau
    sp = O.species(keep, [m], [N], drag=[tab])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED, relativistic=False)
    P = np.zeros((2, N), order="F")
    M = np.zeros((2, N), order="F")
    rng = np.random.default_rng(20260615)
    for k in range(8):
        rc, _ = O.run_cartesian(p, med, sp, P, M, 1 + 100 * k, 100, rng.standard_normal((100, 2, N)))
        assert rc == 0
    assert np.var(M) == pytest.approx(m * T, rel=0.05)


@pytest.mark.parametrize("rel", [True, False])
def test_equilibrium_relativistic_switch_reference_gate(rel):
    # test_relativistic_switch.jl:63-80: d = 2, Juttner 2T(M^2+3MT+3T^2)/(M+T) vs Maxwell 2MT,
    # here through the oracle's own driver (positions stay at r = 0 off-grid is irrelevant: T uniform)
    keep = O.Keep()
    M_, T, DsT = 1.5, 0.3, 0.5
    med, tab = uniform_setup(keep, T, M_, DsT)
    N, nsteps, dt = 8000, 2000, 4e-3
    sp = O.species(keep, [M_], [N], drag=[tab])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED, relativistic=rel)
    rng = np.random.default_rng(20260806)
    P = np.zeros((2, N), order="F")
    M = np.asfortranarray(0.9 * rng.standard_normal((2, N)))
    for k in range(nsteps // 500):
        rc, _ = O.run_cartesian(p, med, sp, P, M, 1 + 500 * k, 500, rng.standard_normal((500, 2, N)))
        assert rc == 0
    p2 = np.mean(np.sum(M * M, axis=0))
    want = 2 * T * (M_**2 + 3 * M_ * T + 3 * T**2) / (M_ + T) if rel else 2 * M_ * T
    assert p2 == pytest.approx(want, rel=0.04)


def test_bjorken_redshift_telescopes_reference_gate():
    # test_momentum_dims3.jl:173-188: momentum_langevin=false, p_z only redshifts: tau0/tau
    keep = O.Keep()
    m, T = 1.5, 0.3
    xg = np.arange(0.0, 400.5, 0.5)
    tg = np.arange(0.4, 4.5, 1.0)
    med = O.radial_medium(keep, xg, tg, np.full((xg.size, tg.size), T), np.zeros((xg.size, tg.size)))
    N, t0, dt = 2000, 0.4, 1e-3
    sp = O.species(keep, [m], [N])
    p = _ffi.make_params(dt=dt, t0=t0, DsT=0.11634, collision=_ffi.LMH_FLUID, bjorken_redshift=True,
                         noise=_ffi.LMH_NOISE_FIXED)
    rng = np.random.default_rng(3)
    P = np.asfortranarray(4.0 * np.ones((2, N)))
    M = np.asfortranarray(np.vstack([0.3 * rng.standard_normal((2, N)), 0.5 * rng.standard_normal((1, N))]))
    pz0 = np.mean(M[2] ** 2)
    rc, _ = O.run_cartesian(p, med, sp, P, M, 1, 2000, np.zeros((2000, 0, N)))
    assert rc == 0
    tau = t0 + 2000 * dt
    assert np.mean(M[2] ** 2) == pytest.approx(pz0 * (t0 / tau) ** 2, rel=1e-9)


def test_set_to_fluid_closed_form():
    keep = O.Keep()
    m, v = 1.5, 0.4
    xg = np.linspace(0, 10, 11)
    tg = np.array([0.0, 1.0])
    med = O.radial_medium(keep, xg, tg, np.full((11, 2), 0.3), np.full((11, 2), v))
    sp = O.species(keep, [m], [1])
    p = _ffi.make_params(dt=0.01, DsT=0.0, collision=_ffi.LMH_FLUID, noise=_ffi.LMH_NOISE_FIXED)
    P = np.asfortranarray([[3.0], [4.0]])
    M = np.asfortranarray([[1.0], [-2.0]])
    rc, _ = O.run_cartesian(p, med, sp, P, M, 1, 1, np.zeros((1, 0, 1)))
    g = 1 / math.sqrt(1 - v * v + 1e-10)
    q = m * g * v * np.array([0.6, 0.8])
    E = math.sqrt(m * m + q @ q)
    assert rc == 0
    assert np.allclose(M[:, 0], q, rtol=4e-15, atol=0)
    assert np.allclose(P[:, 0], [3.0 + 0.01 * q[0] / E, 4.0 + 0.01 * q[1] / E], rtol=4e-15, atol=0)


def test_gc_step_closed_form():
    # fluid at rest, xi = 0: p^r -> g2 * a * g1 * p^r with the 1e-10 / 1e-8 regulator pair
    # (kernels_cpu_GC.jl:73,93), eta_D without m/E; tau advances by dt.
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.6366197723675814, 0.01
    med, tab = uniform_setup(keep, T, m, DsT)
    sp = O.species(keep, [m], [1], drag=[tab])
    p = _ffi.make_params(dt=dt, t0=0.5, DsT=DsT, coords=_ffi.LMH_GENERAL_COORDS, noise=_ffi.LMH_NOISE_FIXED)
    pr = 0.9
    P = np.asfortranarray([[0.5], [2.0]])
    M = np.asfortranarray([[math.sqrt(m * m + pr * pr)], [pr]])
    rc, _ = O.run_gc(p, med, sp, P, M, 1, 1, np.zeros((1, 1, 1)))
    assert rc == 0
    eta = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))
    g1 = 1 / mp.sqrt(1 + mp.mpf("1e-10"))
    g2 = 1 / mp.sqrt(1 + mp.mpf("1e-8"))
    prn = g2 * mp.exp(-eta * dt) * g1 * pr
    ptn = g2 * mp.sqrt(m * m + (mp.exp(-eta * dt) * g1 * pr) ** 2)
    assert M[1, 0] == pytest.approx(float(prn), rel=3e-15)
    assert M[0, 0] == pytest.approx(float(ptn), rel=3e-15)
    assert P[0, 0] == pytest.approx(0.5 + dt, rel=3e-15)
    assert P[1, 0] == pytest.approx(float(2.0 + dt * prn / ptn), rel=3e-15)


def test_nonfinite_stops_at_first_bad_step():
    keep = O.Keep()
    med, tab = uniform_setup(keep)
    sp = O.species(keep, [1.5], [3], drag=[tab])
    p = _ffi.make_params(dt=0.01, DsT=0.6366, noise=_ffi.LMH_NOISE_FIXED)
    P = np.zeros((3, 3), order="F")
    M = np.ones((3, 3), order="F")
    noise = np.zeros((5, 3, 3))
    noise[2, 1, 2] = np.inf
    rc, bad = O.run_cartesian(p, med, sp, P, M, 10, 5, noise)
    assert rc == -4 and bad == 12


def test_grid_sample_affine_and_radial_agreement():
    from lmhip import media
    g = media.hydro_grid(nx=21, ny=31, nt=5, L=5.0)
    # replace the fields by an affine function of (x, y, tau): the sample is exact
    X = g.x0 + g.dx * np.arange(g.nx)
    Y = g.y0 + g.dy * np.arange(g.ny)
    Tg = np.asarray(g.tgrid)
    A = 0.2 + 0.01 * X[:, None, None] - 0.02 * Y[None, :, None] + 0.03 * Tg[None, None, :]
    g.T = np.asfortranarray(A)
    g.vx = np.asfortranarray(0.5 * A)
    g.vy = np.asfortranarray(-A)
    keep = O.Keep()
    gm = O.grid_medium(keep, g)
    out = np.zeros(3)
    rng = np.random.default_rng(1)
    for _ in range(100):
        x, y, t = rng.uniform(-5, 5), rng.uniform(-5, 5), rng.uniform(Tg[0], Tg[-1])
        L.lmo_grid_sample(ctypes.byref(gm), x, y, t, O.dp(out))
        want = 0.2 + 0.01 * x - 0.02 * y + 0.03 * t
        assert out[0] == pytest.approx(want, rel=1e-13)
        assert out[1] == pytest.approx(0.5 * want, rel=1e-13)
        assert out[2] == pytest.approx(-want, rel=1e-13)
    # clamped outside the grid
    L.lmo_grid_sample(ctypes.byref(gm), 50.0, -50.0, 100.0, O.dp(out))
    assert out[0] == pytest.approx(0.2 + 0.05 + 0.1 + 0.03 * Tg[-1], rel=1e-13)


def test_grid_step_equals_radial_step_for_radial_flow():
    # a 2+1D grid holding a radial medium: the v_hat boost equals the r_hat boost when the flow
    # is radial and outward; with a medium linear in r and tau both interpolations are exact,
    # so one fixed-noise step agrees to rounding.
    from lmhip import media
    m, DsT, dt = 1.5, 0.6366197723675814, 0.01
    n = 41
    L_ = 4.0
    x = np.linspace(-L_, L_, n)
    tg = np.array([1.0, 2.0, 3.0])
    X, Y = np.meshgrid(x, x, indexing="ij")
    R = np.sqrt(X * X + Y * Y)
    Tfun = lambda r, t: 0.35 - 0.0 * r - 0.01 * t
    vfun = lambda r, t: 0.05 * r + 0.01 * t
    T = np.stack([Tfun(R, t) for t in tg], axis=2)
    with np.errstate(invalid="ignore"):
        cx, cy = np.where(R > 0, X / R, 0), np.where(R > 0, Y / R, 0)
    # vx, vy = v(r) r_hat is not affine in (x, y); keep particles on the x axis where it is
    vx = np.stack([vfun(R, t) * cx for t in tg], axis=2)
    vy = np.stack([vfun(R, t) * cy for t in tg], axis=2)
    g = media.Grid2p1Medium(float(x[0]), float(x[1] - x[0]), n, float(x[0]), float(x[1] - x[0]), n, tg,
                            np.asfortranarray(T), np.asfortranarray(vx), np.asfortranarray(vy))
    keep = O.Keep()
    gm = O.grid_medium(keep, g)
    rg = np.linspace(0, 6, 61)
    rm = O.radial_medium(keep, rg, tg, np.stack([Tfun(rg, t) for t in tg], 1), np.stack([vfun(rg, t) for t in tg], 1))
    rc, t0, inv, d = build_table(0, m, DsT, 0.3, 0.35, 1024)
    sp = O.species(keep, [m], [3], drag=[(t0, inv, d)])
    p = _ffi.make_params(dt=dt, t0=1.0, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    pos = np.asfortranarray([[0.6, 1.4, 3.0], [0.0, 0.0, 0.0]])     # on the x axis, nodes every 0.2
    mom = np.asfortranarray([[0.5, -0.3, 1.0], [0.2, 0.4, -0.1], [0.1, 0.0, 0.3]])
    noise = np.random.default_rng(2).standard_normal((1, 3, 3))
    P1, M1 = pos.copy(order="F"), mom.copy(order="F")
    P2, M2 = pos.copy(order="F"), mom.copy(order="F")
    assert O.run_grid(p, gm, sp, P1, M1, 1, 1, noise)[0] == 0
    assert O.run_cartesian(p, rm, sp, P2, M2, 1, 1, noise)[0] == 0
    assert np.allclose(M1, M2, rtol=1e-13, atol=1e-15)
    assert np.allclose(P1, P2, rtol=1e-13, atol=1e-15)


def test_observables_vs_numpy():
    rng = np.random.default_rng(4)
    mom = np.asfortranarray(rng.normal(0, 2, (3, 5000)))
    out = O.observables(mom, 0, 5000, 40, 0.0, 8.0, 0.0)
    pt = np.hypot(mom[0], mom[1])
    assert out[0] == 5000
    assert out[1] == pytest.approx(pt.sum(), rel=1e-13)
    assert out[3] == pytest.approx(((mom[0] ** 2 - mom[1] ** 2) / pt ** 2).sum(), rel=1e-12)
    h, _ = np.histogram(pt, bins=40, range=(0, 8))
    assert np.array_equal(out[5:], h.astype(float))
    sm = O.observables(mom, 0, 5000, 40, 0.0, 8.0, 0.3)
    # the smoothed bins of one particle telescope to Phi((8-pT)/s) - Phi((0-pT)/s)
    from scipy.stats import norm
    assert sm[5:].sum() == pytest.approx(np.sum(norm.cdf((8 - pt) / 0.3) - norm.cdf(-pt / 0.3)), rel=1e-10)
    # and each bin is the Gaussian mass of the bin around every particle
    assert sm[5 + 7] == pytest.approx(np.sum(norm.cdf((1.6 - pt) / 0.3) - norm.cdf((1.4 - pt) / 0.3)), rel=1e-9)


# ---------------------------------------------------------------- freeze-out capture ----------
def _bjorken_capture(interp, Tfo=0.2, N=300, dt=0.01, T0=0.4, tau0=0.6):
    xg, tg, Tf, Vf, _ = media.bjorken_bath(T0=T0, tau0=tau0, Tstop=0.15, ntau=4096)
    keep = O.Keep()
    med = O.radial_medium(keep, xg, tg, Tf, Vf)
    drag = lmhip.build_tau_drag_spline(1.5, 0.6366, Tmin=float(Tf.min()), Tmax=float(Tf.max()), nT=1024)
    sp = O.species(keep, [1.5], [N], drag=[drag])
    p = _ffi.make_params(dt=dt, t0=tau0, DsT=0.6366, noise=_ffi.LMH_NOISE_FIXED)
    rng = np.random.default_rng(4)
    pos = np.asfortranarray(rng.normal(0, 2, (3, N)))
    mom = np.asfortranarray(rng.normal(0, 1, (3, N)))
    fo = O.Freezeout(3, 3, N)
    steps = int((tau0 * (T0 / Tfo) ** 3 - tau0) / dt) + 20
    hist = []
    for s in range(1, steps + 1):
        xi = np.ascontiguousarray(rng.normal(size=(1, 3, N)))
        hist.append((pos.copy(), mom.copy()))
        rc, _ = O.run_cartesian(p, med, sp, pos, mom, s, 1, xi)
        assert rc == 0
        fo.capture(med, None, pos, mom, s, dt, tau0, Tfo, interp)
    hist.append((pos.copy(), mom.copy()))
    return fo, hist, tau0 * (T0 / Tfo) ** 3, dt, tau0


def test_freezeout_bjorken_crossing_time():
    """T(tau) = T0 (tau0/tau)^(1/3) at every r: every quark latches at tau* = tau0 (T0/Tfo)^3,
    to O(dt^2) with the interpolation in T and within one step after tau* without it"""
    fo, hist, tstar, dt, tau0 = _bjorken_capture(True)
    assert np.all(fo.fo_flag == 1.0)
    assert np.max(np.abs(fo.fo_tau - tstar)) < 1e-5
    fo0, _, _, _, _ = _bjorken_capture(False)
    lag = fo0.fo_tau - tstar
    assert np.all(lag > 0.0) and np.all(lag <= dt + 1e-12)
    # the latched momentum lies between the states before and after the crossing step
    s = int(round((fo0.fo_tau[0] - tau0) / dt))
    lo, hi = hist[s - 1][1], hist[s][1]
    assert np.all((fo.fo_mom - np.minimum(lo, hi)) >= -1e-12) and np.all((np.maximum(lo, hi) - fo.fo_mom) >= -1e-12)
    assert np.allclose(fo0.fo_mom, hi, rtol=1e-15, atol=1e-15)   # last + 1 * (current - last)


def test_position_diffusion_msd_gives_Ds():
    """Position diffusion (kernels_cpu.jl:591-607): with the momenta glued to a fluid at rest
    (no streaming), the diffusion kicks alone make the 2-D mean squared displacement 4 D t with
    D = (D_s T / T) / hbar c (rtol 3 %; 2e4 quarks x 400 steps, statistical error 0.7 %)"""
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.6366197723675814, 0.01
    med, tab = uniform_setup(keep, T, m, DsT)
    N, steps = 20_000, 400
    sp = O.species(keep, [m], [N], drag=[tab])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_PHILOX, seed=12, collision=_ffi.LMH_FLUID,
                         position_diffusion=True)
    P = np.zeros((2, N), order="F")
    M = np.zeros((2, N), order="F")
    rec = O.noise_record(p, 2, 2, N, 1, steps)
    rc, _ = O.run_cartesian(p, med, sp, P, M, 1, steps, np.ascontiguousarray(rec))
    assert rc == 0
    assert np.all(M == 0.0)
    D = DsT / T / FMGEV
    assert np.mean(np.sum(P * P, axis=0)) == pytest.approx(4 * D * steps * dt, rel=0.03)


def test_reflecting_wall_preserves_momentum_magnitude():
    """The reflecting wall (kernels_cpu.jl:609-634) at r_max, the last radial node: a quark that
    crosses it comes back to 2 r_max - r with its radial momentum flipped, its tangential momentum
    and |p| unchanged -- against the same step without the wall"""
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.6366197723675814, 0.01
    med, tab = uniform_setup(keep, T, m, DsT)               # radial nodes 0 .. 30 fm
    N = 64
    rng = np.random.default_rng(4)
    phi = rng.uniform(0.0, 2 * np.pi, N)
    P0 = np.asfortranarray(29.995 * np.stack([np.cos(phi), np.sin(phi)]))
    pr, pt = rng.uniform(3.0, 6.0, N), rng.uniform(-1.0, 1.0, N)
    M0 = np.asfortranarray(np.stack([pr * np.cos(phi) - pt * np.sin(phi), pr * np.sin(phi) + pt * np.cos(phi)]))
    sp = O.species(keep, [m], [N], drag=[tab])
    out = []
    for wall in (False, True):
        p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED, reflecting_boundary=wall)
        P, M = P0.copy(order="F"), M0.copy(order="F")
        rc, _ = O.run_cartesian(p, med, sp, P, M, 1, 1, np.zeros((1, 2, N)))
        assert rc == 0
        out.append((P, M))
    (Pf, Mf), (Pr, Mr) = out
    rf, rr = np.hypot(Pf[0], Pf[1]), np.hypot(Pr[0], Pr[1])
    assert np.all(rf > 30.0) and np.all(rr <= 30.0)
    np.testing.assert_allclose(rr, 60.0 - rf, rtol=1e-14)
    np.testing.assert_allclose(np.hypot(Mr[0], Mr[1]), np.hypot(Mf[0], Mf[1]), rtol=1e-14)
    nx, ny = Pf[0] / rf, Pf[1] / rf
    np.testing.assert_allclose(Mr[0] * nx + Mr[1] * ny, -(Mf[0] * nx + Mf[1] * ny), rtol=1e-12)
    np.testing.assert_allclose(-Mr[0] * ny + Mr[1] * nx, -Mf[0] * ny + Mf[1] * nx, rtol=1e-12, atol=1e-14)
