"""Pins of the oracle's step functions that the reference's own tests leave open: the RTA/BGK
collision (kernels_cpu.jl:474-508 with the inverse-CDF redraw of kernels_gpu.jl:653-704), the
drift-midpoint integrator (kernels_gpu.jl:366-381), the post-point integrator and the
gamma(p, T) drag (extensions), and the radial-mode Euler kick (kernels_cpu.jl:293-323).

Every expected value comes from the mathematics, not from the oracle's formula: the exact
solution of the drag ODE (order of the local error), the fixed point of the implicit post-point
equation, the stationary distribution of the Fokker-Planck equation each drag defines, binomial
statistics of the collision decision with a deterministic per-particle threshold, the
Juttner quantiles by quadrature, and 50-digit closed forms of one step.  Each check is chosen
so that a plausible mistake (a flipped comparison, a swapped uniform, a missing sweep, a wrong
half step, a dropped m/E) fails it.
"""
import math

import mpmath as mp
import numpy as np
import pytest

import _oracle as O
from lmhip import _ffi

mp.mp.dps = 50
L = O.lib()
FMGEV = 1 / 0.197327
EPS = np.finfo(float).eps


def bath(keep, T=0.3, V=0.0):
    xg = np.arange(0.0, 30.25, 0.5)
    tg = np.arange(0.0, 40.5, 1.0)
    return O.radial_medium(keep, xg, tg, np.full((xg.size, tg.size), T), np.full((xg.size, tg.size), V))


def drag_table(m, DsT, T):
    """the reference's drag table on [T, T + eps]: T itself is node 0 (eval gives vals[0])"""
    vals = np.zeros(1024)
    t0, inv = O.ctypes.c_double(), O.ctypes.c_double()
    rc = L.lmo_build_time_spline(0, m, DsT, T, T + EPS, 1024, 0, 1.765, -0.159, 0.156, 0, 0.0, 1.0,
                                 O.ctypes.byref(t0), O.ctypes.byref(inv), O.dp(vals))
    assert rc == 0
    assert vals[0] == m * DsT / (T * T) / FMGEV           # runtests.jl:96-103
    return t0.value, inv.value, vals


def const_table(val, T, n=8):
    """a relaxation-time table holding one value, with T on its first node (eval is exact there)"""
    return (T, 1.0, np.full(n, float(val)))


def jt_table(m, d, T):
    """inverse-CDF Juttner table whose node 100 is T (node spacing 1e-3)"""
    nT, nU = 201, 2048
    inv = np.zeros(nT * nU)
    t0, idt = O.ctypes.c_double(), O.ctypes.c_double()
    rc = L.lmo_build_juttner_invcdf(m, d, T - 0.1, T + 0.1, nT, nU, 8192, 1e-6, O.dp(inv),
                                    O.ctypes.byref(t0), O.ctypes.byref(idt))
    assert rc == 0
    return inv, nU, nT, t0.value, idt.value


def juttner_p2(m, T, d):
    w = lambda p: p ** (d - 1) * mp.exp(-(mp.sqrt(p * p + m * m) - m) / T)
    lim = [0, 10 * T, 40 * T, mp.inf]
    return float(mp.quad(lambda p: p * p * w(p), lim) / mp.quad(w, lim))


def juttner_quantile(m, T, d, u):
    """the exact inverse CDF of P(p) ~ p^(d-1) exp(-(E-m)/T) on [0, pmax] (the table's support)"""
    from scipy.integrate import quad
    from scipy.optimize import brentq
    w = lambda p: p ** (d - 1) * math.exp(-(math.sqrt(p * p + m * m) - m) / T)
    pmax = 16 * math.sqrt(m * T) + 16 * T
    Z = quad(w, 0, pmax, limit=200)[0]
    return brentq(lambda p: quad(w, 0, p, limit=200)[0] / Z - u, 0, pmax, xtol=1e-14)


# ---------------------------------------------------------------- RTA / BGK collision ------------
def rta_run(keep, N, T, taun, ucol, usamp, g, m=1.5, dt=0.01, DsT=0.5, pdim=3, mom=None, steps=1):
    """RTA-only steps of quarks at the origin (r = 0: no boost), fluid at rest; returns the
    momenta before and after"""
    med = bath(keep, T)
    jt = jt_table(m, pdim, 0.3)
    sp = O.species(keep, [m], [N], rta=[taun], jt=[jt])
    p = _ffi.make_params(dt=dt, DsT=DsT, collision=_ffi.LMH_RTA, noise=_ffi.LMH_NOISE_FIXED)
    pos = np.zeros((2 if pdim == 3 else pdim, N), order="F")
    if mom is None:
        mom = np.asfortranarray(np.tile(np.array([[1.0], [-0.5], [2.0]])[:pdim], (1, N)))
    M = mom.copy(order="F")
    rec = np.zeros((steps, 2 + pdim, N))
    rec[:, 0], rec[:, 1], rec[:, 2:] = ucol, usamp, g
    rc, _ = O.run_cartesian(p, med, sp, pos, M, 1, steps, rec)
    assert rc == 0
    return mom, M, jt


def test_rta_collision_decision_is_u_below_dt_over_tau():
    """kernels_cpu.jl:489-491: collide iff u < clamp(dt/tau_n, 0, 1), particle by particle,
    including uniforms one ulp either side of the threshold; non-colliding momenta untouched"""
    keep = O.Keep()
    N, T, dt, tau = 4000, 0.3, 0.01, 0.01 / 0.3
    P = dt / tau                                               # the threshold the reference forms
    rng = np.random.default_rng(1)
    ucol = rng.random(N)
    ucol[:6] = [P, np.nextafter(P, 0), np.nextafter(P, 1), 0.0, np.nextafter(1.0, 0), P / 2]
    usamp = rng.random(N)
    g = rng.standard_normal((3, N))
    mom, M, _ = rta_run(keep, N, T, const_table(tau, T), ucol, usamp, g, dt=dt)
    changed = np.any(M != mom, axis=0)
    assert np.array_equal(changed, ucol < P)
    assert changed[[1, 3, 5]].all() and not changed[[0, 2, 4]].any()
    frac = changed.mean()
    assert abs(frac - P) < 4 * math.sqrt(P * (1 - P) / N)


@pytest.mark.parametrize("case", ["tau_zero", "tau_negative", "tau_inf", "dt_over_tau_above_1",
                                  "T_zero", "DsT_nonpositive"])
def test_rta_certain_collision_cases(case):
    """tau_n <= 0 or non-finite, T <= 0 or DsT <= 0 give Pcol = 1; dt/tau > 1 clamps to 1
    (kernels_cpu.jl:488-490): every uniform in [0, 1) collides"""
    keep = O.Keep()
    N, dt = 500, 0.01
    T = 0.0 if case == "T_zero" else 0.3
    tau = {"tau_zero": 0.0, "tau_negative": -2.0, "tau_inf": np.inf, "dt_over_tau_above_1": 0.5 * dt,
           "T_zero": 1.0, "DsT_nonpositive": 1.0}[case]
    DsT = -1.0 if case == "DsT_nonpositive" else 0.5
    rng = np.random.default_rng(2)
    ucol = rng.random(N)
    ucol[0] = np.nextafter(1.0, 0)
    mom, M, _ = rta_run(keep, N, T, const_table(tau, max(T, 0.3)), ucol, rng.random(N),
                        rng.standard_normal((3, N)), dt=dt, DsT=DsT)
    assert np.all(np.any(M != mom, axis=0))


def test_rta_redraw_magnitude_and_direction():
    """a colliding quark gets |p| = the Juttner quantile of u_samp at the local T (exact inverse
    CDF by quadrature, table nodes) and the direction of its Gaussian triple g/|g|
    (kernels_gpu.jl:673-701); u_col plays no part in it"""
    keep = O.Keep()
    m, T, N = 1.5, 0.3, 12
    nU = 2048
    us = np.array([50, 200, 600, 1024, 1500, 1900, 2000, 2030, 100, 700, 1300, 1800]) / (nU - 1)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((3, N))
    ucol = np.full(N, 0.25)                                    # all collide (P = 0.3)
    mom, M, jt = rta_run(keep, N, T, const_table(0.01 / 0.3, T), ucol, us, g)
    pm = np.linalg.norm(M, axis=0)
    assert np.allclose(M / pm, g / np.linalg.norm(g, axis=0), rtol=0, atol=4e-16)
    want = np.array([juttner_quantile(m, T, 3, u) for u in us])
    assert np.allclose(pm, want, rtol=2e-5, atol=0)
    # the same uniforms with another u_col below the threshold give the same state
    _, M2, _ = rta_run(O.Keep(), N, T, const_table(0.01 / 0.3, T), np.full(N, 0.01), us, g)
    assert np.array_equal(M, M2)


@pytest.mark.parametrize("pdim", [2, 3])
def test_rta_relaxes_to_juttner_at_rate_pcol(pdim):
    """BGK relaxation: after n steps the untouched fraction is (1 - P)^n (binomial) and the
    redrawn quarks carry the Juttner <p^2> of dimension pdim (quadrature)"""
    keep = O.Keep()
    m, T, N, n, P = 1.5, 0.3, 20000, 60, 0.05
    dt = 0.01
    rng = np.random.default_rng(4 + pdim)
    d = rng.standard_normal((pdim, N))
    mom = np.asfortranarray(4.0 * d / np.linalg.norm(d, axis=0))
    mom0, M, _ = rta_run(keep, N, T, const_table(dt / P, T), rng.random((n, N)), rng.random((n, N)),
                         rng.standard_normal((n, pdim, N)), dt=dt, pdim=pdim, mom=mom, steps=n)
    # untouched momenta keep their value up to the boosts' 1e-10 regulator per step
    untouched = np.all(np.abs(M - mom0) <= 1e-7 * np.abs(mom0) + 1e-12, axis=0)
    q = (1 - P) ** n
    assert abs(untouched.mean() - q) < 4 * math.sqrt(q * (1 - q) / N)
    p2 = np.sum(M[:, ~untouched] ** 2, axis=0)
    want = juttner_p2(m, T, pdim)
    assert p2.mean() == pytest.approx(want, rel=4 * p2.std() / math.sqrt(p2.size) / want + 6e-3)


# ---------------------------------------------------------------- one OU step helpers ------------
def one_step(integrator, q, xi, *, rel=True, drag_model=_ffi.LMH_DRAG_TABLE, dt=0.01, m=1.5, T=0.3,
             DsT=0.6366197723675814, gamma0=5.3, Lambda=5.0):
    """one Langevin step of quarks at the origin (no boost: r = 0), fluid at rest; q (pdim, N)"""
    keep = O.Keep()
    q = np.asfortranarray(q, dtype=float)
    pdim, N = q.shape
    med = bath(keep, T)
    sp = O.species(keep, [m], [N], drag=[drag_table(m, DsT, T)])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED, integrator=integrator,
                         relativistic=rel, drag_model=drag_model, gamma0=gamma0, Lambda=Lambda)
    pos = np.zeros((2 if pdim == 3 else pdim, N), order="F")
    M = q.copy(order="F")
    rc, _ = O.run_cartesian(p, med, sp, pos, M, 1, 1, np.ascontiguousarray(xi, dtype=float).reshape(1, pdim, N))
    assert rc == 0
    return M


def drag_ode_exact(p0, eta_m, t, m):
    """|p|(t) of dp/dt = -eta_D (m/E) p, solved through F(p) = E - m artanh(m/E), dF/dp = E/p:
    F(p(t)) = F(p0) - eta_D m t"""
    F = lambda p: mp.sqrt(p * p + m * m) - m * mp.atanh(m / mp.sqrt(p * p + m * m))
    target = F(p0) - eta_m * t
    return mp.findroot(lambda p: F(p) - target, p0 * mp.exp(-eta_m * t / mp.sqrt(p0 * p0 + m * m)))


# ---------------------------------------------------------------- drift midpoint -----------------
def test_midpoint_is_second_order_for_the_drag():
    """without noise the drift-midpoint step (kernels_gpu.jl:368-381) follows the exact solution of
    dp/dt = -eta_D (m/E) p with a local error O(dt^3): halving dt divides it by ~8, where the
    pre-point step's O(dt^2) error divides by ~4"""
    m, T, DsT = 1.5, 0.3, 0.2
    eta = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))
    p0 = mp.mpf("2.0")
    errs = {}
    for integ in (_ffi.LMH_MIDPOINT, _ffi.LMH_PREPOINT):
        e = []
        for dt in (0.08, 0.04, 0.02):
            M = one_step(integ, [[2.0], [0.0], [0.0]], np.zeros((3, 1)), dt=dt, DsT=DsT)
            exact = drag_ode_exact(p0, eta * m, mp.mpf(dt), m)
            e.append(abs(float(M[0, 0] - exact) / float(exact)))
        errs[integ] = e
    mid, pre = errs[_ffi.LMH_MIDPOINT], errs[_ffi.LMH_PREPOINT]
    for a, b in zip(mid, mid[1:]):
        assert 6.5 < a / b < 9.5, mid
    for a, b in zip(pre, pre[1:]):
        assert 3.3 < a / b < 4.7, pre
    assert mid[-1] < pre[-1] / 20


def test_midpoint_noise_is_the_exact_ou_deviation_at_the_midpoint_drag():
    """the noise part p(xi) - p(0) = sqrt(kappa (1 - a^2)/(2 eta_mid)) xi, a = exp(-eta_mid dt),
    with eta_mid the drag at the noise-free half step e^(-eta0 dt/2) p (not at p, not with xi)"""
    m, T, DsT, dt = 1.5, 0.3, 0.2, 0.05
    q = np.array([[1.2], [-0.7], [0.4]])
    xi = np.array([[0.9], [-1.1], [0.3]])
    d = one_step(_ffi.LMH_MIDPOINT, q, xi, dt=dt, DsT=DsT) - one_step(_ffi.LMH_MIDPOINT, q, 0 * xi, dt=dt, DsT=DsT)
    eta_D = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))
    pq = [mp.mpf(v) for v in q[:, 0]]
    E0 = mp.sqrt(sum(v * v for v in pq) + m * m)
    half = mp.exp(-eta_D * m / E0 * dt / 2)
    Em = mp.sqrt(sum((half * v) ** 2 for v in pq) + m * m)
    eta_mid = eta_D * m / Em
    sd = mp.sqrt(2 * m * T * eta_D * (1 - mp.exp(-2 * eta_mid * dt)) / (2 * eta_mid))
    assert np.allclose(d[:, 0], [float(sd * v) for v in xi[:, 0]], rtol=1e-13, atol=1e-17)


def test_midpoint_equals_prepoint_without_momentum_dependence():
    """non-relativistic drag does not depend on p: the midpoint re-evaluation changes nothing"""
    rng = np.random.default_rng(7)
    q, xi = rng.normal(0, 1.5, (3, 200)), rng.standard_normal((3, 200))
    a = one_step(_ffi.LMH_MIDPOINT, q, xi, rel=False)
    b = one_step(_ffi.LMH_PREPOINT, q, xi, rel=False)
    assert np.array_equal(a, b)
    c = one_step(_ffi.LMH_MIDPOINT, q, xi, rel=True)
    assert not np.array_equal(a, c)


# ---------------------------------------------------------------- post point ---------------------
def test_postpoint_equals_prepoint_without_momentum_dependence():
    rng = np.random.default_rng(8)
    q, xi = rng.normal(0, 1.5, (3, 200)), rng.standard_normal((3, 200))
    assert np.array_equal(one_step(_ffi.LMH_POSTPOINT, q, xi, rel=False),
                          one_step(_ffi.LMH_PREPOINT, q, xi, rel=False))
    big = 1e200
    assert np.array_equal(one_step(_ffi.LMH_POSTPOINT, q, xi, drag_model=_ffi.LMH_DRAG_GAMMA_P, Lambda=big),
                          one_step(_ffi.LMH_PREPOINT, q, xi, drag_model=_ffi.LMH_DRAG_GAMMA_P, Lambda=big))


@pytest.mark.parametrize("drag_model", [_ffi.LMH_DRAG_TABLE, _ffi.LMH_DRAG_GAMMA_P])
def test_postpoint_solves_the_implicit_step(drag_model):
    """the post-point result q is (to the contraction error of three sweeps) the fixed point of
    q = p + (a(q) - 1) p + sqrt((1 - a(q)^2) kappa/(2 eta(q))) xi, a(q) = exp(-eta(q) dt), with the
    drag at q (relativistic eta_D m/E(q), or gamma0 T^2/(1 + q^2/Lambda^2)); one or two sweeps
    leave a residual three to six orders larger"""
    m, T, DsT, dt, g0, Lam = 1.5, 0.3, 0.2, 0.01, 5.3, 2.0
    rng = np.random.default_rng(9)
    q, xi = rng.normal(0, 2.0, (3, 40)), rng.standard_normal((3, 40))
    out = one_step(_ffi.LMH_POSTPOINT, q, xi, dt=dt, DsT=DsT, drag_model=drag_model, gamma0=g0, Lambda=Lam)
    if drag_model == _ffi.LMH_DRAG_TABLE:
        eta_D = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))
        eta = lambda v: eta_D * m / mp.sqrt(sum(c * c for c in v) + m * m)
    else:
        eta_D = mp.mpf(g0) * mp.mpf(T) ** 2
        eta = lambda v: eta_D / (1 + sum(c * c for c in v) / mp.mpf(Lam) ** 2)
    kappa = 2 * m * mp.mpf(T) * eta_D

    def G(p, v, x):
        e = eta(v)
        a = mp.exp(-e * dt)
        s = mp.sqrt((1 - a * a) * kappa / (2 * e))
        return [pi + (a - 1) * pi + s * xv for pi, xv in zip(p, x)]

    worst, one_sweep = 0.0, 0.0
    for i in range(q.shape[1]):
        p = [mp.mpf(v) for v in q[:, i]]
        x = [mp.mpf(v) for v in xi[:, i]]
        v = [mp.mpf(c) for c in out[:, i]]
        res = G(p, v, x)
        nrm = mp.sqrt(sum(c * c for c in v))
        worst = max(worst, float(mp.sqrt(sum((a - b) ** 2 for a, b in zip(res, v))) / nrm))
        q1 = G(p, p, x)                                                # a single sweep
        one_sweep = max(one_sweep, float(mp.sqrt(sum((a - b) ** 2 for a, b in zip(G(p, q1, x), q1))) / nrm))
    # the residual after k sweeps shrinks as L^k with the contraction constant L <= 6e-3 here
    assert worst < 1e-7, worst
    assert one_sweep > 1e3 * worst, (one_sweep, worst)


def _stationary_p2(integrator, drag_model, Lam, seed, M0, rel=True):
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.3, 0.02
    N, nsteps = M0.shape[1], 1200
    med = bath(keep, T)
    sp = O.species(keep, [m], [N], drag=[drag_table(m, DsT, T)])
    p = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED, integrator=integrator,
                         relativistic=rel, drag_model=drag_model, gamma0=5.3, Lambda=Lam)
    rng = np.random.default_rng(seed)
    P = np.zeros((2, N), order="F")
    M = np.asfortranarray(M0)
    for k in range(nsteps // 200):
        rc, _ = O.run_cartesian(p, med, sp, P, M, 1 + 200 * k, 200, rng.standard_normal((200, 3, N)))
        assert rc == 0
    return np.sum(M * M, axis=0)


def test_postpoint_relaxes_to_juttner():
    """with the relativistic table drag (kappa = 2 m T eta_D constant, eta = eta_D m/E) the
    Fokker-Planck stationary state is Juttner exp(-E/T) whatever point the drag is taken at:
    <p^2> = 3 m T K3/K2 (quadrature)"""
    rng = np.random.default_rng(21)
    p2 = _stationary_p2(_ffi.LMH_POSTPOINT, _ffi.LMH_DRAG_TABLE, 5.0, 11,
                        rng.normal(0, math.sqrt(1.5 * 0.3), (3, 10000)))
    want = juttner_p2(1.5, 0.3, 3)
    assert p2.mean() == pytest.approx(want, rel=0.03)


@pytest.mark.parametrize("integrator", [_ffi.LMH_PREPOINT, _ffi.LMH_POSTPOINT])
def test_gamma_p_drag_stationary_state(integrator):
    """gamma(p, T) = gamma0 T^2/(1 + p^2/Lambda^2) with kappa = 2 m T gamma0 T^2: the stationary
    Fokker-Planck solution is f ~ (1 + p^2/Lambda^2)^(-nu), nu = Lambda^2/(2 m T), whose 3-D
    second moment is 3 Lambda^2/(2 nu - 5) (33 % above Maxwell's 3 m T at Lambda = 3)"""
    m, T, Lam = 1.5, 0.3, 3.0
    nu = Lam * Lam / (2 * m * T)
    # start from that distribution itself (a 3-D Student t: p = Lambda Z/sqrt(W), W ~ chi^2 with
    # 2 nu - 3 degrees of freedom), so that 12 relaxation times test that the dynamics keeps it
    rng = np.random.default_rng(22 + integrator)
    N = 20000
    M0 = Lam * rng.standard_normal((3, N)) / np.sqrt(rng.chisquare(2 * nu - 3, N))
    p2 = _stationary_p2(integrator, _ffi.LMH_DRAG_GAMMA_P, Lam, 12 + integrator, M0)
    want = 3 * Lam * Lam / (2 * nu - 5)
    assert p2.mean() == pytest.approx(want, rel=0.04)
    assert abs(p2.mean() - 3 * m * T) > 0.2 * 3 * m * T


def test_gamma_p_large_lambda_is_the_constant_drag_ou():
    """Lambda -> infinity: eta = gamma0 T^2 for every p (no m/E), one exact-OU step
    p' = a p + sqrt(kappa (1 - a^2)/(2 eta)) xi with kappa = 2 m T gamma0 T^2 (mpmath)"""
    m, T, dt, g0 = 1.5, 0.3, 0.02, 5.3
    q = np.array([[1.2, 3.0], [-0.7, 0.1], [0.4, -2.5]])
    xi = np.array([[0.9, -0.2], [-1.1, 0.6], [0.3, 1.7]])
    for integ in (_ffi.LMH_PREPOINT, _ffi.LMH_POSTPOINT, _ffi.LMH_MIDPOINT):
        out = one_step(integ, q, xi, dt=dt, drag_model=_ffi.LMH_DRAG_GAMMA_P, gamma0=g0, Lambda=1e200)
        eta = mp.mpf(g0) * mp.mpf(T) ** 2
        a = mp.exp(-eta * dt)
        sd = mp.sqrt(2 * m * mp.mpf(T) * eta * (1 - a * a) / (2 * eta))
        want = [[float(a * mp.mpf(q[d, i]) + sd * mp.mpf(xi[d, i])) for i in range(2)] for d in range(3)]
        assert np.allclose(out, want, rtol=1e-14, atol=0)


# ---------------------------------------------------------------- radial-mode Euler kick ----------
def test_radial_mode_euler_kick_closed_form():
    """dims = pdim = 1 (kernels_cpu.jl:293-323, :406-417, boosts :114-118): with the fluid at rest
    the two boosts scale p by g = 1/sqrt(1 + 1e-10) each; the kick is the Euler step
    p + (-eta_D (m/E) p dt + sqrt(dt) kT xi) with E from the LRF momentum, kT = sqrt(2 m T eta_D)
    (kL == kT, so the projector term vanishes); then r += (p/E) dt"""
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.6366197723675814, 0.01
    med = bath(keep, T)
    sp = O.species(keep, [m], [3], drag=[drag_table(m, DsT, T)])
    prm = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    P = np.asfortranarray([[2.0, 5.0, 0.7]])
    M = np.asfortranarray([[0.8, -1.7, 3.2]])
    xi = np.array([0.4, -1.3, 0.05])
    P1, M1 = P.copy(order="F"), M.copy(order="F")
    rc, _ = O.run_cartesian(prm, med, sp, P1, M1, 1, 1, xi.reshape(1, 1, 3))
    assert rc == 0
    eta_D = mp.mpf(T) ** 2 * mp.mpf(FMGEV) / (mp.mpf(m) * mp.mpf(DsT))
    kT = mp.sqrt(2 * m * mp.mpf(T) * eta_D)
    g = 1 / mp.sqrt(1 + mp.mpf("1e-10"))
    for i in range(3):
        p = g * mp.mpf(M[0, i])
        E = mp.sqrt(p * p + m * m)
        p = p + (-eta_D * m / E * p * dt + mp.sqrt(dt) * kT * mp.mpf(xi[i]))
        p = g * p
        r = mp.mpf(P[0, i]) + p / mp.sqrt(p * p + m * m) * dt
        assert M1[0, i] == pytest.approx(float(p), rel=2e-15)
        assert P1[0, i] == pytest.approx(float(r), rel=2e-15)


def test_radial_mode_euler_equilibrium_is_1d_juttner():
    """the radial Euler kick with the relativistic drag relaxes to the 1-D Juttner <p^2>
    (quadrature), not to Maxwell's m T"""
    keep = O.Keep()
    m, T, DsT, dt = 1.5, 0.3, 0.3, 0.005
    N, nsteps = 10000, 2400
    med = bath(keep, T)
    sp = O.species(keep, [m], [N], drag=[drag_table(m, DsT, T)])
    prm = _ffi.make_params(dt=dt, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    rng = np.random.default_rng(13)
    P = np.asfortranarray(5.0 + rng.random((1, N)))
    M = np.asfortranarray(rng.normal(0, math.sqrt(m * T), (1, N)))
    for k in range(nsteps // 400):
        rc, _ = O.run_cartesian(prm, med, sp, P, M, 1 + 400 * k, 400, rng.standard_normal((400, 1, N)))
        assert rc == 0
    want = juttner_p2(m, T, 1)
    assert np.mean(M * M) == pytest.approx(want, rel=0.03)
    assert abs(want - m * T) > 0.06 * m * T


# ---------------------------------------------------------------- Milne coordinates --------------
def milne_run(params_kw, pos, mom, steps, noise=None, T=0.3, V=0.0, m=1.5, DsT=0.5, t0=0.6, dt=0.005):
    keep = O.Keep()
    med = bath(keep, T, V)
    N = pos.shape[1]
    sp = O.species(keep, [m], [N], drag=[drag_table(m, DsT, T)])
    kw = dict(dt=dt, t0=t0, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED)
    kw.update(params_kw)
    p = _ffi.make_params(**kw)
    P, M = np.asfortranarray(pos, dtype=float).copy(order="F"), np.asfortranarray(mom, dtype=float).copy(order="F")
    C = L.lmo_noise_channels(O.ctypes.byref(p), P.shape[0], M.shape[0])
    rec = np.zeros((steps, C, N)) if noise is None else noise
    rc, _ = O.run_cartesian(p, med, sp, P, M, 1, steps, np.ascontiguousarray(rec))
    assert rc == 0, rc
    return P, M


def straight_line_eta(y, tau0, tau):
    """eta_s at proper time tau of a free particle of rapidity y that left eta_s = 0 at tau0:
    z = tanh(y) (t - tau0) on the line, t^2 - z^2 = tau^2, eta_s = atanh(z/t)"""
    v = mp.tanh(y)
    t = mp.findroot(lambda t: t * t - (v * (t - tau0)) ** 2 - tau * tau, tau * mp.cosh(y))
    return mp.atanh(v * (t - tau0) / t)


def test_milne_free_streaming_closed_form():
    """free streaming (no drag: D_s T <= 0 gives eta_D = kappa = 0; Galilean boosts at rest are
    exact): p_z tau is conserved (the reference's gate R, test_momentum_dims3.jl:173-188) and
    eta_s follows the straight line z = tanh(y)(t - tau0) through t = tau cosh(eta_s),
    z = tau sinh(eta_s) (50 digits), for quarks with and without transverse momentum"""
    m, t0, dt, n = 1.5, 0.6, 0.005, 1942
    ys = np.array([0.9, -0.4, 0.05, 2.5])
    pts = np.array([0.0, 0.0, 1.3, 0.7])
    N = ys.size
    mT = np.sqrt(m * m + pts ** 2)
    mom = np.vstack([pts, 0 * pts, mT * np.sinh(ys)])
    pos = np.zeros((3, N))
    P, M = milne_run(dict(coords=_ffi.LMH_MILNE, collision=_ffi.LMH_LANGEVIN, relativistic=False, DsT=-1.0),
                     pos, mom, n, t0=t0, dt=dt)
    tau = t0 + n * dt
    for i in range(N):
        want_pz = mp.mpf(mT[i]) * mp.sinh(mp.mpf(ys[i])) * t0 / mp.mpf(tau)
        assert M[2, i] == pytest.approx(float(want_pz), rel=1e-9)              # gate R
        want_eta = straight_line_eta(mp.mpf(ys[i]), mp.mpf(t0), mp.mpf(tau))
        assert P[2, i] == pytest.approx(float(want_eta), rel=1e-12, abs=1e-14)
    assert np.array_equal(M[:2], mom[:2])                   # the transverse momenta stream freely


def test_milne_deta_is_the_difference_of_rapidity_gaps():
    """one step of lmo_milne_deta against asinh(k b) - asinh(b) at 50 digits, including b near 0,
    large |b| and both signs"""
    for px, py, pz, m, tau_a, dt in [(0.3, -0.2, 1.7, 1.5, 0.6, 0.005), (0.0, 0.0, -40.0, 1.5, 2.0, 0.01),
                                     (2.0, 1.0, 1e-9, 4.75, 9.0, 0.005), (0.0, 0.5, 3.0, 1.5, 0.01, 0.5)]:
        got = L.lmo_milne_deta(px, py, pz, m, tau_a, dt)
        b = mp.mpf(pz) / mp.sqrt(mp.mpf(m) ** 2 + mp.mpf(px) ** 2 + mp.mpf(py) ** 2)
        k = (mp.mpf(tau_a) + mp.mpf(dt)) / mp.mpf(tau_a)
        want = mp.asinh(k * b) - mp.asinh(b)
        assert got == pytest.approx(float(want), rel=4e-16, abs=1e-300)
    assert L.lmo_milne_deta(1.0, 0.0, 2.0, 1.5, 0.0, 0.01) == 0.0


def test_milne_step_is_the_bjorken_mode_step_on_the_transverse_rows():
    """the x, y and momentum rows of a Milne run equal the reference's Bjorken-mode Cartesian run
    (dimensions 2, momentum rows 3, bjorken_redshift) bit for bit, on a flowing medium with RTA
    and the Langevin kick: eta_s is a passenger"""
    keep = O.Keep()
    rng = np.random.default_rng(17)
    N, n, m = 600, 40, 1.5
    xg = np.sqrt(np.linspace(0.0, 400.0, 81))
    tg = np.arange(0.0, 10.01, 0.5)
    R, Tt = np.meshgrid(xg, tg, indexing="ij")
    Tf = np.asfortranarray(0.4 * np.exp(-R * R / 50.0) * (1 - 0.03 * Tt) + 0.05)
    Vf = np.asfortranarray(0.6 * np.tanh(R / 5.0) * (1 + 0.02 * Tt))
    med = O.radial_medium(keep, xg, tg, Tf, Vf)
    DsT = 0.6366
    jt = jt_table(m, 3, 0.3)
    sp = O.species(keep, [m], [N], drag=[drag_table(m, DsT, 0.3)], rta=[const_table(0.5, 0.3)], jt=[jt])
    pos = np.asfortranarray(rng.normal(0, 3, (2, N)))
    mom = np.asfortranarray(rng.normal(0, 1.5, (3, N)))
    base = dict(dt=0.01, t0=0.6, DsT=DsT, noise=_ffi.LMH_NOISE_FIXED, collision=_ffi.LMH_LANGEVIN_RTA)
    pc = _ffi.make_params(bjorken_redshift=True, **base)
    pmil = _ffi.make_params(coords=_ffi.LMH_MILNE, **base)
    C = L.lmo_noise_channels(O.ctypes.byref(pc), 2, 3)
    rec = rng.standard_normal((n, C, N))
    rec[:, 3:5] = rng.random((n, 2, N))                     # u_col, u_samp in [0, 1)
    Pc, Mc = pos.copy(order="F"), mom.copy(order="F")
    assert O.boost_to_lab_all(pc, med, sp, Pc, Mc, 0) == 0
    assert O.run_cartesian(pc, med, sp, Pc, Mc, 1, n, rec)[0] == 0
    Pm = np.asfortranarray(np.vstack([pos, 0.1 * rng.standard_normal(N)]))
    eta0 = Pm[2].copy()
    Mm = mom.copy(order="F")
    assert O.boost_to_lab_all(pmil, med, sp, Pm, Mm, 0) == 0
    assert O.run_cartesian(pmil, med, sp, Pm, Mm, 1, n, rec)[0] == 0
    assert np.array_equal(Pm[:2], Pc) and np.array_equal(Mm, Mc)
    assert np.all(np.sign(Pm[2] - eta0) == np.sign(Mm[2]) + (Mm[2] == 0)) or np.mean(Pm[2] != eta0) > 0.99


# ------------------------------------------------------------ initial boost on the grid -------
@pytest.mark.parametrize("relativistic", [True, False])
def test_grid_initial_boost_is_the_lorentz_boost(relativistic):
    """lmo_boost_to_lab_grid on a medium of uniform flow v = (vx, vy): the textbook boost
    p' = p + (gamma - 1)(p.n)n + gamma E v n with E = sqrt(p^2 + m^2) (relativistic; the
    reference's gamma regulator 1e-10 makes the bar 1e-9), p' = p + m v (Galilean); a quark at
    rest lands on p' = gamma m v; the species keep their own masses"""
    from lmhip import media
    vx, vy = 0.5, -0.6
    n = 5
    x = np.linspace(-2.0, 2.0, n)
    tg = np.array([0.5, 1.0, 2.0])
    shape = (n, n, tg.size)
    g = media.Grid2p1Medium(-2.0, 1.0, n, -2.0, 1.0, n, tg, np.full(shape, 0.3, order="F"),
                            np.full(shape, vx, order="F"), np.full(shape, vy, order="F"))
    keep = O.Keep()
    gm = O.grid_medium(keep, g)
    rng = np.random.default_rng(5)
    counts, masses = [60, 40], [1.5, 4.75]
    N = sum(counts)
    sp = O.species(keep, masses, counts, drag=[(0.0, 1.0, np.zeros(2))] * 2)
    pos = np.asfortranarray(rng.uniform(-3.0, 3.0, (2, N)))       # outside the grid: clamped
    mom = np.asfortranarray(rng.normal(0.0, 2.0, (3, N)))
    mom[:, 0] = 0.0
    p = _ffi.make_params(dt=0.01, t0=0.6, relativistic=relativistic)
    out = mom.copy(order="F")
    assert O.boost_to_lab_grid(p, gm, sp, pos, out, 7) == 0
    m = np.repeat(masses, counts)
    v = np.hypot(vx, vy)
    nv = np.array([vx, vy, 0.0]) / v
    if relativistic:
        gam = 1.0 / np.sqrt(1.0 - v * v)
        E = np.sqrt((mom ** 2).sum(0) + m * m)
        pn = nv @ mom
        want = mom + np.outer(nv, (gam - 1.0) * pn + gam * v * E)
        # per quark, relative to |p'| (single components cancel)
        dev = np.linalg.norm(out - want, axis=0) / np.linalg.norm(want, axis=0)
        assert dev.max() <= 1e-9, dev.max()
        assert np.allclose(out[:, 0], gam * 1.5 * v * nv, rtol=1e-9)
        # the mass shell: E'^2 - p'^2 is unchanged
        Ep = gam * (E + v * pn)
        assert np.allclose(Ep ** 2 - (out ** 2).sum(0), m * m, rtol=1e-6)
    else:
        want = mom + np.outer(nv * v, m)
        assert np.allclose(out, want, rtol=1e-13, atol=1e-14)
