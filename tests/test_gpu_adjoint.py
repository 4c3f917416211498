"""Pathwise adjoint of the 2+1D step (lmh_adjoint_forward / lmh_adjoint_backward).

The reference gradient is the oracle (the C restatement of the CPU kernels) differentiated by
central finite differences at fixed Philox key: the same noise record feeds every evaluation, so
the finite difference is the derivative of the CPU path along the adjoint's noise path.  With a
relative step of 1e-5 no grid-cell crossing, position clamp or collision decision changes between
the two evaluations at these sizes, so the finite difference is the smooth pathwise derivative to
O(h^2).  Bar: |adjoint - FD| <= 1e-4 |FD| + 1e-10 (the adjoint-vs-FD tolerance of the
differentiation workload).
"""
from __future__ import annotations

import numpy as np
import pytest

import _oracle as O
import lmhip
from helpers import small_grid
from lmhip import _ffi, media, transport
from lmhip.adjoint import GridEnsemble

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
DT, T0 = 0.02, 0.6
OBS = dict(nbins=12, pt_min=0.0, pt_max=6.0, sigma=0.5)


def _setup(N=400, seed=5, two_species=True):
    g = small_grid()
    rng = np.random.default_rng(seed)
    pos = media.hydro_positions(N, rng) * 2.0
    if two_species:
        nc = int(0.7 * N)
        counts = [nc, N - nc]
        masses = [1.5, 4.75]
        mom = np.asfortranarray(np.hstack([media.pt_rapidity_momenta(nc, 1.5, rng),
                                           media.pt_rapidity_momenta(N - nc, 4.75, rng)]))
    else:
        counts, masses = [N], [1.5]
        mom = media.pt_rapidity_momenta(N, 1.5, rng)
    return g, masses, counts, pos, mom


def _tminmax(g):
    Tmin = max(float(np.min(g.T)), 0.0)
    return Tmin, max(float(np.max(g.T)), Tmin + EPS)


def _rta_tables(g, masses):
    Tmin, Tmax = _tminmax(g)
    rt, jt = [], []
    for m in masses:
        rt.append(transport.build_relaxation_spline(media.tau_R_law(0.16), Tmin=Tmin, Tmax=Tmax, nT=1024))
        jt.append(lmhip.build_juttner_invcdf(m, 3, Tmin=Tmin, Tmax=Tmax, nT=64, nU=512, np_=2048))
    return rt, jt


def _cotangents(N, nbins, seed=3):
    rng = np.random.default_rng(seed)
    w_pt = np.zeros(5 + nbins); w_pt[1] = 1.0 / N
    w_c2 = np.zeros(5 + nbins); w_c2[3] = 1.0 / N
    w_sp = np.zeros(5 + nbins); w_sp[5:] = rng.normal(size=nbins) / N
    w_all = rng.normal(size=5 + nbins) / N
    w_all[0] = 0.0
    return {"mean_pt": w_pt, "cos2phi": w_c2, "spectrum": w_sp, "mixed": w_all}


class OracleEnsemble:
    """The same ensemble through the oracle: observables as a function of theta."""

    def __init__(self, g, masses, counts, pos, mom, steps, drag_model, collision, seed, rta):
        self.g, self.masses, self.counts = g, masses, counts
        self.pos, self.mom, self.steps = pos, mom, steps
        self.drag_model, self.collision, self.seed = drag_model, collision, seed
        self.rt, self.jt = _rta_tables(g, masses) if rta else (None, None)
        p = self.params((0.6, 5.3, 5.0))
        self.noise = np.ascontiguousarray(O.noise_record(p, 2, 3, pos.shape[1], 1, steps))

    def params(self, theta):
        return _ffi.make_params(dt=DT, t0=T0, DsT=theta[0], gamma0=theta[1], Lambda=theta[2],
                                seed=self.seed, collision=self.collision,
                                noise=_ffi.LMH_NOISE_PHILOX, drag_model=self.drag_model,
                                check_finite=False)

    def obs(self, theta, pos=None, mom=None, only=None):
        """the observables of the final state; only = i: of particle i alone (the particles are
        independent, so a change of particle i's initial state changes these and nothing else)"""
        p = self.params(theta)
        Tmin, Tmax = _tminmax(self.g)
        drag = [transport.build_tau_drag_spline(m, theta[0], Tmin=Tmin, Tmax=Tmax, nT=1024)
                for m in self.masses]
        keep = O.Keep()
        gm = O.grid_medium(keep, self.g)
        sp = O.species(keep, self.masses, self.counts, drag=drag, rta=self.rt, jt=self.jt)
        x = (self.pos if pos is None else pos).copy(order="F")
        q = (self.mom if mom is None else mom).copy(order="F")
        rc, bad = O.run_grid(p, gm, sp, x, q, 1, self.steps, self.noise)
        assert rc == 0, (rc, bad)
        i0, i1 = (0, x.shape[1]) if only is None else (only, only + 1)
        return O.observables(q, i0, i1, OBS["nbins"], OBS["pt_min"], OBS["pt_max"], OBS["sigma"])


def _gpu_ensemble(ctx, g, masses, counts, pos, mom, steps, drag_model, collision, seed, K, rta):
    rt, jt = _rta_tables(g, masses) if rta else (None, None)
    return GridEnsemble(ctx, g, masses, counts, pos, mom, dt=DT, t0=T0, nsteps=steps,
                        drag_model=drag_model, collision=collision, seed=seed,
                        checkpoint_every=K, rta_tables=rt, juttner_tables=jt,
                        boost_initial=False, **OBS)


def _close(adj, fd, what):
    err = np.abs(adj - fd)
    tol = 1e-4 * np.abs(fd) + 1e-10
    assert np.all(err <= tol), f"{what}: adjoint {adj} fd {fd} err {err}"


CASES = {
    "table_langevin": (_ffi.LMH_DRAG_TABLE, _ffi.LMH_LANGEVIN, False),
    "table_langevin_rta": (_ffi.LMH_DRAG_TABLE, _ffi.LMH_LANGEVIN_RTA, True),
    "gammap_langevin": (_ffi.LMH_DRAG_GAMMA_P, _ffi.LMH_LANGEVIN, False),
    "gammap_langevin_rta": (_ffi.LMH_DRAG_GAMMA_P, _ffi.LMH_LANGEVIN_RTA, True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_matches_oracle_finite_differences(gpu_ok, case):
    drag_model, collision, rta = CASES[case]
    steps, K, seed = 70, 16, 11
    g, masses, counts, pos, mom = _setup()
    theta = np.array([0.6366, 5.3, 5.0])
    orc = OracleEnsemble(g, masses, counts, pos, mom, steps, drag_model, collision, seed, rta)
    ws = _cotangents(pos.shape[1], OBS["nbins"])
    # central differences of the oracle, relative step 1e-5
    fd_obs = []
    for k in range(3):
        h = 1e-5 * theta[k]
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd_obs.append((orc.obs(tp) - orc.obs(tm)) / (2 * h))
    fd_obs = np.array(fd_obs)                       # (3, 5 + nbins)
    with lmhip.Context(0) as ctx:
        ens = _gpu_ensemble(ctx, g, masses, counts, pos, mom, steps, drag_model, collision, seed, K, rta)
        obs = ens.forward(theta)
        # the forward half reproduces the oracle's observables
        o_obs = orc.obs(theta)
        assert np.allclose(obs, o_obs, rtol=1e-12, atol=1e-12)
        for name, w in ws.items():
            grad = ens.pullback(w)
            _close(grad, fd_obs @ w, f"{case}/{name}")
        if drag_model == _ffi.LMH_DRAG_TABLE:
            assert np.all(fd_obs[1:] == 0.0)
        else:
            assert np.all(fd_obs[0] == 0.0)


def test_adjoint_forward_state_is_the_run_state(gpu_ok):
    """lmh_adjoint_forward leaves the state lmh_run leaves with the default tuning (one step per
    launch, sliced medium), bit for bit, for any checkpoint cadence and whatever tuning is set"""
    g, masses, counts, pos, mom = _setup(N=3000)
    rt, jt = _rta_tables(g, masses)
    theta = (0.6366, 5.3, 5.0)
    res = []
    with lmhip.Context(0) as ctx:
        for K in (0, 1, 7, 40, -7):
            if K < 0:
                ctx.set_tuning(block=256, steps_per_launch=4, slice=0)
            ens = GridEnsemble(ctx, g, masses, counts, pos, mom, dt=DT, t0=T0, nsteps=40,
                               collision=_ffi.LMH_LANGEVIN_RTA, seed=3, checkpoint_every=max(abs(K), 1),
                               rta_tables=rt, juttner_tables=jt, **OBS)
            if K == 0:
                ens.run_plain(theta)
            else:
                ens.forward(theta)
            res.append(ctx.get_particles())
        # the tuning set by hand is kept
        assert ctx.get_tuning()["steps_per_launch"] == 4 and ctx.get_tuning()["slice"] == 0
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0]) and np.array_equal(r[1], res[0][1])


def test_adjoint_checkpoint_cadence_and_repeat(gpu_ok):
    """the gradient does not depend on the checkpoint cadence beyond rounding, and a second
    pullback with the same cotangent gives the same bits"""
    g, masses, counts, pos, mom = _setup(N=2000)
    w = _cotangents(2000, OBS["nbins"])["mixed"]
    theta = (0.6366, 5.3, 5.0)
    grads = []
    with lmhip.Context(0) as ctx:
        for K in (1, 9, 50):
            ens = _gpu_ensemble(ctx, g, masses, counts, pos, mom, 50, _ffi.LMH_DRAG_GAMMA_P,
                                _ffi.LMH_LANGEVIN, 4, K, False)
            ens.forward(theta)
            g1 = ens.pullback(w)
            g2 = ens.pullback(w)
            assert np.array_equal(g1, g2)
            grads.append(g1)
    for gk in grads[1:]:
        assert np.allclose(gk, grads[0], rtol=1e-12, atol=0)


def test_adjoint_state_gradient_matches_oracle(gpu_ok):
    """dL/d(initial momentum and position) of single particles against oracle differences"""
    steps, K, seed = 40, 12, 21
    g, masses, counts, pos, mom = _setup(N=200)
    theta = np.array([0.6366, 5.3, 5.0])
    orc = OracleEnsemble(g, masses, counts, pos, mom, steps, _ffi.LMH_DRAG_TABLE, _ffi.LMH_LANGEVIN, seed, False)
    w = _cotangents(200, OBS["nbins"])["mixed"]
    with lmhip.Context(0) as ctx:
        ens = _gpu_ensemble(ctx, g, masses, counts, pos, mom, steps, _ffi.LMH_DRAG_TABLE,
                            _ffi.LMH_LANGEVIN, seed, K, False)
        ens.forward(theta)
        _, dpos, dmom = ens.pullback(w, state_grad=True)
    # the reference: central differences of particle i's own observables (the other particles'
    # sums, unchanged, would add their rounding to the difference) at h = 1e-5, where the
    # difference quotients have converged: below 1e-6 the trajectory's rounding dominates them
    # (~1e-10 on a 2.5e-7 derivative), above 3e-5 the +-h paths straddle the piecewise-linear
    # medium's cell edges
    def cd(fn, base, d, i, h=1e-5):
        a, b = base.copy(order="F"), base.copy(order="F")
        a[d, i] += h
        b[d, i] -= h
        return (fn(a) - fn(b)) @ w / (2 * h)
    for i in (0, 57, 150, 199):
        for d in range(3):
            fd = cd(lambda m_: orc.obs(theta, mom=m_, only=i), mom, d, i)
            assert abs(dmom[d, i] - fd) <= 1e-4 * abs(fd) + 1e-10, (i, d, dmom[d, i], fd)
        for d in range(2):
            fd = cd(lambda p_: orc.obs(theta, pos=p_, only=i), pos, d, i)
            assert abs(dpos[d, i] - fd) <= 1e-4 * abs(fd) + 1e-10, (i, d, dpos[d, i], fd)


def test_adjoint_gpu_finite_differences_larger(gpu_ok):
    """2e4 quarks x 150 steps with collisions: the adjoint against central differences of the
    plain GPU run (lmh_run) at relative step 1e-4"""
    steps, K, seed = 150, 50, 17
    g, masses, counts, pos, mom = _setup(N=20000, seed=9)
    w = _cotangents(20000, OBS["nbins"])["mean_pt"]
    rt, jt = _rta_tables(g, masses)
    for drag_model, theta in ((_ffi.LMH_DRAG_TABLE, np.array([0.6366, 0.0, 5.0])),
                              (_ffi.LMH_DRAG_GAMMA_P, np.array([0.6366, 5.3, 5.0]))):
        with lmhip.Context(0) as ctx:
            ens = GridEnsemble(ctx, g, masses, counts, pos, mom, dt=DT, t0=T0, nsteps=steps,
                               drag_model=drag_model, collision=_ffi.LMH_LANGEVIN, seed=seed,
                               checkpoint_every=K, rta_tables=rt, juttner_tables=jt, **OBS)
            ens.forward(theta)
            grad = ens.pullback(w)
            ks = (0,) if drag_model == _ffi.LMH_DRAG_TABLE else (1, 2)
            for k in ks:
                h = 1e-4 * theta[k]
                tp, tm = theta.copy(), theta.copy()
                tp[k] += h
                tm[k] -= h
                fd = (ens.run_plain(tp) - ens.run_plain(tm)) @ w / (2 * h)
                assert abs(grad[k] - fd) <= 1e-3 * abs(fd) + 1e-9, (drag_model, k, grad[k], fd)


def test_adjoint_backward_needs_its_forward_state(gpu_ok):
    g, masses, counts, pos, mom = _setup(N=500)
    w = _cotangents(500, OBS["nbins"])["mean_pt"]
    with lmhip.Context(0) as ctx:
        ens = _gpu_ensemble(ctx, g, masses, counts, pos, mom, 10, _ffi.LMH_DRAG_TABLE,
                            _ffi.LMH_LANGEVIN, 1, 4, False)
        with pytest.raises(lmhip.LMHError):
            ctx.adjoint_backward(w, **OBS)
        ens.forward((0.6366, 0.0, 5.0))
        g0 = ens.pullback(w)
        ctx.run(ens.params((0.6366, 0.0, 5.0)), 11, 1)
        # the library refuses a record whose state changed; the ensemble's pullback runs its
        # forward pass again and gives the same gradient
        with pytest.raises(lmhip.LMHError, match="changed since"):
            ctx.adjoint_backward(w, **OBS)
        assert np.array_equal(ens.pullback(w), g0)
        bad = _ffi.make_params(dt=DT, t0=T0, DsT=0.6366, noise=_ffi.LMH_NOISE_FIXED)
        with pytest.raises(lmhip.LMHError, match="Philox"):
            ctx.adjoint_forward(bad, 1, 5, 2)


@pytest.mark.parametrize("drag_model,rta", [(_ffi.LMH_DRAG_TABLE, True), (_ffi.LMH_DRAG_GAMMA_P, False)])
def test_adjoint_keep_tape_equals_checkpointed(gpu_ok, drag_model, rta):
    """keep_tape: the forward pass writes the whole tape; its final state is lmh_run's and its
    gradients (parameters and initial state) are the checkpointed pullback's, bit for bit"""
    N, steps = 3000, 45
    g, masses, counts, pos, mom = _setup(N=N)
    rt, jt = _rta_tables(g, masses) if rta else (None, None)
    coll = _ffi.LMH_LANGEVIN_RTA if rta else _ffi.LMH_LANGEVIN
    w = _cotangents(N, OBS["nbins"])["mixed"]
    theta = (0.6366, 5.3, 5.0)
    out = {}
    with lmhip.Context(0) as ctx:
        for keep in (False, True):
            ens = GridEnsemble(ctx, g, masses, counts, pos, mom, dt=DT, t0=T0, nsteps=steps,
                               drag_model=drag_model, collision=coll, seed=8, checkpoint_every=10,
                               rta_tables=rt, juttner_tables=jt, keep_tape=keep, **OBS)
            obs = ens.forward(theta)
            state = ctx.get_particles()
            grad, dpos, dmom = ens.pullback(w, state_grad=True)
            out[keep] = (obs, state[0], state[1], grad, dpos, dmom, ens.pullback(w))
        ens.run_plain(theta)
        plain = ctx.get_particles()
        # the whole tape is not kept under freeze-out capture
        ctx.freezeout_enable(0.2)
        with pytest.raises(lmhip.LMHError, match="freeze-out"):
            ens.forward(theta)
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[True][1], plain[0]) and np.array_equal(out[True][2], plain[1])


def test_adjoint_cell_order_independent(gpu_ok):
    """the checkpoints keep the cell order their segment runs in and the backward pass re-runs and
    reverses each segment in it: the gradients (parameters and initial state) do not depend on the
    sort cadence -- never, mid-segment, at every checkpoint, every other -- bit for bit, and equal
    the whole-tape pullback's (caller's order throughout)"""
    N, steps, K = 3000, 45, 10
    g, masses, counts, pos, mom = _setup(N=N)
    rt, jt = _rta_tables(g, masses)
    w = _cotangents(N, OBS["nbins"])["mixed"]
    theta = (0.6366, 5.3, 5.0)
    out = []
    with lmhip.Context(0) as ctx:
        for sort_every, keep in ((0, False), (3, False), (10, False), (20, False), (3, True)):
            ctx.set_tuning(block=256, steps_per_launch=1, slice=1, defer=0, sort_every=sort_every)
            ens = GridEnsemble(ctx, g, masses, counts, pos, mom, dt=DT, t0=T0, nsteps=steps,
                               drag_model=_ffi.LMH_DRAG_TABLE, collision=_ffi.LMH_LANGEVIN_RTA, seed=8,
                               checkpoint_every=K, rta_tables=rt, juttner_tables=jt, keep_tape=keep, **OBS)
            obs = ens.forward(theta)
            state = ctx.get_particles()
            grad, dpos, dmom = ens.pullback(w, state_grad=True)
            out.append((obs, state[0], state[1], grad, dpos, dmom))
    for r in out[1:]:
        for a, b in zip(out[0], r):
            assert np.array_equal(a, b)


def test_torch_autograd_function(gpu_ok):
    import torch
    from lmhip.adjoint import observables
    g, masses, counts, pos, mom = _setup(N=1000)
    with lmhip.Context(0) as ctx:
        ens = _gpu_ensemble(ctx, g, masses, counts, pos, mom, 20, _ffi.LMH_DRAG_GAMMA_P,
                            _ffi.LMH_LANGEVIN, 2, 8, False)
        theta = torch.tensor([0.6366, 5.3, 5.0], dtype=torch.float64, requires_grad=True)
        obs = observables(theta, ens)
        loss = obs[1] / obs[0] + 0.5 * obs[3] / obs[0]
        loss.backward()
        w = np.zeros(5 + OBS["nbins"])
        n = float(obs[0].detach())
        w[1], w[3] = 1.0 / n, 0.5 / n
        ref = ens.pullback(w)
        assert np.allclose(theta.grad.numpy(), ref, rtol=1e-14, atol=0)


def test_two_forwards_before_backward(gpu_ok):
    """two autograd nodes at different theta whose forwards both run before either backward (on
    one ensemble, and on two ensembles of one context): each node gets its own theta's gradient,
    bit for bit the gradient of a forward + pullback at that theta alone"""
    import torch
    from lmhip.adjoint import observables
    g, masses, counts, pos, mom = _setup(N=1000)
    th = [np.array([0.6366, 5.3, 5.0]), np.array([0.8, 4.1, 3.5])]
    w = np.zeros(5 + OBS["nbins"])
    with lmhip.Context(0) as ctx:
        ens = [_gpu_ensemble(ctx, g, masses, counts, pos, mom, 20, _ffi.LMH_DRAG_GAMMA_P,
                             _ffi.LMH_LANGEVIN, 2, 8, False) for _ in range(2)]
        alone = []
        for k in range(2):
            obs = ens[0].forward(th[k])
            w[:] = 0.0
            w[1] = 1.0 / obs[0]
            alone.append(ens[0].pullback(w))
        for owners in ((0, 0), (0, 1)):
            t = [torch.tensor(th[k], dtype=torch.float64, requires_grad=True) for k in range(2)]
            obs = [observables(t[k], ens[owners[k]]) for k in range(2)]
            (obs[0][1] / obs[0][0] + obs[1][1] / obs[1][0]).backward()
            for k in range(2):
                assert np.array_equal(t[k].grad.numpy(), alone[k]), (owners, k)
