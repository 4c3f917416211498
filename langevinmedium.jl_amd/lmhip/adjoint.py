"""Differentiable final-state observables of a 2+1D ensemble (extension).

`GridEnsemble` holds a fixed initial ensemble, a 2+1D medium and the step options;
`GridEnsemble.forward(theta)` runs the ensemble for `nsteps` steps with
theta = (DsT, gamma0, Lambda) and returns the observables of the final lab momenta (the
`lmh_observables` vector), keeping checkpoints; `GridEnsemble.pullback(cotangent)` returns
dL/dtheta for L = cotangent . observables from the pathwise adjoint (liblmhip's
`lmh_adjoint_backward`).  `observables(theta, ensemble)` wraps the pair as a
torch.autograd.Function, the PyTorch counterpart of a ChainRulesCore.rrule.  With
keep_tape=True the forward pass keeps every step's tape row in device memory (nsteps x
(2 + 2 pdim [+1]) x N doubles) and the pullback runs the reverse sweeps only.

DsT enters through the drag-time table (rebuilt for each DsT, `transport.jl:165-209`);
gamma0 and Lambda through the GAMMA_P drag eta = gamma0 T^2/(1 + p^2/Lambda^2).  The noise is
Philox with a fixed key, so forward runs at different theta share their random numbers
(common random numbers) and central finite differences at fixed seed are meaningful.
"""
from __future__ import annotations

import numpy as np

from . import _ffi, transport
from .context import Context


class GridEnsemble:
    def __init__(self, ctx: Context, medium, masses, counts, pos, mom, *, dt, t0, nsteps,
                 drag_model=_ffi.LMH_DRAG_TABLE, collision=_ffi.LMH_LANGEVIN, seed=0, stream=0,
                 checkpoint_every=50, nbins=40, pt_min=0.0, pt_max=10.0, sigma=0.0, species=-1,
                 rta_tables=None, juttner_tables=None, nT=1024, boost_initial=True,
                 bjorken_redshift=False, keep_tape=False):
        self.ctx = ctx
        self.medium = medium
        self.masses = [float(m) for m in masses]
        self.counts = [int(c) for c in counts]
        self.pos0 = np.asfortranarray(pos, dtype=np.float64)
        self.mom0 = np.asfortranarray(mom, dtype=np.float64)
        self.dt, self.t0, self.nsteps = float(dt), float(t0), int(nsteps)
        self.drag_model, self.collision = int(drag_model), int(collision)
        self.seed, self.stream = int(seed), int(stream)
        self.K = int(checkpoint_every)
        self.keep_tape = bool(keep_tape)
        self.obs = dict(nbins=int(nbins), pt_min=float(pt_min), pt_max=float(pt_max),
                        sigma=float(sigma), species=int(species))
        self.nT = int(nT)
        self.boost_initial = bool(boost_initial)
        self.bjorken_redshift = bool(bjorken_redshift)
        T = np.asarray(medium.T)
        eps = np.finfo(float).eps
        self.Tmin = max(float(np.min(T)), 0.0)
        self.Tmax = max(float(np.max(T)), self.Tmin + eps)
        g = medium
        ctx.set_medium_grid2p1(g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, g.tgrid, g.T, g.vx, g.vy)
        for k in range(len(self.masses)):
            if rta_tables is not None:
                ctx.set_time_table(_ffi.LMH_TABLE_RTA, k, *rta_tables[k])
            if juttner_tables is not None:
                ctx.set_juttner_table(k, *juttner_tables[k])
        self._dst_loaded = None
        self._theta = None

    def params(self, theta):
        DsT, gamma0, Lambda = (float(v) for v in theta)
        return _ffi.make_params(dt=self.dt, t0=self.t0, DsT=DsT, gamma0=gamma0, Lambda=Lambda,
                                seed=self.seed, stream=self.stream, collision=self.collision,
                                noise=_ffi.LMH_NOISE_PHILOX, drag_model=self.drag_model,
                                bjorken_redshift=self.bjorken_redshift, check_finite=False)

    def _load_drag(self, DsT):
        if self.drag_model != _ffi.LMH_DRAG_TABLE or DsT == self._dst_loaded:
            return
        for k, m in enumerate(self.masses):
            tab = transport.build_tau_drag_spline(m, DsT, Tmin=self.Tmin, Tmax=self.Tmax, nT=self.nT)
            self.ctx.set_time_table(_ffi.LMH_TABLE_DRAG, k, *tab)
        self._dst_loaded = DsT

    def reset(self, theta):
        """Upload the initial ensemble and the theta-dependent tables."""
        p = self.params(theta)
        self._load_drag(p.DsT)
        self.ctx.set_particles(self.pos0, self.mom0, self.masses, self.counts)
        if self.boost_initial:
            self.ctx.boost_to_lab(p, 0)
        return p

    def run_plain(self, theta):
        """The same steps through lmh_run (no checkpoints): the final observables."""
        p = self.reset(theta)
        self.ctx.run(p, 1, self.nsteps, None, 0)
        return self.ctx.observables(**self.obs)

    def forward(self, theta):
        p = self.reset(theta)
        self.ctx.adjoint_forward(p, 1, self.nsteps, self.K, keep_tape=self.keep_tape)
        self._theta = tuple(float(v) for v in theta)
        return self.ctx.observables(**self.obs)

    def pullback(self, cotangent, state_grad=False):
        if self._theta is None:
            raise RuntimeError("pullback before forward")
        return self.ctx.adjoint_backward(np.asarray(cotangent, dtype=np.float64), state_grad=state_grad,
                                         **self.obs)


def _torch_function():
    import torch

    class _Observables(torch.autograd.Function):
        @staticmethod
        def forward(fctx, theta, ensemble):
            obs = ensemble.forward(theta.detach().cpu().numpy())
            fctx.ensemble = ensemble
            return torch.from_numpy(obs).to(theta.dtype)

        @staticmethod
        def backward(fctx, gobs):
            g = fctx.ensemble.pullback(gobs.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g).to(gobs.dtype), None

    return _Observables


_FN = None


def observables(theta, ensemble: GridEnsemble):
    """Final-state observables of `ensemble` as a differentiable function of the tensor
    theta = (DsT, gamma0, Lambda)."""
    global _FN
    if _FN is None:
        _FN = _torch_function()
    return _FN.apply(theta, ensemble)
