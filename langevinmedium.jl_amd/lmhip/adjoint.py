"""Differentiable final-state observables of a 2+1D ensemble (extension).

`GridEnsemble` holds a fixed initial ensemble, a 2+1D medium and the step options;
`GridEnsemble.forward(theta)` runs the ensemble for `nsteps` steps with
theta = (DsT, gamma0, Lambda) and returns the observables of the final lab momenta (the
`lmh_observables` vector), keeping checkpoints; `GridEnsemble.pullback(cotangent)` returns
dL/dtheta for L = cotangent . observables from the pathwise adjoint (liblmhip's
`lmh_adjoint_backward`).  `observables(theta, ensemble)` wraps the pair as a
torch.autograd.Function, the PyTorch counterpart of a ChainRulesCore.rrule.  With
keep_tape=True the forward pass keeps every step's tape row in device memory (nsteps x
(2 + pdim) x N doubles: the state; the noise is drawn again from its Philox counters) and the
pullback runs the reverse sweeps only.

A context holds one adjoint record, the last forward pass's.  Each forward pass gets a record
id; a pullback that names the forward it belongs to (theta and record, as the autograd node
keeps them) re-runs that forward first when the context's record is another's (a second
forward before the backward, on this or another ensemble of the context, or a plain run that
replaced the state), so every node gets its own theta's gradient.

DsT enters through the drag-time table (rebuilt for each DsT, `transport.jl:165-209`);
gamma0 and Lambda through the GAMMA_P drag eta = gamma0 T^2/(1 + p^2/Lambda^2).  The noise is
Philox with a fixed key, so forward runs at different theta share their random numbers
(common random numbers) and central finite differences at fixed seed are meaningful.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import _ffi, transport
from .context import Context, LMHError

_RECORDS = itertools.count(1)


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
        self.last_record = None

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
        # the context's adjoint record no longer describes its state
        self.ctx._adjoint_record = None
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
        """The final observables at theta; keeps the context's adjoint record (its id in
        last_record)."""
        p = self.reset(theta)
        self.run_forward(p, theta)
        return self.ctx.observables(**self.obs)

    def run_forward(self, p, theta, keep_tape=None):
        """The checkpointed forward pass from the state reset(theta) uploaded (p its
        parameters); registers the context's adjoint record as this ensemble's."""
        keep = self.keep_tape if keep_tape is None else bool(keep_tape)
        self.ctx.adjoint_forward(p, 1, self.nsteps, self.K, keep_tape=keep)
        self.keep_tape = keep
        self._theta = tuple(float(v) for v in theta)
        self.last_record = next(_RECORDS)
        self.ctx._adjoint_record = (id(self), self.last_record)

    def pullback(self, cotangent, state_grad=False, theta=None, record=None):
        """dL/dtheta (and with state_grad the state gradients) of the forward pass `record` at
        `theta`; without them, of the last forward pass of this ensemble.  When the context's
        record is not that pass's, the pass is run again at its theta first."""
        if record is None:
            if self._theta is None:
                raise RuntimeError("pullback before forward")
            theta, record = self._theta, self.last_record
        w = np.asarray(cotangent, dtype=np.float64)
        if getattr(self.ctx, "_adjoint_record", None) != (id(self), record):
            self._rerun(theta, record)
        try:
            return self.ctx.adjoint_backward(w, state_grad=state_grad, **self.obs)
        except LMHError as e:
            # the library found the record stale (the state, medium or tables changed since)
            if e.code != _ffi.LMH_E_STATE:
                raise
            self._rerun(theta, record)
            return self.ctx.adjoint_backward(w, state_grad=state_grad, **self.obs)

    def _rerun(self, theta, record):
        self.forward(theta)
        self.last_record = record
        self.ctx._adjoint_record = (id(self), record)


def _torch_function():
    import torch

    class _Observables(torch.autograd.Function):
        @staticmethod
        def forward(fctx, theta, ensemble):
            th = theta.detach().cpu().numpy().astype(np.float64)
            obs = ensemble.forward(th)
            fctx.ensemble, fctx.theta, fctx.record = ensemble, th.copy(), ensemble.last_record
            return torch.from_numpy(obs).to(theta.dtype)

        @staticmethod
        def backward(fctx, gobs):
            g = fctx.ensemble.pullback(gobs.detach().cpu().numpy().astype(np.float64),
                                       theta=fctx.theta, record=fctx.record)
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
