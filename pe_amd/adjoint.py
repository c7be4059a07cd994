"""Differentiable time step: the adjoint of ``pehip_adjoint_*`` wrapped as a
``torch.autograd.Function``.

``DifferentiableRun`` owns a ``World`` built from a scene, runs it forward with checkpoints and
runs the reverse pass on the GPU.  ``FinalPositions.apply(v0, mu, run)`` maps the initial sphere
velocities and the friction coefficient to the final body positions; its backward calls the
adjoint.  The gradient is the derivative of the converged step at fixed active set, so the
forward runs each step's sweeps to a tight threshold.
"""
import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import check, ptr
from .world import World, ApproximateInelasticCoulombContactByDecoupling


class DifferentiableRun:
    def __init__(self, scene, steps, dt=None, checkpoint_every=20, model=ApproximateInelasticCoulombContactByDecoupling,
                 max_iterations=4000, abort_threshold=1e-13, adjoint_sweeps=4000, adjoint_tol=1e-13, device=0,
                 max_sub_iterations=20):
        self.scene = scene
        self.steps = int(steps)
        self.dt = scene.dt if dt is None else dt
        self.every = int(checkpoint_every)
        self.model = model
        self.max_iterations = max_iterations
        self.abort_threshold = abort_threshold
        self.adjoint_sweeps = adjoint_sweeps
        self.adjoint_tol = adjoint_tol
        self.device = device
        self.max_sub_iterations = int(max_sub_iterations)
        self.world = None
        self.last_stats = None

    def make_world(self, v0, mu):
        """A World of the scene with sphere velocities v0 (n_spheres x 3) and pair friction mu."""
        w = World(device=self.device)
        self.scene.to_world(w)
        w.setMaxIterations(self.max_iterations)
        w.setRelaxationModel(self.model)
        w.setAbortThreshold(self.abort_threshold)
        w.setMaxSubIterations(self.max_sub_iterations)
        nm = len(self.scene.materials)
        for i in range(nm):
            for j in range(nm):
                w.setPairFriction(i, j, float(mu), float(mu))
        a = w.host_arrays()
        npl = len(self.scene.planes)
        a["v"][npl:npl + len(v0)] = np.asarray(v0, float)
        return w

    def forward(self, v0, mu):
        """Runs the scene forward; returns the final state (dict of arrays)."""
        if self.world is not None:
            self.world.close()
        self.world = self.make_world(v0, mu)
        w = self.world
        w._ensure_device()
        check(_abi.load().pehip_adjoint_forward(w._ctx, float(self.dt), self.steps, self.every), w._ctx)
        w._state_on_device = True
        return w.getState()

    def backward(self, dL_dpos, dL_dv=None, dL_dw=None):
        """dL/d(final pos, v, w) -> (dL/d pos0, dL/d v0, dL/d w0, dL/d mu) for all bodies."""
        w = self.world
        n = w.size
        gp = np.ascontiguousarray(dL_dpos, float).reshape(n, 3)
        gv = None if dL_dv is None else np.ascontiguousarray(dL_dv, float).reshape(n, 3)
        gw = None if dL_dw is None else np.ascontiguousarray(dL_dw, float).reshape(n, 3)
        out = {"pos": np.zeros((n, 3)), "v": np.zeros((n, 3)), "w": np.zeros((n, 3))}
        dmu = np.zeros(1)
        st = _abi.AdjointStats()
        check(_abi.load().pehip_adjoint_backward(w._ctx, ptr(gp), ptr(gv), ptr(gw), ptr(out["pos"]), ptr(out["v"]),
                                                 ptr(out["w"]), ptr(dmu), int(self.adjoint_sweeps),
                                                 float(self.adjoint_tol), ctypes.byref(st)), w._ctx)
        self.last_stats = {k: getattr(st, k) for k, _ in st._fields_}
        return out["pos"], out["v"], out["w"], float(dmu[0])

    def close(self):
        if self.world is not None:
            self.world.close()
            self.world = None


class FinalPositions(torch.autograd.Function):
    """(v0 [n_spheres x 3], mu [scalar]) -> final positions of all bodies [n x 3] (float64, CPU)."""

    @staticmethod
    def forward(ctx, v0, mu, run):
        st = run.forward(v0.detach().cpu().numpy(), float(mu.detach().cpu()))
        ctx.run = run
        ctx.nv = v0.shape[0]
        ctx.v0_dtype, ctx.v0_device = v0.dtype, v0.device
        return torch.from_numpy(st["pos"].copy())

    @staticmethod
    def backward(ctx, grad_pos):
        run = ctx.run
        _, gv, _, gmu = run.backward(grad_pos.detach().cpu().numpy())
        npl = len(run.scene.planes)
        g_v0 = torch.from_numpy(gv[npl:npl + ctx.nv].copy()).to(ctx.v0_dtype).to(ctx.v0_device)
        return g_v0, torch.tensor(gmu, dtype=torch.float64), None
