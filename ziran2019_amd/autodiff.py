"""torch.autograd binding of the differentiable explicit MLS rollout.

A thin test/driver binding over the C-ABI (zr_adjoint_*, zr_checkpoint_*):
the forward pass runs the GPU step ``steps`` times from (X0, V0, C0, F0) with
per-particle Young's modulus E (fixed-corotated, Poisson ratio nu), saving a
device checkpoint every ``checkpoint_every`` steps; the backward pass walks the
segments in reverse, recomputing each segment's step inputs from its
checkpoint and running the hand-written adjoint kernels step by step.  The
reference has no differentiable path; the gradients are pinned by FP64
central differences and by the CPU adjoint oracle in the tests.
"""
import numpy as np
import torch

from .engine import MpmEngine, grid_log2_for


def lame_from_E(E, nu):
    """StvkWithHencky/NeoHookean setLameParameters (StvkWithHencky.cpp:16-21)."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lam


class Rollout:
    """One engine holding a rollout problem: fixed X0, C0, F0, mass, volume,
    colliders and dt; E and V0 are the differentiable inputs."""

    def __init__(self, scene, steps, dt, nu=0.3, checkpoint_every=8, colliders=None):
        self.scene = scene
        self.steps = int(steps)
        self.dt = float(dt)
        self.nu = float(nu)
        self.every = int(checkpoint_every)
        self.colliders = list(scene.get("colliders", []) if colliders is None else colliders)
        self.eng = MpmEngine(dx=scene["dx"], gravity=scene.get("gravity", (0.0, -9.8, 0.0)),
                             grid_log2=grid_log2_for(scene["dx"]))
        self.nseg = (self.steps + self.every - 1) // self.every

    def close(self):
        self.eng.close()

    def forward(self, E, V0):
        """Final positions (n, 3) from per-particle E (n,) and V0 (n, 3)."""
        mu, lam = lame_from_E(np.asarray(E, dtype=np.float64), self.nu)
        s = dict(self.scene, V=np.ascontiguousarray(V0, dtype=np.float64), mu=mu, lam=lam)
        eng = self.eng
        eng.upload(s)
        eng.set_colliders(self.colliders)
        eng.set_time(0.0)
        eng.rollout_forward(self.steps, self.dt, self.every, 0)
        return eng.download()["X"]

    def backward(self, gX, gV=None):
        """Gradients (dE (n,), dV0 (n, 3)) of a loss whose gradient in the
        final (X, V) is (gX, gV); runs after forward()."""
        eng = self.eng
        eng.adjoint_set(X=gX, V=gV)
        # per segment from the last: restore its checkpoint, recompute its
        # steps' inputs into the scratch slots, then the adjoint steps in
        # reverse (zr_rollout_backward)
        eng.rollout_backward(self.steps, self.dt, self.every, 0, self.nseg)
        a = eng.adjoint_get()
        nu = self.nu
        gE = a["mu"] / (2.0 * (1.0 + nu)) + a["lam"] * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        return gE, a["V"]


class MpmRolloutFunction(torch.autograd.Function):
    """X_T = rollout(E, V0); float64 tensors on any device (the engine works
    on its own device buffers; tensors cross through host memory)."""

    @staticmethod
    def forward(ctx, E, V0, problem):
        X = problem.forward(E.detach().cpu().numpy(), V0.detach().cpu().numpy())
        ctx.problem = problem
        return torch.from_numpy(X).to(device=E.device, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gX):
        gE, gV = ctx.problem.backward(gX.detach().cpu().numpy())
        dev = gX.device
        return torch.from_numpy(gE).to(dev), torch.from_numpy(gV).to(dev), None


def rollout(E, V0, problem):
    return MpmRolloutFunction.apply(E, V0, problem)
