"""Differentiable rollouts as a torch.autograd.Function.

A thin shim: the rollout and its adjoint run in the HIP library
(ne_gpu_rollout / ne_gpu_rollout_backward in include/ne_gpu.h); PyTorch only
holds the device buffers.  The gradient is that of the step with each step's
contact set (pairs, normals, points, depths) held fixed -- integrate, the
constraint build, the warm start and the fixed-iteration impulse solve are
differentiated exactly -- with respect to the initial velocities and each
world's default friction and restitution.
"""
from __future__ import annotations

import torch

from .gpu import GpuWorld
from .scene import Batch, Scene


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, mu, e, world, nsteps, every):
        inputs = tuple(t.detach().double().cpu().numpy().copy() for t in (v0, mu, e))
        ctx.world = world
        ctx.gen = world._run(inputs, nsteps, every)[1]
        ctx.inputs, ctx.nsteps, ctx.every = inputs, nsteps, every
        ctx.nw = mu.numel()
        ctx.devs = (v0.device, mu.device, e.device)
        return torch.from_numpy(world._last_x).to(v0.device)

    @staticmethod
    def backward(ctx, gx):
        # The library holds the checkpoints of the last rollout only: when
        # another rollout ran on these worlds since this node's forward, this
        # node's rollout is recomputed (bit-identically) before its adjoint.
        if ctx.world._generation != ctx.gen:
            ctx.gen = ctx.world._run(ctx.inputs, ctx.nsteps, ctx.every)[1]
        gv, gmu, ge = ctx.world.gpu.rollout_backward(gx.detach().double().cpu().numpy(), ctx.nw)
        return (torch.from_numpy(gv).to(ctx.devs[0]), torch.from_numpy(gmu).to(ctx.devs[1]),
                torch.from_numpy(ge).to(ctx.devs[2]), None, None, None)


class DifferentiableWorlds:
    """Worlds resident on the GPU whose rollouts are differentiable.

    `rollout(v0, mu, e)` runs `nsteps` steps from the scene's initial poses
    with initial velocities v0 (n x 3 float64 tensor) and per-world default
    friction / restitution mu, e (one entry per world) and returns the final
    positions (n x 3) as part of the autograd graph.  The tensors live on the
    host: the rollout's state stays resident in the HIP library, which copies
    only v0 in and x out (this also keeps PyTorch's own HIP runtime out of
    the process, which would otherwise be loaded beside the library's)."""

    def __init__(self, scene: Scene | Batch, nsteps: int = 120, every: int = 16):
        self.gpu = GpuWorld()
        if isinstance(scene, Batch):
            self.scene = scene.scene
            self.gpu.set_scene(scene.scene)
            self.gpu.set_worlds(scene.offsets, scene.half, scene.restitution, scene.friction)
            self.nworlds = scene.nworlds
        else:
            self.scene = scene
            self.gpu.set_scene(scene)
            self.nworlds = 1
        self.n = self.scene.n
        self.nsteps = nsteps
        self.every = every
        self._generation = 0   # which rollout the library's checkpoints belong to
        self._last_x = None

    def _reset(self):
        s = self.scene
        self.gpu.set_state(s.x, s.v, s.q, s.w)

    def _run(self, inputs, nsteps, every):
        """Rollout from the scene's initial poses; returns (final x, generation)."""
        v0, mu, e = inputs
        self._reset()
        self._last_x = self.gpu.rollout(v0, mu, e, nsteps, every)
        self._generation += 1
        return self._last_x, self._generation

    def rollout(self, v0: torch.Tensor, mu: torch.Tensor, e: torch.Tensor) -> torch.Tensor:
        return _Rollout.apply(v0, mu, e, self, self.nsteps, self.every)

    def close(self):
        self.gpu.close()
