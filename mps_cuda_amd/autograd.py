"""torch.autograd.Function over the hand-written adjoint (mps_adjoint_*).

    pos = rollout(nu, c, vel0, scene, params, nsteps, checkpoint_every=25)

runs `nsteps` explicit MPS steps of `scene` on the GPU with kinematic
viscosity `nu`, sound speed `c` and initial fluid velocities `vel0`
([n_fluid, dim], fluid particles in creation order) and returns the final
fluid positions [n_fluid, dim] (CPU tensor, the dtype of vel0).  Backward
runs the C-ABI reverse sweep (checkpoint every `checkpoint_every` steps,
recompute between checkpoints) and returns dL/dnu, dL/dc and dL/dvel0.  The
derivative is that of the step with its discrete choices (neighbour, mirror
and contact sets, clip masks, speed clamp) frozen.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import ffi


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nu, c, vel0, scene, params, nsteps, checkpoint_every):
        p = ffi.Params.from_buffer_copy(params)
        p.nu = float(nu)
        p.c = float(c)
        fluid = np.where(scene.typ == ffi.FLD)[0]
        dim = p.dim
        if tuple(vel0.shape) != (len(fluid), dim):
            raise ValueError(f"vel0 must be [{len(fluid)}, {dim}]")
        vel = np.zeros((3, scene.n)) if scene.vel is None else np.array(scene.vel, np.float64)
        vel[:dim, fluid] = vel0.detach().cpu().double().numpy().T
        sc = copy.copy(scene)
        sc.vel = vel
        h = ffi.Handle(p, sc)
        h.adjoint_forward(int(nsteps), int(checkpoint_every))
        st = h.state()
        ctx.h, ctx.fluid, ctx.dim, ctx.dtype = h, fluid, dim, vel0.dtype
        return torch.from_numpy(st["pos"][:dim, fluid].T.copy()).to(vel0.dtype)

    @staticmethod
    def backward(ctx, gpos):
        h = ctx.h
        xbar = np.zeros((3, h.n))
        xbar[:ctx.dim, ctx.fluid] = gpos.detach().cpu().double().numpy().T
        g = h.adjoint_backward(pos_bar=xbar)
        h.close()
        ctx.h = None
        gv = torch.from_numpy(g["vel0"][:ctx.dim, ctx.fluid].T.copy()).to(ctx.dtype)
        return (torch.tensor(g["nu"], dtype=ctx.dtype), torch.tensor(g["c"], dtype=ctx.dtype), gv,
                None, None, None, None)


def rollout(nu, c, vel0, scene, params, nsteps, checkpoint_every=25):
    return Rollout.apply(nu, c, vel0, scene, params, nsteps, checkpoint_every)
