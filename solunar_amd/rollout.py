"""A differentiable WCSPH rollout for torch: a thin harness over the C ABI's
tape (sfl_wcsph_record / sfl_wcsph_adjoint).

`rollout(v0, alpha, positions, settings, dt, steps, every, build)` makes a
world with `build(world)` (containers, walls), adds the particles at
`positions` with velocities `v0`, records `steps` steps of `dt` with a
checkpoint every `every` steps, and returns the final positions per particle
id. Its backward runs the hand-written adjoint and returns dL/dv0 and
dL/dalpha (alpha is the artificial-viscosity setting). The world and the
tape live as long as the autograd graph holds the node.
"""
from __future__ import annotations

from dataclasses import replace
from typing import Callable, Optional

import numpy as np
import torch

from .fluid import FluidWorld, Tape
from .types import FluidSettings


class WcsphRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0: torch.Tensor, alpha: torch.Tensor, positions: np.ndarray, settings: FluidSettings,
                dt: float, steps: int, every: int, build: Optional[Callable]):
        s = replace(settings, viscosity=float(alpha.detach().cpu()))
        world = FluidWorld(s)
        if build is not None:
            build(world)
        world.add_particles(np.asarray(positions, np.float32), v0.detach().cpu().numpy().astype(np.float32))
        tape = Tape(world, dt, steps, every)
        ids = world.ids()
        x = np.empty((len(ids), 3), np.float32)
        x[ids] = world.positions()
        ctx.tape = tape
        ctx.alpha_dtype = alpha.dtype
        return torch.from_numpy(x).to(v0.device)

    @staticmethod
    def backward(ctx, grad_x: torch.Tensor):
        xbar = grad_x.detach().cpu().numpy().astype(np.float32)
        _, v0_bar, alpha_bar = ctx.tape.adjoint(xbar)
        dv0 = torch.from_numpy(v0_bar).to(grad_x.device)
        dalpha = torch.tensor(alpha_bar, dtype=ctx.alpha_dtype, device=grad_x.device)
        return dv0, dalpha, None, None, None, None, None, None


def rollout(v0: torch.Tensor, alpha: torch.Tensor, positions: np.ndarray, settings: FluidSettings, dt: float,
            steps: int, every: int = 20, build: Optional[Callable] = None) -> torch.Tensor:
    """Final positions (N, 3) per id after `steps` WCSPH steps; differentiable
    in `v0` (N, 3) and `alpha` (a 0-d tensor)."""
    return WcsphRollout.apply(v0, alpha, positions, settings, dt, steps, every, build)
