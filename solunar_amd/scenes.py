"""Seeded synthetic scenes of the benchmark and test workloads.

Every builder works on any object with the FluidWorld methods (the device
world or the CPU oracle), so both sides get identical inputs.

Workloads (dam break in a 1.6 x 1.6 x 0.8 m container, water block at the
low corner, lattice sites nudged by up to +-0.05 spacing per axis):
  dam_small   WCSPH, spacing 0.02, 0.4 x 0.8 x 0.4 block: 16,000 particles
  dam_1m      DFSPH, spacing 0.005, same block: 1,024,000 particles
  dam_16m     DFSPH, spacing 0.002: 16,000,000 particles
The Wendland C2 kernel with support 2 stands for "h = 2 dx"; containers have
ceil(support) = 2 wall layers (src/fluid/boundary.rs:41-68).
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np

from .types import DFSPH, WCSPH, FluidSettings

BOX = (1.6, 1.6, 0.8)
BLOCK = (0.4, 0.8, 0.4)
JITTER = 0.05
SEED = 0x9E3779B9


def settings_dam_small() -> FluidSettings:
    """WCSPH dam break, spacing 0.02, alpha 0.05, c0 20 m/s."""
    return FluidSettings(solver=WCSPH, spacing=0.02, rest_density=1000.0, support=2.0, sound_speed=20.0,
                         viscosity=0.05, gravity=(0.0, -9.81, 0.0), max_particles=1 << 22)


def settings_dam_1m() -> FluidSettings:
    """DFSPH dam break, spacing 0.005; divergence to 0.1 %, density to 0.01 %."""
    return FluidSettings(solver=DFSPH, spacing=0.005, rest_density=1000.0, support=2.0, sound_speed=20.0,
                         viscosity=0.05, gravity=(0.0, -9.81, 0.0), divergence_tolerance=1e-3,
                         density_tolerance=1e-4, max_iterations=100, max_particles=1 << 22)


def settings_dam_16m() -> FluidSettings:
    return replace(settings_dam_1m(), spacing=0.002, max_particles=1 << 25)


def dam_break(world, block=BLOCK, box=BOX, jitter: float = JITTER, seed: int = SEED) -> int:
    """Container, then the jittered water block; returns the particle count."""
    world.add_box_container((0.0, 0.0, 0.0), box)
    return world.fill_box_jittered((0.0, 0.0, 0.0), block, jitter, seed)


def shaken_positions(count: int) -> np.ndarray:
    """The reference tests' `shaken` lattice (src/fluid/mod.rs:752-774) in f32."""
    seed = np.uint32(0x9E3779B9)
    side = int(np.ceil(np.cbrt(np.float32(count))))
    out = np.zeros((count, 3), dtype=np.float32)
    umax = np.float32(4294967295.0)
    half = np.float32(0.5)
    amp = np.float32(0.02)
    step = np.float32(0.05)

    def nxt():
        nonlocal seed
        s = int(seed)
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        seed = np.uint32(s)
        return (np.float32(s) / umax - half) * amp

    for i in range(count):
        lattice = np.array([i % side, i // side % side, i // side // side], dtype=np.float32) * step
        j = np.array([nxt(), nxt(), nxt()], dtype=np.float32)
        out[i] = lattice + j
    return out


def shaken(world, count: int):
    world.add_particles(shaken_positions(count))
    return world
