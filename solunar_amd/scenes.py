"""Seeded synthetic scenes of the benchmark and test workloads.

Every builder works on any object with the FluidWorld methods (the device
world or the CPU oracle), so both sides get identical inputs.

Workloads (dam break in a 1.6 x 1.6 x 0.8 m container, water block at the
low corner, lattice sites nudged by up to +-0.05 spacing per axis):
  dam_small   WCSPH, spacing 0.02, 0.4 x 0.8 x 0.4 block: 16,000 particles
  dam_1m      DFSPH, spacing 0.005, same block: 1,024,000 particles
  dam_16m     DFSPH, spacing 0.002: 16,000,000 particles
  adjoint_64k WCSPH, spacing 0.01, 0.4 x 0.4 x 0.4 block: 64,000 particles with
              seeded N(0, 0.1 m/s) initial velocities, 200 steps of 1e-4 s,
              loss = mean |x_T - x_T*|^2 against the viscosity-0.08 rollout
  trench_720k WCSPH at spacing 0.005: a dam break of 720,000 particles (1 x
              0.3 x 0.3 m) down a 2 m trench 0.3 m deep and wide in flat
              ground, one terrain region of 15.6M nodes
  bodies_1m   dam_1m under WCSPH with 200 coupled spheres (radii uniform in
              0.02 to 0.08 m): 100 resting on the water block's top, half a
              spacing clear of it, settling in at 0.1 m/s, and 100 on the dry
              floor downstream; drifting up to 0.1 m/s and turning about y at
              up to 1 rad/s; every other one dynamic
  batch_256   256 WCSPH dam breaks at spacing 0.02, block sides uniform in
              [0.2, 0.8] m per axis, redrawn until the block holds 2,000 to
              64,000 particles
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


def sphere_points(center, radius: float, spacing: float, reach: float) -> np.ndarray:
    """Boundary points of a sphere the way the reference samples a body
    (shape_points, src/fluid/couple.rs:452-507): lattice points `spacing`
    apart, centred on the body, kept when 0 <= radius - |p| <= reach; then
    moved to `center`. f32 throughout."""
    f = np.float32
    sp, r, reach = f(spacing), f(radius), f(reach)
    cells = int(min(max(int(np.floor(f(2.0) * r / sp)), 0), 1 << 12))
    start = -(f(cells) - f(1.0)) * sp / f(2.0)
    idx = np.arange(cells, dtype=np.float32)
    lat = start + idx * sp
    # loop order y, z, x as in the reference
    y, z, x = np.meshgrid(lat, lat, lat, indexing="ij")
    at = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float32)
    length = np.sqrt((at[:, 0] * at[:, 0] + at[:, 1] * at[:, 1]) + at[:, 2] * at[:, 2]).astype(np.float32)
    depth = (r - length).astype(np.float32)
    keep = (depth >= 0) & (depth <= reach)
    return (at[keep] + np.asarray(center, dtype=np.float32)).astype(np.float32)


def obstacle_field(count: int, spacing: float, seed: int = 7, region=((0.5, 0.05, 0.05), (1.5, 0.8, 0.75)),
                   radii=(0.02, 0.08)) -> np.ndarray:
    """`count` spheres with radii uniform in `radii`, centres uniform in
    `region` (downstream of the water block), sampled with shell depth of one
    kernel radius (2 spacings); all their points in one array."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(region[0]), np.asarray(region[1])
    parts = []
    for _ in range(count):
        r = rng.uniform(*radii)
        c = rng.uniform(lo + r, hi - r)
        parts.append(sphere_points(c, r, spacing, 2.0 * spacing))
    return np.concatenate(parts) if parts else np.zeros((0, 3), np.float32)


def settings_bodies_1m() -> FluidSettings:
    """dam_1m's water under WCSPH: the reference couples bodies with the
    weakly compressible solver only (src/fluid/couple.rs:223-225)."""
    return replace(settings_dam_1m(), solver=WCSPH)


def body_field(count: int = 200, seed: int = 11, spacing: float = 0.005, block=BLOCK, box=BOX,
               radii=(0.02, 0.08)):
    """[(handle, radius, centre, velocity, spin)] of `count` spheres: the
    first half on the block's top surface (its particles' centres reach
    block_y - spacing / 2), the rest on the floor past the block."""
    rng = np.random.default_rng(seed)
    out = []
    top = block[1] - 0.5 * spacing
    for k in range(count):
        r = float(rng.uniform(*radii))
        if k < count // 2:
            c = np.array([rng.uniform(0.0, block[0]), top + 0.5 * spacing + r + 0.5 * spacing,
                          rng.uniform(0.0, block[2])])
            v = np.array([rng.uniform(-0.1, 0.1), -0.1, rng.uniform(-0.1, 0.1)])
        else:
            c = np.array([rng.uniform(block[0] + 0.1 + r, box[0] - r), r + spacing, rng.uniform(r, box[2] - r)])
            v = np.array([rng.uniform(-0.1, 0.1), 0.0, rng.uniform(-0.1, 0.1)])
        spin = float(rng.uniform(-1.0, 1.0))
        out.append(((k, 0), r, c, v, spin))
    return out


def body_poses(field, t: float):
    """The field's poses at time t: centres moved on, turned about y."""
    from .types import BodyPose

    poses = []
    for handle, _, c, v, spin in field:
        half = 0.5 * spin * t
        poses.append(BodyPose(handle, tuple(float(x) for x in c + v * t), (0.0, float(np.sin(half)), 0.0,
                                                                              float(np.cos(half))),
                              tuple(float(x) for x in v), (0.0, spin, 0.0), dynamic=handle[0] % 2 == 0))
    return poses


def body_poses_array(field, t: float):
    """body_poses as one numpy array of POSE_DTYPE (the C layout)."""
    from .types import POSE_DTYPE

    out = np.zeros(len(field), POSE_DTYPE)
    idx = np.array([h for h, *_ in field], np.uint32)
    c = np.array([f[2] for f in field], np.float64)
    v = np.array([f[3] for f in field], np.float64)
    spin = np.array([f[4] for f in field], np.float64)
    half = 0.5 * spin * t
    out["index"], out["generation"] = idx[:, 0], idx[:, 1]
    out["position"] = c + v * t
    out["rotation"][:, 1] = np.sin(half)
    out["rotation"][:, 3] = np.cos(half)
    out["linear_velocity"] = v
    out["angular_velocity"][:, 1] = spin
    out["dynamic"] = (idx[:, 0] % 2 == 0).astype(np.int32)
    return out


def dam_with_obstacles(world, obstacles: int = 200, seed: int = 7) -> int:
    """The dam break plus `obstacles` spheres added as one boundary."""
    n = dam_break(world)
    pts = obstacle_field(obstacles, world.settings.spacing, seed)
    if len(pts):
        world.add_boundary_points(pts)
    return n


def batch_blocks(count: int = 256, seed: int = SEED, spacing: float = 0.02, low: int = 2000,
                 high: int = 64000):
    """Block sides of `count` dam-break scenes: uniform in [0.2, 0.8] m per
    axis, drawn again until fill_box's count (floor(side / spacing) per axis,
    in f32) lies in [low, high]. Returns [(sides, particles)]."""
    rng = np.random.default_rng(seed)
    sp = np.float32(spacing)
    out = []
    while len(out) < count:
        sides = rng.uniform(0.2, 0.8, size=3).astype(np.float32)
        n = int(np.prod(np.floor(sides / sp).astype(np.int64)))
        if low <= n <= high:
            out.append((tuple(float(x) for x in sides), n))
    return out


def batch_scene(world, sides, index: int, seed: int = SEED) -> int:
    """Scene `index` of a batch: the container, then its jittered block."""
    world.add_box_container((0.0, 0.0, 0.0), BOX)
    return world.fill_box_jittered((0.0, 0.0, 0.0), sides, JITTER, (seed + 7919 * (index + 1)) & 0xFFFFFFFF)


def batch_worlds(make_world, count: int = 256, seed: int = SEED, settings: FluidSettings = None):
    """`count` worlds made by `make_world(settings)`, each one batch scene."""
    settings = settings or settings_dam_small()
    worlds = []
    for k, (sides, n) in enumerate(batch_blocks(count, seed, settings.spacing)):
        w = make_world(settings)
        got = batch_scene(w, sides, k, seed)
        assert got == n, (got, n)
        worlds.append(w)
    return worlds


ADJOINT_BLOCK = (0.4, 0.4, 0.4)
ADJOINT_SEED = 2024


def settings_adjoint_64k(viscosity: float = 0.05) -> FluidSettings:
    """WCSPH at spacing 0.01 for the differentiable rollout."""
    return replace(settings_dam_small(), spacing=0.01, viscosity=viscosity)


def adjoint_positions(spacing: float = 0.01) -> np.ndarray:
    """The jittered 0.4 m cube of the differentiable rollout, in fill order
    (ids 0..63,999 at spacing 0.01), as fill_box_jittered lays it out."""
    f = np.float32
    sp = f(spacing)
    n = [int(np.floor(f(side) / sp)) for side in ADJOINT_BLOCK]
    half = sp * f(0.5)
    state = np.uint32(SEED)
    amp = f(2.0) * f(JITTER) * sp
    umax = f(4294967295.0)
    out = np.zeros((n[0] * n[1] * n[2], 3), np.float32)
    k = 0

    def draw():
        nonlocal state
        s = int(state)
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        state = np.uint32(s)
        return (f(s) / umax - f(0.5)) * amp

    for y in range(n[1]):
        for z in range(n[2]):
            for x in range(n[0]):
                at = [half + f(x) * sp, half + f(y) * sp, half + f(z) * sp]
                j = [draw(), draw(), draw()]
                out[k] = [at[0] + j[0], at[1] + j[1], at[2] + j[2]]
                k += 1
    return out


def adjoint_velocities(n: int, seed: int = ADJOINT_SEED) -> np.ndarray:
    """Initial velocities N(0, 0.1 m/s) per component, f32."""
    return np.random.default_rng(seed).normal(0.0, 0.1, size=(n, 3)).astype(np.float32)


def container(world) -> None:
    world.add_box_container((0.0, 0.0, 0.0), BOX)


# ---- terrain ground (src/fluid/terrain.rs tests) ----

# The region the reference's trench tests sample (terrain.rs:356-357).
TRENCH_LOW = (-0.4, -0.7, -0.5)
TRENCH_HIGH = (2.2, 0.3, 0.5)


def trench_distance(points) -> np.ndarray:
    """Signed distance of flat ground at y = 0 with a trench along x from
    -0.1 to 1.9, its bed at y = -0.4, 0.6 m wide (the reference's test
    terrain, terrain.rs:360-381, as an exact field rather than its 10 cm
    voxels): max(plane, -box), positive above the ground."""
    p = np.asarray(points, np.float64)
    plane = p[:, 1]
    c, h = np.array([0.9, -0.2, 0.0]), np.array([1.0, 0.2, 0.3])
    q = np.abs(p - c) - h
    box = np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
    return np.maximum(plane, -box).astype(np.float32)


def in_trench(world) -> int:
    """add_terrain over the trench region; returns the terrain's boundary id."""
    return world.add_terrain(TRENCH_LOW, TRENCH_HIGH, trench_distance)


# The trench_720k terrain: a 2 m trench, bed at y = -0.3, 0.3 m wide, in
# flat ground at y = 0; its region a little round it, under 2^24 nodes at a
# spacing of 0.005.
LONG_TRENCH_LOW = (-0.04, -0.32, -0.17)
LONG_TRENCH_HIGH = (2.04, 0.02, 0.17)


def long_trench_distance(points) -> np.ndarray:
    p = np.asarray(points, np.float64)
    c, h = np.array([1.0, -0.15, 0.0]), np.array([1.0, 0.15, 0.15])
    q = np.abs(p - c) - h
    box = np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
    return np.maximum(p[:, 1], -box).astype(np.float32)


def settings_trench_720k() -> FluidSettings:
    return FluidSettings(solver=WCSPH, spacing=0.005, rest_density=1000.0, support=2.0, sound_speed=20.0,
                         viscosity=0.05, gravity=(0.0, -9.81, 0.0), max_particles=1 << 22)


def trench_dam(world) -> int:
    """The trench_720k scene: the terrain, then the water block at the
    trench's closed end; returns the particle count."""
    world.add_terrain(LONG_TRENCH_LOW, LONG_TRENCH_HIGH, long_trench_distance)
    return world.fill_box((0.0, -0.3, -0.15), (1.0, 0.0, 0.15))
