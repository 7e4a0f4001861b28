"""Coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) on the device,
against the oracle on the same poses: reports, every body's lift and damping
impulses, and the particles, bit for bit, step after step; refusals as the
reference refuses; snapshots with coupled bodies crossing between the two.
"""
import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, BodyPose, FluidSettings, PhysicsError, Shape

pytestmark = pytest.mark.gpu

TICK = 1.0 / 60.0


def pool(w, size=(1.0, 1.0, 0.8), water=(1.0, 0.4, 0.8)):
    w.add_box_container((0, 0, 0), size)
    w.fill_box((0, 0, 0), water)
    return w


def pair(settings=None, **kw):
    s = settings or FluidSettings()
    return pool(FluidWorld(s), **kw), pool(OracleWorld(s, threads=8), **kw)


def same(gpu, cpu, poses, k):
    g = gpu.step_with_bodies(TICK, poses)
    c = cpu.step_with_bodies(TICK, poses)
    assert g[0] == c[0], k
    assert g[1] == c[1], k
    assert np.array_equal(gpu.ids(), cpu.ids()), k
    assert np.array_equal(gpu.positions(), cpu.positions()), k
    assert np.array_equal(gpu.velocities(), cpu.velocities()), k
    return g


def test_a_turning_box_and_a_kinematic_ball_push_the_water_as_the_oracle_does():
    print("case: box + ball, 6 ticks", flush=True)
    gpu, cpu = pair()
    for w in (gpu, cpu):
        w.couple((3, 0), Shape.cuboid((0.12, 0.1, 0.15)))
        w.couple((1, 2), Shape.sphere(0.08))
    lifted = False
    for k in range(6):
        poses = [BodyPose((3, 0), (0.5, 0.5 - 0.04 * k, 0.4), (0.0, 0.0998334, 0.0, 0.9950042),
                          (0.0, -2.4, 0.0), (0.0, 1.0, 0.3)),
                 BodyPose((1, 2), (0.2, 0.35, 0.3), linear_velocity=(0.6, 0.0, 0.0), dynamic=False)]
        _, pushes = same(gpu, cpu, poses, k)
        # handle order: the ball (1, 2), then the box (3, 0)
        lifted = lifted or pushes[1].lifted
    assert lifted and not pushes[0].lifted
    assert gpu.state_hash() == cpu.state_hash()


def test_many_bodies_across_many_chunks():
    print("case: 12 bodies, 64k water", flush=True)
    s = FluidSettings(spacing=0.02)
    gpu, cpu = pair(s, size=(1.6, 1.0, 0.8), water=(1.6, 0.4, 0.8))
    shapes = [Shape.sphere(0.05), Shape.cuboid((0.06, 0.03, 0.05)), Shape.capsule(0.03, 0.05)]
    bodies = [(b, 1) for b in range(12)]
    for w in (gpu, cpu):
        for b in bodies:
            w.couple(b, shapes[b[0] % 3])
    for k in range(3):
        poses = [BodyPose(b, (0.12 + 0.12 * b[0], 0.38 + 0.01 * k, 0.2 + 0.03 * b[0]),
                          (0.0, np.sin(0.1 * b[0]), 0.0, np.cos(0.1 * b[0])),
                          (0.3 - 0.05 * b[0], -0.2, 0.1), (0.0, 0.5, 0.0), dynamic=b[0] % 2 == 0)
                 for b in bodies]
        _, pushes = same(gpu, cpu, poses, k)
    assert sum(p.impulse != (0.0, 0.0, 0.0) for p in pushes) >= 6


def test_a_fast_ball_raises_the_sound_speed_only_near_water():
    print("case: 24 m/s ball", flush=True)
    gpu, cpu = pair(size=(1.0, 1.2, 1.0), water=(1.0, 0.6, 1.0))
    for w in (gpu, cpu):
        w.couple((1, 0), Shape.sphere(0.1))
    speed = 24.0
    far = [BodyPose((1, 0), (0.5, 5.0, 0.5), linear_velocity=(0, -speed, 0))]
    r, _ = same(gpu, cpu, far, "unsorted")
    assert r.sound_speed == np.float32(4 * speed)
    r, _ = same(gpu, cpu, far, "far")
    assert r.sound_speed == np.float32(20.0)
    for k in range(2):
        y = 0.75 - speed * TICK * k
        r, _ = same(gpu, cpu, [BodyPose((1, 0), (0.5, y, 0.5), linear_velocity=(0, -speed, 0))], k)
    assert r.sound_speed == np.float32(4 * speed)


def test_coupled_steps_are_refused_as_the_reference_refuses_them():
    print("case: refusals", flush=True)
    w = pool(FluidWorld(FluidSettings(solver=DFSPH)))
    w.couple((1, 0), Shape.sphere(0.1))
    with pytest.raises(PhysicsError) as e:
        w.step_with_bodies(TICK, [BodyPose((1, 0), (0.5, 0.3, 0.4))])
    assert e.value.name == "InvalidSettings"
    w = pool(FluidWorld(FluidSettings()))
    with pytest.raises(PhysicsError) as e:
        w.couple((1, 0), Shape.sphere(0.001))
    assert e.value.name == "InvalidShape"
    w.couple((1, 0), Shape.sphere(0.1))
    with pytest.raises(PhysicsError) as e:
        w.step_with_bodies(TICK, [BodyPose((1, 0), (0.5, 0.3, 0.4))] * 2)
    assert e.value.name == "InvalidHandle"
    with pytest.raises(PhysicsError) as e:
        w.step_with_bodies(TICK, [BodyPose((1, 0), (0.5, float("inf"), 0.4))])
    assert e.value.name == "NonFinite"
    assert w.coupled_len() == 1
    r, pushes = w.step_with_bodies(TICK, [])
    assert pushes == [] and w.coupled_len() == 0
    assert w.uncouple((1, 0)) is False


def test_coupled_bodies_cross_between_device_and_oracle_snapshots():
    print("case: snapshots with bodies", flush=True)
    gpu, cpu = pair()
    for w in (gpu, cpu):
        w.couple((3, 0), Shape.cuboid((0.15, 0.15, 0.15)))
        w.couple((4, 1), Shape.cuboid((1.0, 0.05, 0.05)))
    poses = lambda k: [BodyPose((3, 0), (0.5, 0.45 - 0.02 * k, 0.4), linear_velocity=(0.0, -1.2, 0.0)),  # noqa
                       BodyPose((4, 1), (0.5, 0.1, 0.6), dynamic=False)]
    same(gpu, cpu, poses(0), 0)
    assert gpu.save_snapshot() == cpu.save_snapshot()
    back = FluidWorld.load_snapshot(cpu.save_snapshot())
    assert back.coupled_len() == 2
    for k in range(1, 3):
        assert back.step_with_bodies(TICK, poses(k)) == cpu.step_with_bodies(TICK, poses(k))
    assert back.state_hash() == cpu.state_hash()


def test_a_step_that_outgrows_the_cell_table_carries_the_bodies_as_the_oracle_does():
    # no container, so the device's cell table is sized by the particle count
    # (max(4096, 2n) cells and a quarter more): water thrown apart at 80 m/s
    # spreads over many times that within the tick, the step's launches
    # abort and run again from the last finished substep, and the body
    # points must stand where the oracle's are after every substep
    print("case: exploding block beside a sphere", flush=True)
    s = FluidSettings()
    rng = np.random.default_rng(7)
    x = np.stack(np.meshgrid(*[np.arange(20) * 0.05 + 0.025] * 3, indexing="ij"), -1).reshape(-1, 3)
    v = rng.uniform(-80.0, 80.0, x.shape)
    gpu, cpu = FluidWorld(s), OracleWorld(s, threads=8)
    for w in (gpu, cpu):
        w.add_particles(x, v)
        w.couple((2, 0), Shape.sphere(0.2))
    poses = [BodyPose((2, 0), (0.5, 0.5, 1.25), linear_velocity=(0.0, 0.0, -3.0))]
    r, _ = same(gpu, cpu, poses, "explode")
    assert r.substeps > 40
    p = gpu.positions()
    cells = np.floor(p / np.float32(s.radius())).astype(np.int64)
    volume = np.prod(cells.max(axis=0) - cells.min(axis=0) + 1)
    assert volume > 2 * (2 * len(x) * 1.25 + 4096)
    same(gpu, cpu, poses, "after")


def test_poses_as_an_array_step_as_poses_as_values():
    print("case: array poses", flush=True)
    from solunar_amd import scenes
    from solunar_amd.types import POSE_DTYPE

    field = scenes.body_field(count=12, spacing=0.05, block=(0.5, 0.4, 0.8), box=(1.0, 1.0, 0.8), radii=(0.1, 0.2))
    worlds = []
    for _ in range(2):
        w = pool(FluidWorld(FluidSettings()))
        for handle, r, *_ in field:
            w.couple(handle, Shape.sphere(r))
        worlds.append(w)
    for k in range(3):
        a = worlds[0].step_with_bodies(TICK, scenes.body_poses(field, k * TICK))
        arr = scenes.body_poses_array(field, k * TICK)
        assert arr.dtype == POSE_DTYPE
        b = worlds[1].step_with_bodies(TICK, arr)
        assert a[0] == b[0]
        assert [p.impulse for p in a[1]] == [tuple(x) for x in b[1]["impulse"].tolist()]
        assert [p.lift_impulse for p in a[1]] == [tuple(x) for x in b[1]["lift_impulse"].tolist()]
    assert worlds[0].state_hash() == worlds[1].state_hash()


def test_a_step_whose_first_substep_outgrows_the_cell_table():
    # a cluster round a sphere and a few particles 6 m out: the device sizes
    # its cell table by the particle count (no container), so the very first
    # substep overflows it and the step runs again from there, carrying the
    # body points again from their placement
    print("case: first substep overflows", flush=True)
    s = FluidSettings()
    cluster = np.stack(np.meshgrid(*[np.arange(10) * 0.05 + 0.275] * 3, indexing="ij"), -1).reshape(-1, 3)
    far = np.array([[x, y, z] for x in (-3.0, 3.0) for y in (-3.0, 3.0) for z in (-3.0, 3.0)])
    x = np.concatenate([cluster, far]).astype(np.float32)
    gpu, cpu = FluidWorld(s), OracleWorld(s, threads=8)
    for w in (gpu, cpu):
        w.add_particles(x, np.zeros_like(x))
        w.couple((2, 0), Shape.sphere(0.15))
    cells = np.floor(x / np.float32(s.radius())).astype(np.int64)
    volume = np.prod(cells.max(axis=0) - cells.min(axis=0) + 1)
    assert volume > 4 * (2 * len(x) * 1.25 + 4096)
    for k in range(2):
        poses = [BodyPose((2, 0), (0.5, 0.5, 0.5 - 0.02 * k), linear_velocity=(0.0, 0.0, -1.5))]
        same(gpu, cpu, poses, k)


def pair_room_retries(gpu, rep):
    """(first, mid): whether this step's pair records outgrew their room in
    its first substep or a later one. An outgrown substep aborts the rest of
    the step's launches and the step runs again from the last finished
    substep: the force pass is enqueued substeps + (substeps - finished)
    times."""
    extra = gpu.profile_read(reset=True)["force"][1] - rep.substeps
    return extra == rep.substeps, 0 < extra < rep.substeps


def test_steps_that_outgrow_the_pair_room_carry_on_as_the_oracle_does(monkeypatch):
    # no slack over the last step's pair count and a first room of 16
    # records per body point: a ball coupled deep under water outgrows the
    # room in its first substep; a ball falling into water outgrows it
    # mid-step as more of it goes under. Reports, pushes and particles must
    # still equal the oracle's, which has no room to outgrow.
    monkeypatch.setenv("SOLUNAR_BODY_PAIR_SLACK", "0")
    print("case: tight pair room", flush=True)
    firsts = mids = 0
    for start, speed, ticks in ((0.25, 0.0, 2), (0.62, -3.0, 8)):
        gpu, cpu = pair()
        for w in (gpu, cpu):
            w.couple((2, 0), Shape.sphere(0.12))
        gpu.profile(True)
        gpu.profile_read(reset=True)
        for k in range(ticks):
            poses = [BodyPose((2, 0), (0.5, start + speed * TICK * k, 0.4), linear_velocity=(0.0, speed, 0.0))]
            rep, _ = same(gpu, cpu, poses, (start, k))
            first, mid = pair_room_retries(gpu, rep)
            print(f"ball from {start}, tick {k}: {rep.substeps} substeps, first {first}, mid {mid}", flush=True)
            firsts += first
            mids += mid
        assert gpu.state_hash() == cpu.state_hash()
    assert firsts >= 1 and mids >= 1


def test_lifts_read_by_sorting_columns_push_as_the_oracle_does(monkeypatch):
    # a body whose column box is too sparse for a column table sorts its
    # water's column keys instead; here every body takes that path
    monkeypatch.setenv("SOLUNAR_LIFT_COLUMN_TABLES", "0")
    print("case: sorted lift columns", flush=True)
    test_a_turning_box_and_a_kinematic_ball_push_the_water_as_the_oracle_does()
    test_many_bodies_across_many_chunks()


def test_bodies_in_a_large_pool_push_as_the_oracle_does():
    # ~330k particles: past 262,144 the force pass with body points takes
    # its two-phase walk for the water pairs
    print("case: 2 bodies, 330k water", flush=True)
    s = FluidSettings(spacing=0.01)
    gpu, cpu = pair(s, size=(1.0, 1.0, 0.8), water=(1.0, 0.42, 0.8))
    assert len(gpu) > 262144
    for w in (gpu, cpu):
        w.couple((1, 0), Shape.sphere(0.08))
        w.couple((2, 0), Shape.cuboid((0.1, 0.06, 0.08)))
    for k in range(2):
        poses = [BodyPose((1, 0), (0.3, 0.42 - 0.02 * k, 0.4), linear_velocity=(0.0, -1.2, 0.0)),
                 BodyPose((2, 0), (0.7, 0.4, 0.4), (0.0, 0.0998334, 0.0, 0.9950042), (0.2, 0.0, 0.0),
                          (0.0, 0.5, 0.0))]
        same(gpu, cpu, poses, k)
    assert gpu.state_hash() == cpu.state_hash()
