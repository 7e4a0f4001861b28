"""Coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) in the CPU oracle.

The reference's own coupling tests need its rigid world, which is not part
of this path: here the bodies stand where a pose puts them. What pins the
restatement without the reference: lattice counts of the sampled shapes
worked out in integers, rotations against scipy, the body points' density
and neighbour lists against brute force, Archimedes' buoyancy and the drag
law in closed form, the momentum a dynamic body's lift takes out of the
water, thread-count invariance (couple.rs:250-264), the fast-body sound
speed rule (couple.rs:24-56), and the refusals of couple.rs:268-291.
"""
import itertools

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle.oracle import OracleWorld
from solunar_amd.types import DFSPH, BodyPose, FluidSettings, PhysicsError, Shape, SnapshotError

TICK = 1.0 / 60.0


def still_pool(threads=8, settings=None):
    # water at rest on its lattice, 0.4 m deep (surface particles at 0.375)
    w = OracleWorld(settings or FluidSettings(), threads=threads)
    w.add_box_container((0, 0, 0), (1.0, 1.0, 0.8))
    w.fill_box((0, 0, 0), (1.0, 0.4, 0.8))
    return w


def lattice_count(c, keep):
    """Points of a c^3 lattice at half-integer offsets (i - (c-1)/2) whose
    integer 4*|p/s|^2 satisfies keep."""
    half = [(2 * i - (c - 1)) for i in range(c)]  # twice the offset in spacings
    return sum(1 for a, b, d in itertools.product(half, half, half) if keep(a * a + b * b + d * d))


def test_shapes_sample_and_bad_ones_are_refused():
    w = still_pool(threads=1)
    for shape in (Shape.sphere(0.2), Shape.capsule(0.1, 0.2), Shape.cuboid((0.3, 0.1, 0.1))):
        w.couple((7, 1), shape)
        assert len(w.coupled()) == 1
    with pytest.raises(PhysicsError) as e:
        w.couple((7, 1), Shape.sphere(0.001))
    assert e.value.name == "InvalidShape"
    with pytest.raises(PhysicsError) as e:
        w.couple((7, 1), Shape.sphere(-1.0))
    assert e.value.name == "InvalidShape"
    with pytest.raises(PhysicsError) as e:
        w.couple((7, 1), Shape(9, 0.2))
    assert e.value.name == "InvalidShape"
    assert w.uncouple((7, 1))
    assert not w.uncouple((7, 1))
    # a body no pose names is dropped by the next coupled step
    w.couple((7, 1), Shape.sphere(0.2))
    w.step_with_bodies(TICK, [])
    assert w.coupled() == []


def test_sampled_shapes_hold_the_lattice_counts():
    w = still_pool(threads=1)
    # cuboid, half 0.2, spacing 0.05: 8 per axis at +-0.025 .. +-0.175; the
    # wall points are those within r = 0.1 of a face, i.e. some |x| >= 0.125
    w.couple((1, 0), Shape.cuboid((0.2, 0.2, 0.2)))
    # sphere of 0.3: 12 per axis; |p| <= 0.3 is sum of (2i - 11)^2 <= 144,
    # wall points |p| >= 0.2 are sums >= 64 (the sums are 3 mod 8, no ties)
    w.couple((2, 0), Shape.sphere(0.3))
    # a 2 m cube: probes coarsen by 1.25 until floor(2/s)^3 <= 4096: 16 per axis
    w.couple((3, 0), Shape.cuboid((1.0, 1.0, 1.0)))
    (b1, p1, q1, s1), (b2, p2, q2, s2), (b3, p3, q3, s3) = w.coupled()
    assert (b1, b2, b3) == ((1, 0), (2, 0), (3, 0))
    assert len(p1) == 8 ** 3 - 4 ** 3 and len(q1) == 8 ** 3 and s1 == np.float32(0.05)
    inside = lattice_count(12, lambda n: n <= 144)
    assert len(q2) == inside and len(p2) == inside - lattice_count(12, lambda n: n < 64)
    assert len(q3) == 16 ** 3 and abs(s3 - 0.05 * 1.25 ** 4) < 1e-6
    assert np.isclose(q3.min(axis=0), -7.5 * s3, atol=1e-5).all()
    # every point lies in the shape
    assert (np.abs(p1) <= 0.2).all() and (np.linalg.norm(q2.astype(np.float64), axis=1) <= 0.3 + 1e-6).all()


def test_body_points_stand_and_move_where_the_pose_says():
    w = still_pool(threads=1)
    w.couple((1, 0), Shape.cuboid((0.1, 0.05, 0.15)))
    w.couple((4, 2), Shape.capsule(0.06, 0.08))
    (_, local1, _, _), (_, local2, _, _) = w.coupled()
    # identity rotation: exactly the centre plus the point
    still = [BodyPose((1, 0), (0.5, 0.5, 0.4), linear_velocity=(0.5, -1.0, 0.0)),
             BodyPose((4, 2), (0.2, 0.7, 0.3))]
    w.place_movers(still)
    x, v, owner = w.movers()
    expect = np.concatenate([np.float32([0.5, 0.5, 0.4]) + local1, np.float32([0.2, 0.7, 0.3]) + local2])
    got = {tuple(r) for r in x.tolist()}
    assert got == {tuple(r) for r in expect.tolist()} and len(x) == len(expect)
    assert (v[owner == 0] == np.float32([0.5, -1.0, 0.0])).all() and (v[owner == 1] == 0).all()
    # sorted by cell (x, y, z), then owner
    cells = np.floor(x / np.float32(0.1)).astype(np.int64)
    keys = list(zip(cells[:, 0], cells[:, 1], cells[:, 2], owner))
    assert keys == sorted(keys)
    # a turned, spinning body, against scipy's rotation and cross product
    q = Rotation.from_euler("xyz", [0.3, -1.1, 0.7]).as_quat()  # x, y, z, w
    omega = np.array([0.4, -2.0, 1.5])
    turned = [BodyPose((1, 0), (0.5, 0.5, 0.4), tuple(q), (0.1, 0.2, -0.3), tuple(omega)),
              BodyPose((4, 2), (0.2, 0.7, 0.3))]
    w.place_movers(turned)
    x, v, owner = w.movers()
    arm = Rotation.from_quat(q).apply(local1.astype(np.float64))
    want_x = np.array([0.5, 0.5, 0.4]) + arm
    want_v = np.array([0.1, 0.2, -0.3]) + np.cross(omega, arm)
    mine = owner == 0
    order = np.lexsort(x[mine].T[::-1])
    worder = np.lexsort(want_x.T[::-1])
    assert np.allclose(x[mine][order], want_x[worder], atol=2e-6)
    assert np.allclose(v[mine][order], want_v[worder], atol=2e-6)


def test_body_points_count_in_density_and_lists_as_brute_force_says():
    w = still_pool(threads=4)
    w.couple((1, 0), Shape.sphere(0.12))
    w.update_neighbors()
    w.compute_densities()
    before = w.densities().astype(np.float64)
    w.place_movers([BodyPose((1, 0), (0.45, 0.3, 0.37), linear_velocity=(0.2, 0, 0))])
    w.update_neighbors()
    w.compute_densities()
    after = w.densities().astype(np.float64)
    mx, _, _ = w.movers()
    x = w.positions().astype(np.float64)
    r = float(w.settings.radius())
    start, items = w.mover_neighbors()
    d2 = ((x[:, None, :] - mx[None, :, :].astype(np.float64)) ** 2).sum(-1)
    for i in range(len(x)):
        brute = set(np.nonzero(d2[i] < r * r)[0].tolist())
        assert set(items[start[i]:start[i + 1]].tolist()) == brute, i
    # the density gains mass * sum W over the body points (Wendland C2)
    q = np.sqrt(d2) / r
    wv = 21.0 / (2 * np.pi * r ** 3) * np.where(q < 1, (1 - q) ** 4 * (1 + 4 * q), 0.0)
    gain = w.particle_mass() * wv.sum(axis=1)
    assert (gain > 0).sum() > 100
    assert np.allclose(after - before, gain, rtol=1e-4, atol=1e-3)


def pose_at(y, velocity=(0.0, 0.0, 0.0), dynamic=True):
    return BodyPose((1, 0), (0.5, y, 0.4), linear_velocity=velocity, dynamic=dynamic)


def test_a_body_under_still_water_is_lifted_by_the_water_it_displaces():
    w = still_pool()
    w.couple((1, 0), Shape.cuboid((0.1, 0.1, 0.1)))
    (_, _, probes, s), = w.coupled()
    s = float(s)
    rho, g = 1000.0, 9.81
    # deep under (level 0.4): every probe whole, no flow, no waterline
    lift, = w.lifts(TICK, [pose_at(0.2)])
    assert lift.lifted
    want = rho * g * len(probes) * s ** 3 * TICK
    assert abs(lift.lift_impulse[1] - want) <= 2e-6 * want
    assert abs(lift.lift_impulse[0]) < 1e-9 and abs(lift.lift_impulse[2]) < 1e-9
    assert np.abs(lift.lift_angular).max() < 1e-8
    # centred on the surface: the lower two of its four layers are under
    lift, = w.lifts(TICK, [pose_at(0.4)])
    assert abs(lift.lift_impulse[1] - want / 2) <= 2e-6 * want
    # moving at 0.8 m/s through still water, under: drag 1/2 rho v^2 V^(2/3)
    v = 0.8
    lift, = w.lifts(TICK, [pose_at(0.2, (v, 0.0, 0.0))])
    volume = len(probes) * s ** 3
    drag = 0.5 * rho * v * v * volume ** (2.0 / 3.0) * TICK
    assert abs(lift.lift_impulse[0] + drag) <= 1e-5 * drag
    assert abs(lift.lift_impulse[1] - want) <= 2e-6 * want
    # clear of the water: nothing
    lift, = w.lifts(TICK, [pose_at(2.0)])
    assert not lift.lifted


def test_the_lift_a_dynamic_body_takes_leaves_the_water():
    def run(dynamic):
        w = still_pool()
        w.couple((1, 0), Shape.cuboid((0.1, 0.1, 0.1)))
        report, pushes = w.step_with_bodies(TICK, [pose_at(0.3, (0.0, -0.5, 0.0), dynamic)])
        return w, pushes[0]

    a, pa = run(True)
    b, pb = run(False)
    assert pa.lifted and not pb.lifted and pa.impulse == pb.impulse
    assert np.array_equal(a.ids(), b.ids())
    dv = a.velocities().astype(np.float64) - b.velocities().astype(np.float64)
    taken = a.particle_mass() * dv.sum(axis=0)
    assert np.allclose(taken, -np.array(pa.lift_impulse), rtol=1e-4, atol=1e-7)
    assert np.count_nonzero(np.abs(dv).sum(axis=1)) > 100


def test_coupling_is_the_same_on_any_thread_count():
    def go(threads):
        w = still_pool(threads=threads)
        w.couple((1, 0), Shape.cuboid((0.15, 0.15, 0.15)))
        w.couple((2, 5), Shape.sphere(0.1))
        out = []
        for k in range(6):
            poses = [BodyPose((1, 0), (0.5, 0.55 - 0.05 * k, 0.4), (0.0, 0.0998334, 0.0, 0.9950042),
                              (0.0, -3.0, 0.0), (0.0, 1.0, 0.0)),
                     BodyPose((2, 5), (0.2, 0.35, 0.2), linear_velocity=(1.0, 0.0, 0.0), dynamic=False)]
            out.append(w.step_with_bodies(TICK, poses))
        return out, w.state_hash()

    one, eight = go(1), go(8)
    assert one == eight
    assert any(p.impulse != (0.0, 0.0, 0.0) for _, pushes in one[0] for p in pushes)


def test_a_fast_body_near_water_raises_the_sound_speed_and_one_in_the_air_does_not():
    # couple.rs:24-56: a 24 m/s ball into the pool
    w = OracleWorld(FluidSettings(), threads=8)
    w.add_box_container((0, 0, 0), (1.0, 1.2, 1.0))
    w.fill_box((0, 0, 0), (1.0, 0.6, 1.0))
    w.couple((1, 0), Shape.sphere(0.1))
    speed = 24.0
    ball = [BodyPose((1, 0), (0.5, 5.0, 0.5), linear_velocity=(0, -speed, 0))]
    # before the water's first sort there is no grid to ask: any water counts
    first, _ = w.step_with_bodies(TICK, ball)
    assert first.sound_speed == np.float32(4 * speed)
    far, _ = w.step_with_bodies(TICK, ball)
    assert far.sound_speed == np.float32(20.0)
    louder, squeeze = 0.0, 0.0
    for k in range(3):
        y = 0.75 - speed * TICK * k
        r, _ = w.step_with_bodies(TICK, [BodyPose((1, 0), (0.5, y, 0.5), linear_velocity=(0, -speed, 0))])
        louder, squeeze = max(louder, r.sound_speed), max(squeeze, r.max_compression)
    assert louder >= 4.0 * speed
    assert squeeze < 0.2


def test_coupled_steps_are_refused_as_the_reference_refuses_them():
    w = still_pool(threads=1, settings=FluidSettings(solver=DFSPH))
    w.couple((1, 0), Shape.sphere(0.1))
    with pytest.raises(PhysicsError) as e:
        w.step_with_bodies(TICK, [pose_at(0.3)])
    assert e.value.name == "InvalidSettings"
    w.step(TICK)  # a plain step runs without the bodies
    w = still_pool(threads=1)
    w.couple((1, 0), Shape.sphere(0.1))
    with pytest.raises(PhysicsError) as e:
        w.step_with_bodies(TICK, [pose_at(0.3), pose_at(0.2)])
    assert e.value.name == "InvalidHandle"
    with pytest.raises(PhysicsError) as e:
        w.step_with_bodies(TICK, [pose_at(float("nan"))])
    assert e.value.name == "NonFinite"
    assert len(w.coupled()) == 1


def test_coupled_bodies_travel_with_the_snapshot():
    w = still_pool()
    w.couple((1, 0), Shape.cuboid((0.15, 0.15, 0.15)))
    w.couple((2, 3), Shape.cuboid((1.0, 0.05, 0.05)))  # coarse probes
    poses = lambda k: [pose_at(0.45 - 0.02 * k, (0.0, -1.2, 0.0)),  # noqa: E731
                       BodyPose((2, 3), (0.5, 0.1, 0.6), dynamic=False)]
    w.step_with_bodies(TICK, poses(0))
    saved = w.save_snapshot()
    loaded = OracleWorld.load_snapshot(saved, threads=8)
    assert loaded.save_snapshot() == saved
    for k in range(1, 4):
        assert w.step_with_bodies(TICK, poses(k)) == loaded.step_with_bodies(TICK, poses(k))
    assert w.state_hash() == loaded.state_hash()
    (_, _, _, s1), (_, _, _, s2) = loaded.coupled()
    assert s1 == np.float32(0.05) and s2 > 0.05


def test_coupled_bodies_out_of_order_in_a_snapshot_are_refused():
    import struct

    def fnv1a(data):
        h = 0xcbf29ce484222325
        for b in data:
            h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
        return h

    w = OracleWorld(FluidSettings())
    w.couple((1, 0), Shape.sphere(0.06))
    w.couple((2, 0), Shape.sphere(0.06))
    b = bytearray(w.save_snapshot()[:-8])
    first = b.index(struct.pack("<II", 1, 0))
    second = b.index(struct.pack("<II", 2, 0), first + 8)
    b[second:second + 8] = struct.pack("<II", 1, 0)
    b += struct.pack("<Q", fnv1a(bytes(b)))
    with pytest.raises(SnapshotError) as e:
        OracleWorld.load_snapshot(bytes(b))
    assert e.value == SnapshotError(6, 0, "fluid body order")
