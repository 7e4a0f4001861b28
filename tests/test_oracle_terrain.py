"""Terrain ground (src/fluid/terrain.rs) in the CPU oracle.

The reference's own terrain tests (terrain.rs:389-536) on the oracle's
restatement. Their terrain comes from the reference's signed-distance
volumes, which are not part of this path: here the same trench is an exact
field (`scenes.trench_distance`), sampled at the nodes add_terrain samples.
"""
import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.types import DFSPH, FluidSettings, PhysicsError

TICK = 1.0 / 60.0


def trenched(threads=8, settings=None):
    w = OracleWorld(settings or FluidSettings(), threads=threads)
    scenes.in_trench(w)
    return w


def test_the_field_matches_the_terrain_between_nodes():
    w = trenched(1)
    for at in [(0.5, -0.3, 0.0), (1.03, 0.07, 0.41), (0.2, -0.5, 0.2)]:
        distance, _ = w.terrain_sample(0, at)
        truth = float(scenes.trench_distance([at])[0])
        assert abs(distance - truth) < 0.02, (at, distance, truth)
    # on the trench bed the way out is up, on its side wall it is across
    _, up = w.terrain_sample(0, (1.0, -0.38, 0.0))
    assert up[1] > 0.9
    _, across = w.terrain_sample(0, (1.0, -0.2, 0.28))
    assert across[2] < -0.9
    assert w.terrain_sample(0, (9.0, 0.0, 0.0)) is None


def test_ground_gives_what_a_box_wall_gives():
    box = OracleWorld(FluidSettings())
    box.add_box_container((0, 0, 0), (0.5, 0.5, 0.5))
    at = np.array([0.225, 0.025, 0.225])
    walls = box.wall_positions().astype(np.float64)
    r = FluidSettings().radius()
    d = np.linalg.norm(walls - at, axis=1)
    d = d[d < r]
    q = d / r
    wall = box.particle_mass() * (21.0 / (2 * np.pi * r ** 3) * (1 - q) ** 4 * (1 + 4 * q)).sum()
    w = trenched(1)
    density, gradient = w.ground_near((1.0, -0.375, 0.0))
    assert abs(density - wall) / wall < 0.05, (density, wall)
    # it pushes up, away from the bed. The exact field puts every site half a
    # spacing off the particle exactly in or out of the ground there (no
    # partial share, so no gradient), which the reference's voxel field never
    # does: look 5 mm higher, where the sites straddle the surface
    _, gradient = w.ground_near((1.0, -0.37, 0.0))
    assert gradient[1] < 0.0 and abs(gradient[0]) < abs(gradient[1]) * 0.1
    # in the corner of bed and side there is more of it
    corner, _ = w.ground_near((1.0, -0.375, 0.275))
    assert corner > density * 1.2


def test_water_set_down_in_a_trench_does_not_jump():
    w = trenched()
    w.fill_box((-0.1, -0.4, -0.3), (0.5, -0.1, 0.3))
    r = w.step(TICK)
    assert r.max_speed < 0.5 and r.max_compression < 0.05, r


def test_water_rests_in_a_trench_and_does_not_sink_into_it():
    w = trenched()
    w.fill_box((0.0, -0.4, -0.3), (0.6, -0.1, 0.3))
    count = len(w)
    held = 0
    for _ in range(150):
        held += w.step(TICK).held_back
    x = w.positions()
    assert len(w) == count
    assert held == 0, "water got under the ground"
    assert x[:, 1].min() > -0.4
    # it spread along the trench, and stayed between its walls
    assert x[:, 0].max() > 1.0
    assert (np.abs(x[:, 2]) < 0.35).all()


def test_a_terrain_region_is_checked_and_can_go_away():
    w = OracleWorld(FluidSettings())
    with pytest.raises(PhysicsError):
        w.add_terrain((1, 1, 1), (0, 0, 0), scenes.trench_distance)
    with pytest.raises(PhysicsError) as e:
        w.add_terrain((np.nan, 0, 0), (1, 1, 1), np.zeros(8))
    assert e.value.name == "NonFinite"
    with pytest.raises(PhysicsError) as e:
        w.add_terrain((0, 0, 0), (1e6, 1e6, 1e6), np.zeros(8))
    assert e.value.name == "FluidFull"
    with pytest.raises(PhysicsError) as e:
        w.add_terrain((0, 0, 0), (1, 1, 1), np.zeros(7))  # not the lattice's count
    assert e.value.name == "OutOfRange"
    t = w.add_terrain((0, 0, 0), (1, 1, 1), scenes.trench_distance)
    assert w.terrain_len() == 1
    assert w.remove_boundary(t)
    assert w.terrain_len() == 0


def test_terrain_fields_come_back_from_a_snapshot():
    w = trenched()
    w.fill_box((0.6, 0.05, -0.2), (1.0, 0.4, 0.2))
    for _ in range(10):
        w.step(TICK)
    loaded = OracleWorld.load_snapshot(w.save_snapshot(), threads=8)
    assert loaded.save_snapshot() == w.save_snapshot()
    for _ in range(20):
        assert w.step(TICK) == loaded.step(TICK)
    assert loaded.state_hash() == w.state_hash()


def test_the_divergence_free_solver_refuses_ground():
    w = trenched(1, FluidSettings(solver=DFSPH))
    w.fill_box((0.0, -0.4, -0.3), (0.3, -0.2, 0.3))
    with pytest.raises(PhysicsError) as e:
        w.step(TICK)
    assert e.value.name == "InvalidSettings"


def test_terrain_steps_are_the_same_on_any_thread_count():
    def go(threads):
        w = trenched(threads)
        w.fill_box((-0.1, -0.4, -0.3), (0.5, -0.1, 0.3))
        for _ in range(8):
            w.step(TICK)
        return w.state_hash()

    assert go(1) == go(7)
