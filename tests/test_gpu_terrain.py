"""Terrain ground (src/fluid/terrain.rs) on the device, against the oracle on
the same field: reports, ids, positions and velocities bit for bit, step
after step; the ground's densities through the parity hooks; refusals;
snapshots with terrain crossing between the two.
"""
import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, FluidSettings, PhysicsError

pytestmark = pytest.mark.gpu

TICK = 1.0 / 60.0


def pair(settings=None):
    s = settings or FluidSettings()
    gpu, cpu = FluidWorld(s), OracleWorld(s, threads=8)
    for w in (gpu, cpu):
        scenes.in_trench(w)
    return gpu, cpu


def same(gpu, cpu, k):
    g, c = gpu.step(TICK), cpu.step(TICK)
    assert g == c, k
    assert np.array_equal(gpu.ids(), cpu.ids()), k
    assert np.array_equal(gpu.positions(), cpu.positions()), k
    assert np.array_equal(gpu.velocities(), cpu.velocities()), k
    return g


def test_water_set_down_in_a_trench_steps_as_the_oracle_does():
    print("case: trench, set down", flush=True)
    gpu, cpu = pair()
    for w in (gpu, cpu):
        w.fill_box((-0.1, -0.4, -0.3), (0.5, -0.1, 0.3))
    r = same(gpu, cpu, 0)
    assert r.max_speed < 0.5 and r.max_compression < 0.05


def test_water_resting_in_a_trench_steps_as_the_oracle_does():
    print("case: trench, 30 ticks", flush=True)
    gpu, cpu = pair()
    for w in (gpu, cpu):
        w.fill_box((0.0, -0.4, -0.3), (0.6, -0.1, 0.3))
    for k in range(30):
        same(gpu, cpu, k)
    assert gpu.state_hash() == cpu.state_hash()


def test_water_thrown_at_the_ground_is_put_back_as_the_oracle_puts_it():
    # fast water into the bed and the side walls, the lowest layer starting
    # under the bed: the contact puts it back and counts what was under, with
    # the walls of a container around it all
    print("case: thrown into the trench", flush=True)
    gpu, cpu = pair()
    rng = np.random.default_rng(3)
    x = np.stack(np.meshgrid(np.arange(0.3, 0.7, 0.05), np.arange(-0.43, 0.1, 0.05), np.arange(-0.25, 0.25, 0.05),
                             indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    v = (rng.uniform(-1, 1, x.shape) * np.float32([5.0, 2.0, 20.0]) + np.float32([0.0, -40.0, 0.0])).astype(np.float32)
    for w in (gpu, cpu):
        w.add_box_container((-0.4, -0.6, -0.45), (2.2, 0.6, 0.45))
        w.add_particles(x, v)
    held = 0
    for k in range(6):
        held += same(gpu, cpu, k).held_back
    assert held > 0


def test_the_grounds_density_through_the_parity_hooks():
    print("case: densities", flush=True)
    gpu, cpu = pair()
    for w in (gpu, cpu):
        w.fill_box((0.0, -0.4, -0.3), (0.6, -0.1, 0.3))
        w.update_neighbors()
        w.compute_densities()
    assert np.array_equal(gpu.densities(), cpu.densities())


def test_terrain_travels_with_the_snapshot_and_is_refused_as_the_reference_refuses():
    print("case: snapshots and refusals", flush=True)
    gpu, cpu = pair()
    for w in (gpu, cpu):
        w.fill_box((0.0, -0.4, -0.3), (0.4, -0.2, 0.3))
    for k in range(3):
        same(gpu, cpu, k)
    assert gpu.save_snapshot() == cpu.save_snapshot()
    back = FluidWorld.load_snapshot(cpu.save_snapshot())
    assert back.terrain_len() == 1
    for k in range(3):
        assert back.step(TICK) == cpu.step(TICK), k
    assert back.state_hash() == cpu.state_hash()
    d = FluidWorld(FluidSettings(solver=DFSPH))
    scenes.in_trench(d)
    d.fill_box((0.0, -0.4, -0.3), (0.3, -0.2, 0.3))
    with pytest.raises(PhysicsError) as e:
        d.step(TICK)
    assert e.value.name == "InvalidSettings"
    w = FluidWorld(FluidSettings())
    with pytest.raises(PhysicsError) as e:
        w.add_terrain((0, 0, 0), (1, 1, 1), np.zeros(7))
    assert e.value.name == "OutOfRange"
    t = w.add_terrain((0, 0, 0), (1, 1, 1), scenes.trench_distance)
    assert w.terrain_len() == 1 and w.remove_boundary(t) and w.terrain_len() == 0
