"""The searched cell table on the device against the oracle.

The reference keeps a dense cell index only while the cells' box holds at
most 64 cells per occupied cell plus 4096, and otherwise binary-searches the
occupied cells (src/fluid/mod.rs:38-41, 214-236). Either index gives the
same neighbours in the same order. The device keeps a dense table while the
box holds at most max(64 n + 4096, 2^26) cells and past that searches the
occupied cells' 64-bit slots. A WCSPH world whose water spreads that thin,
and walls spanning that much space, must step bit for bit as the oracle
steps. DFSPH water keeps a dense table (its stop sum runs over the dense
box's tiles) and is refused past 2^31 cells.
"""
import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd.fluid import FluidBatch, FluidWorld, Tape
from solunar_amd.types import DFSPH, OUT_OF_RANGE, UNSUPPORTED, FluidSettings, PhysicsError

pytestmark = pytest.mark.gpu

FAR = 600.0  # metres: a box of (600 / 0.04)^3 ~ 3e12 cells at spacing 0.02


def pair(settings, build):
    print(f"case: {settings}", flush=True)
    gpu, cpu = FluidWorld(settings), OracleWorld(settings, threads=8, residual_mode=1)
    build(gpu)
    build(cpu)
    return gpu, cpu


def assert_same(gpu, cpu):
    assert np.array_equal(gpu.ids(), cpu.ids())
    assert np.array_equal(gpu.positions(), cpu.positions())
    assert np.array_equal(gpu.velocities(), cpu.velocities())
    assert np.array_equal(gpu.densities(), cpu.densities())
    assert gpu.state_hash() == cpu.state_hash()


def clumps(w):
    # two blocks of water far apart on every axis, nothing between them
    w.fill_box_jittered((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 0.05, 3)
    w.fill_box_jittered((FAR, FAR, FAR), (FAR + 0.16, FAR + 0.2, FAR + 0.12), 0.05, 5)


def test_two_clumps_far_apart_step_as_the_oracle_does():
    gpu, cpu = pair(FluidSettings(spacing=0.02), clumps)
    for k in range(8):
        rg, rc = gpu.step(1e-3), cpu.step(1e-3)
        assert rg == rc, (k, rg, rc)
    assert_same(gpu, cpu)
    # the parity hooks read the searched table as the step left it
    for w in (gpu, cpu):
        w.update_neighbors()
        w.compute_densities()
    assert_same(gpu, cpu)
    for a, b in ((gpu.neighbors(), cpu.neighbors()), (gpu.wall_neighbors(), cpu.wall_neighbors())):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for i in (0, 17, len(gpu) - 1):
        assert np.array_equal(gpu.neighbors_of(i), cpu.neighbors_of(i))


def test_an_emitter_grows_a_searched_world_as_the_oracle_does():
    # the searched table's buffers follow the particle count between steps
    def build(w):
        clumps(w)
        w.add_emitter((0.0, 0.3, 0.0), (0.1, 0.34, 0.1), (0.0, -1.0, 0.0))

    gpu, cpu = pair(FluidSettings(spacing=0.02), build)
    start, emitted = len(gpu), 0
    for k in range(6):
        rg, rc = gpu.step(2e-3), cpu.step(2e-3)
        assert rg == rc, (k, rg, rc)
        emitted += rg.emitted
    assert emitted > 0 and len(gpu) == start + emitted
    assert_same(gpu, cpu)


def containers(w):
    # two containers a kilometre apart, water in both: the walls' box and the
    # water's are both far past a dense table
    for origin in ((0.0, 0.0, 0.0), (1000.0, 0.0, 1000.0)):
        lo = np.asarray(origin, np.float32)
        w.add_box_container(tuple(lo), tuple(lo + np.float32([0.4, 0.4, 0.4])))
        w.fill_box_jittered(tuple(lo), tuple(lo + np.float32([0.2, 0.24, 0.2])), 0.05, 7)


def test_two_containers_a_kilometre_apart_step_as_the_oracle_does():
    gpu, cpu = pair(FluidSettings(spacing=0.02), containers)
    assert np.array_equal(gpu.wall_positions(), cpu.wall_positions())
    for k in range(8):
        rg, rc = gpu.step(2e-3), cpu.step(2e-3)
        assert rg == rc, (k, rg, rc)
    assert_same(gpu, cpu)


def test_water_drained_back_together_returns_to_a_dense_table():
    # a clump and one particle far off, which a drain takes out after the
    # first step: the table is searched, then dense again (no container
    # here: a container's bounds would pull the far particle in)
    def build(w):
        w.fill_box_jittered((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 0.05, 9)
        w.add_particles(np.array([[FAR, FAR, FAR]], np.float32))
        w.add_drain((FAR - 1.0, FAR - 1.0, FAR - 1.0), (FAR + 1.0, FAR + 1.0, FAR + 1.0))

    gpu, cpu = pair(FluidSettings(spacing=0.02), build)
    drained = []
    for k in range(6):
        rg, rc = gpu.step(1e-3), cpu.step(1e-3)
        assert rg == rc, (k, rg, rc)
        drained.append(rg.drained)
    assert drained == [1, 0, 0, 0, 0, 0] and len(gpu) == len(cpu) == 1000
    assert_same(gpu, cpu)


def test_dfsph_with_far_walls_and_refused_far_water():
    # walls far apart go to the searched table under DFSPH too; the water
    # stays dense
    def build(w):
        w.add_box_container((0.0, 0.0, 0.0), (0.4, 0.4, 0.4))
        w.add_wall((FAR, FAR, FAR), (FAR + 0.1, FAR + 0.1, FAR + 0.1))
        w.fill_box_jittered((0.0, 0.0, 0.0), (0.2, 0.24, 0.2), 0.05, 11)

    gpu, cpu = pair(FluidSettings(solver=DFSPH, spacing=0.02), build)
    for k in range(5):
        rg, rc = gpu.step(5e-4), cpu.step(5e-4)
        assert rg == rc, (k, rg, rc)
    assert_same(gpu, cpu)
    w = FluidWorld(FluidSettings(solver=DFSPH, spacing=0.02))
    clumps(w)
    before = w.state_hash()
    with pytest.raises(PhysicsError) as e:
        w.step(5e-4)
    assert e.value.code == OUT_OF_RANGE
    assert w.state_hash() == before


def test_the_batch_and_the_tape_refuse_a_searched_world():
    w = FluidWorld(FluidSettings(spacing=0.02))
    clumps(w)
    w.step(1e-3)  # the sort goes to the searched table
    with pytest.raises(PhysicsError) as e:
        FluidBatch([w])
    assert e.value.code == UNSUPPORTED
    with pytest.raises(PhysicsError) as e:
        Tape(w, 1e-3, 2, every=1)
    assert e.value.code == UNSUPPORTED


def test_coupled_bodies_over_two_far_pools_push_as_the_oracle_does():
    # a fast ball over the far pool asks the searched table whether water
    # is in its reach (the sound speed of the step depends on it); a slow
    # box sits in the near pool
    from solunar_amd.types import BodyPose, Shape

    tick = 1.0 / 60.0

    def build(w):
        for x0 in (0.0, FAR):
            w.add_box_container((x0, 0.0, x0), (x0 + 1.0, 1.2, x0 + 1.0))
            w.fill_box((x0, 0.0, x0), (x0 + 1.0, 0.6, x0 + 1.0))
        w.couple((1, 0), Shape.sphere(0.1))
        w.couple((2, 0), Shape.cuboid((0.1, 0.08, 0.1)))

    gpu, cpu = pair(FluidSettings(), build)
    speed = 24.0
    for k in range(4):
        poses = [BodyPose((1, 0), (FAR + 0.5, 0.9 - speed * tick * k, FAR + 0.5), linear_velocity=(0.0, -speed, 0.0)),
                 BodyPose((2, 0), (0.5, 0.55, 0.5), linear_velocity=(0.1, 0.0, 0.0))]
        g = gpu.step_with_bodies(tick, poses)
        c = cpu.step_with_bodies(tick, poses)
        assert g[0] == c[0], k
        assert g[1] == c[1], k
    assert g[0].sound_speed == np.float32(4 * speed)
    assert_same(gpu, cpu)


def test_a_large_world_on_the_searched_table_steps_as_the_oracle_does():
    # ~155k particles in two blocks 600 m apart: the radix sort and the
    # binary-searched walks at a size where every kernel runs many blocks
    def build(w):
        w.add_box_container((0.0, 0.0, 0.0), (0.5, 0.6, 0.5))
        w.fill_box_jittered((0.0, 0.0, 0.0), (0.3, 0.45, 0.3), 0.05, 13)
        w.fill_box_jittered((FAR, FAR, FAR), (FAR + 0.3, FAR + 0.3, FAR + 0.3), 0.05, 17)

    gpu, cpu = pair(FluidSettings(spacing=0.0075), build)
    assert len(gpu) > 150000
    for k in range(2):
        rg, rc = gpu.step(1e-4), cpu.step(1e-4)
        assert rg == rc, (k, rg, rc)
    assert_same(gpu, cpu)
