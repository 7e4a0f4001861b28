"""Flows (emitters, drains) and the domain on the device against the oracle.

Scenes from the reference's own tests (src/fluid/flow.rs:225-339,
src/fluid/save.rs:207-256). Each test steps the device and the oracle side
by side and requires the same report (emitted, drained, lost included), the
same ids in the same order and bit-equal positions and velocities.
"""
import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, FluidSettings, PhysicsError

pytestmark = pytest.mark.gpu

TICK = 1.0 / 60.0


def both(settings):
    print(f"case: {settings}", flush=True)
    return FluidWorld(settings), OracleWorld(settings, residual_mode=1, threads=8)


def same_state(gpu, cpu, where):
    assert np.array_equal(gpu.ids(), cpu.ids()), where
    assert np.array_equal(gpu.positions(), cpu.positions()), where
    assert np.array_equal(gpu.velocities(), cpu.velocities()), where


def channel(w):
    w.add_box_container((0, 0, 0), (2.0, 0.6, 0.3))


def pool(w):
    # src/fluid/save.rs:207-229
    w.add_box_container((0, 0, 0), (1.0, 1.0, 0.6))
    w.add_wall((0.45, 0.0, 0.0), (0.55, 0.3, 0.6))
    w.fill_box((0, 0, 0), (0.4, 0.5, 0.6))
    w.add_emitter((0.05, 0.7, 0.2), (0.2, 0.8, 0.4), (1.0, 0.0, 0.0))
    w.add_drain((0.8, 0.0, 0.0), (1.0, 0.1, 0.6))


def test_an_emitter_fills_its_box_once_and_tops_it_up():
    gpu, cpu = both(FluidSettings())
    for w in (gpu, cpu):
        channel(w)
        w.add_emitter((0, 0, 0), (0.2, 0.3, 0.3), (0, 0, 0))
    total = 0
    for k in range(31):
        rg, rc = gpu.step(TICK), cpu.step(TICK)
        assert rg == rc, (k, rg, rc)
        if k == 0:
            assert rg.emitted == 4 * 6 * 6 and len(gpu) == 4 * 6 * 6
        total += rg.emitted
        same_state(gpu, cpu, k)
    assert len(gpu) == total > 4 * 6 * 6


def test_a_river_from_an_emitter_to_a_drain_matches_the_oracle():
    gpu, cpu = both(FluidSettings())
    for w in (gpu, cpu):
        channel(w)
        w.add_emitter((0, 0, 0), (0.15, 0.3, 0.3), (1.0, 0.0, 0.0))
        w.add_drain((1.8, 0.0, 0.0), (2.0, 0.6, 0.3))
    emitted = drained = 0
    for k in range(120):
        rg, rc = gpu.step(TICK), cpu.step(TICK)
        assert rg == rc, (k, rg, rc)
        emitted += rg.emitted
        drained += rg.drained
    same_state(gpu, cpu, "end")
    assert len(gpu) == emitted - drained
    assert gpu.state_hash() == cpu.state_hash()


@pytest.mark.parametrize("solver", [0, DFSPH])
def test_the_pool_with_its_domain_matches_the_oracle(solver):
    gpu, cpu = both(FluidSettings(solver=solver))
    for w in (gpu, cpu):
        pool(w)
        w.set_domain(((-1.0, -1.0, -1.0), (2.0, 0.3, 2.0)))
    lost = 0
    for k in range(6):
        rg, rc = gpu.step(TICK), cpu.step(TICK)
        assert rg == rc, (k, rg, rc)
        lost += rg.lost
        same_state(gpu, cpu, k)
    assert lost > 0
    assert (gpu.positions()[:, 1] <= np.float32(0.3)).all()
    gpu.set_domain(None)
    cpu.set_domain(None)
    rg, rc = gpu.step(TICK), cpu.step(TICK)
    assert rg == rc and rg.lost == 0
    same_state(gpu, cpu, "no domain")


def test_removed_flows_stop_and_a_zero_step_runs_none():
    gpu, cpu = both(FluidSettings())
    ids = []
    for w in (gpu, cpu):
        pool(w)
        ids.append(w.add_drain((0.0, 0.0, 0.0), (0.1, 0.1, 0.6)))
    assert ids[0] == ids[1] == 2
    for w in (gpu, cpu):
        assert w.remove_flow(0) and not w.remove_flow(0)
    rg, rc = gpu.step(0.0), cpu.step(0.0)
    assert rg == rc and (rg.emitted, rg.drained, rg.lost) == (0, 0, 0)
    rg, rc = gpu.step(TICK), cpu.step(TICK)
    assert rg == rc and rg.emitted == 0 and rg.drained > 0
    same_state(gpu, cpu, "drains only")


def test_a_full_world_stops_the_emitter_with_an_error():
    gpu, cpu = both(FluidSettings(max_particles=10))
    for w in (gpu, cpu):
        w.add_emitter((0, 0, 0), (0.2, 0.2, 0.2), (0, 0, 0))
        with pytest.raises(PhysicsError) as e:
            w.step(TICK)
        assert e.value.name == "FluidFull"
        assert len(w) == 10
    same_state(gpu, cpu, "full")
