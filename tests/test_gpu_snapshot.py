"""Fluid snapshots (SLFL version 3) of the device world.

The reference's own snapshot tests (src/fluid/save.rs:258-314) on the
device, the device writing the oracle's bytes for the same state, and
snapshots crossing between the two.
"""
import struct

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, FluidSettings, SnapshotError

pytestmark = pytest.mark.gpu

TICK = 1.0 / 60.0


def pool(w):
    # src/fluid/save.rs:207-229
    w.add_box_container((0, 0, 0), (1.0, 1.0, 0.6))
    w.add_wall((0.45, 0.0, 0.0), (0.55, 0.3, 0.6))
    w.fill_box((0, 0, 0), (0.4, 0.5, 0.6))
    w.add_emitter((0.05, 0.7, 0.2), (0.2, 0.8, 0.4), (1.0, 0.0, 0.0))
    w.add_drain((0.8, 0.0, 0.0), (1.0, 0.1, 0.6))
    return w


def fnv1a(data: bytes) -> int:
    h = 0xcbf29ce484222325
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_a_loaded_fluid_steps_on_exactly_as_the_saved_one():
    print("case: pool, 20 + 30 ticks", flush=True)
    fluid = pool(FluidWorld(FluidSettings()))
    for _ in range(20):
        fluid.step(TICK)
    saved = fluid.save_snapshot()
    loaded = FluidWorld.load_snapshot(saved)
    assert loaded.save_snapshot() == saved
    assert loaded.state_hash() == fluid.state_hash()
    for _ in range(30):
        assert fluid.step(TICK) == loaded.step(TICK)
    assert loaded.state_hash() == fluid.state_hash()
    assert len(loaded) == len(fluid)


@pytest.mark.parametrize("solver", [0, DFSPH])
def test_the_device_writes_the_oracles_bytes(solver):
    print(f"case: pool, solver {solver}", flush=True)
    s = FluidSettings(solver=solver)
    gpu = pool(FluidWorld(s))
    cpu = pool(OracleWorld(s, residual_mode=1, threads=8))
    for w in (gpu, cpu):
        w.set_domain(((-1.0, -1.0, -1.0), (2.0, 0.45, 2.0)))
    assert gpu.save_snapshot() == cpu.save_snapshot()
    for k in range(4):
        assert gpu.step(TICK) == cpu.step(TICK), k
        assert gpu.save_snapshot() == cpu.save_snapshot(), k


def test_snapshots_cross_between_the_device_and_the_oracle():
    print("case: pool crossing", flush=True)
    gpu = pool(FluidWorld(FluidSettings()))
    for _ in range(5):
        gpu.step(TICK)
    cpu = OracleWorld.load_snapshot(gpu.save_snapshot(), threads=8, residual_mode=1)
    for k in range(5):
        assert gpu.step(TICK) == cpu.step(TICK), k
    assert np.array_equal(gpu.positions(), cpu.positions())
    back = FluidWorld.load_snapshot(cpu.save_snapshot())
    assert back.state_hash() == cpu.state_hash() == gpu.state_hash()


def test_a_cut_or_flipped_fluid_snapshot_is_refused_without_a_panic():
    print("case: refusals", flush=True)
    fluid = FluidWorld(FluidSettings())
    fluid.add_box_container((0, 0, 0), (0.3, 0.3, 0.3))
    fluid.fill_box((0, 0, 0), (0.15, 0.15, 0.15))
    fluid.add_drain((0, 0, 0), (0.05, 0.05, 0.05))
    saved = fluid.save_snapshot()
    for end in range(0, len(saved), 7):
        with pytest.raises(SnapshotError):
            FluidWorld.load_snapshot(saved[:end])
    for at in range(0, len(saved), 5):
        flipped = bytearray(saved)
        flipped[at] ^= 0x41
        with pytest.raises(SnapshotError):
            FluidWorld.load_snapshot(bytes(flipped))
    bad = bytearray(saved[:-8])
    bad[8] = 7
    bad += struct.pack("<Q", fnv1a(bytes(bad)))
    with pytest.raises(SnapshotError) as e:
        FluidWorld.load_snapshot(bytes(bad))
    assert e.value == SnapshotError(6, 0, "fluid solver")
    later = bytearray(saved[:-8])
    later[4:8] = struct.pack("<I", 4)
    later += struct.pack("<Q", fnv1a(bytes(later)))
    with pytest.raises(SnapshotError) as e:
        FluidWorld.load_snapshot(bytes(later))
    assert e.value == SnapshotError(4, 4, "")
