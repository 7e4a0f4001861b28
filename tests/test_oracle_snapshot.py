"""Fluid snapshots (SLFL version 3) in the CPU oracle.

The reference's own snapshot tests (src/fluid/save.rs:237-314) on the
oracle's restatement, plus the header's layout read back field by field as
save.rs:20-40 writes it.
"""
import struct

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd.types import DFSPH, FluidSettings, PhysicsError, SnapshotError

TICK = 1.0 / 60.0


def pool(settings=None):
    # src/fluid/save.rs:207-229
    w = OracleWorld(settings or FluidSettings(), threads=8)
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


def small():
    w = OracleWorld(FluidSettings())
    w.add_box_container((0, 0, 0), (0.3, 0.3, 0.3))
    w.fill_box((0, 0, 0), (0.15, 0.15, 0.15))
    w.add_drain((0, 0, 0), (0.05, 0.05, 0.05))
    return w


def test_a_loaded_fluid_steps_on_exactly_as_the_saved_one():
    fluid = pool()
    for _ in range(20):
        fluid.step(TICK)
    saved = fluid.save_snapshot()
    loaded = OracleWorld.load_snapshot(saved, threads=8)
    assert loaded.save_snapshot() == saved
    assert loaded.state_hash() == fluid.state_hash()
    for _ in range(30):
        fluid.step(TICK)
        loaded.step(TICK)
    assert loaded.state_hash() == fluid.state_hash()
    assert len(loaded) == len(fluid)


def test_a_cut_or_flipped_fluid_snapshot_is_refused_without_a_panic():
    saved = small().save_snapshot()
    for end in range(len(saved)):
        with pytest.raises(SnapshotError):
            OracleWorld.load_snapshot(saved[:end])
    for at in range(len(saved)):
        flipped = bytearray(saved)
        flipped[at] ^= 0x41
        with pytest.raises(SnapshotError):
            OracleWorld.load_snapshot(bytes(flipped))
    # a good checksum over bad content is still refused
    bad = bytearray(saved[:-8])
    bad[8] = 7  # the solver tag, right after magic and version
    bad += struct.pack("<Q", fnv1a(bytes(bad)))
    with pytest.raises(SnapshotError) as e:
        OracleWorld.load_snapshot(bytes(bad))
    assert e.value == SnapshotError(6, 0, "fluid solver")
    later = bytearray(saved[:-8])
    later[4:8] = struct.pack("<I", 4)
    later += struct.pack("<Q", fnv1a(bytes(later)))
    with pytest.raises(SnapshotError) as e:
        OracleWorld.load_snapshot(bytes(later))
    assert e.value == SnapshotError(4, 4, "")


def test_the_header_holds_the_settings_as_save_rs_writes_them():
    s = FluidSettings(solver=DFSPH, spacing=0.02, viscosity=0.07, max_substeps=77, max_particles=12345)
    w = OracleWorld(s)
    w.fill_box((0, 0, 0), (0.1, 0.1, 0.1))
    b = w.save_snapshot()
    assert b[:4] == b"SLFL" and struct.unpack_from("<I", b, 4)[0] == 3 and b[8] == 1
    f = np.frombuffer(b[9:33], "<f4")
    assert list(f) == list(np.float32([s.spacing, s.rest_density, s.support, s.sound_speed, s.viscosity, s.courant]))
    assert struct.unpack_from("<Q", b, 33)[0] == 12345
    assert list(np.frombuffer(b[41:53], "<f4")) == list(np.float32(s.gravity))
    assert struct.unpack_from("<I", b, 53)[0] == 77
    n = struct.unpack_from("<I", b, 57)[0]
    assert n == len(w)
    # particles: id, position, velocity, 28 bytes each, in the world's order
    first = struct.unpack_from("<I3f3f", b, 61)
    assert first[0] == w.ids()[0] and np.array_equal(np.float32(first[1:4]), w.positions()[0])
    assert struct.unpack_from("<Q", b, len(b) - 8)[0] == fnv1a(b[:-8])


def test_the_domain_and_the_flows_travel_with_the_snapshot():
    w = pool()
    w.set_domain(((-1.0, -1.0, -1.0), (2.0, 0.3, 2.0)))
    w.step(TICK)
    loaded = OracleWorld.load_snapshot(w.save_snapshot())
    r1, r2 = w.step(TICK), loaded.step(TICK)
    assert r1 == r2 and r1.emitted > 0
    assert w.state_hash() == loaded.state_hash()
    # an emptied domain flag byte is refused with a good checksum
    b = bytearray(w.save_snapshot()[:-8])
    b[-25] = 9  # the domain flag: 1 byte before two vec3
    b += struct.pack("<Q", fnv1a(bytes(b)))
    with pytest.raises(SnapshotError) as e:
        OracleWorld.load_snapshot(bytes(b))
    assert e.value == SnapshotError(6, 0, "fluid domain flag")


def test_a_coupled_body_without_points_is_refused():
    b = bytearray(small().save_snapshot()[:-8])
    # the coupled count sits after the flows: 4 + 4 + 1 (no domain flag set) from the end
    assert b[-9:] == b"\0" * 9
    b[-9:-5] = struct.pack("<I", 1)
    b[-5:-5] = b"\0" * 12  # handle (0, 0), no points
    b += struct.pack("<Q", fnv1a(bytes(b)))
    with pytest.raises(SnapshotError) as e:
        OracleWorld.load_snapshot(bytes(b))
    assert e.value == SnapshotError(6, 0, "fluid body")
