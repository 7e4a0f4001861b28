"""Batched WCSPH: every scene of a batch steps bit-equal to its world alone.

The single-world device step is pinned to the CPU oracle by
test_gpu_parity.py; here each batch scene is compared with the same world
stepped on its own (ids, positions, velocities, densities and the last
report, all exact), and one small batch is compared with the oracle
directly. The cases cover uneven scene sizes, empty scenes, scenes that take
different substep counts in one step (the slot sequence grows), a scene
without a container whose cell box outgrows the first table (the table
grows and the graph is captured again), graph against plain launches, the
autotuner leaving the state alone, the hand-back to worlds, and a
non-finite scene next to a healthy one.
"""
import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.fluid import FluidBatch, FluidWorld
from solunar_amd.types import FluidSettings, PhysicsError

pytestmark = pytest.mark.gpu


def same_scene(batch, k, world, report=None, wreport=None):
    ids, x, v, rho = batch.state(k)
    assert np.array_equal(ids, world.ids()), f"scene {k}: order differs"
    assert np.array_equal(x, world.positions()), f"scene {k}: positions differ"
    assert np.array_equal(v, world.velocities()), f"scene {k}: velocities differ"
    assert np.array_equal(rho, world.densities()), f"scene {k}: densities differ"
    if report is not None:
        assert report == wreport, f"scene {k}: report {report} != {wreport}"


def small_batch(count=6, seed=11):
    # smaller blocks than the benchmark's so the test stays quick
    blocks = scenes.batch_blocks(count, seed, 0.02, low=200, high=6000)
    return [b for b, _ in blocks]


def make(settings, sides, k):
    w = FluidWorld(settings)
    scenes.batch_scene(w, sides, k)
    return w


def test_uneven_scenes_step_like_their_worlds():
    s = scenes.settings_dam_small()
    blocks = small_batch()
    print(f"case: {len(blocks)} scenes", flush=True)
    worlds = [make(s, b, k) for k, b in enumerate(blocks)]
    alone = [make(s, b, k) for k, b in enumerate(blocks)]
    worlds.insert(2, FluidWorld(s))  # an empty scene
    alone.insert(2, FluidWorld(s))
    batch = FluidBatch(worlds)
    assert batch.scenes() == len(worlds) and batch.len(2) == 0
    reports = batch.step(1e-4, 25)
    last = []
    for w in alone:
        for _ in range(25):
            r = w.step(1e-4)
        last.append(r)
    assert reports[2].substeps == 0
    for k, w in enumerate(alone):
        same_scene(batch, k, w, reports[k], last[k])
        assert batch.state_hash(k) == w.state_hash()


def test_a_batch_scene_matches_the_oracle():
    s = scenes.settings_dam_small()
    blocks = small_batch(3, seed=5)
    print("case: oracle batch", flush=True)
    worlds = [make(s, b, k) for k, b in enumerate(blocks)]
    cpus = []
    for k, b in enumerate(blocks):
        o = OracleWorld(s, threads=8)
        scenes.batch_scene(o, b, k)
        cpus.append(o)
    batch = FluidBatch(worlds)
    batch.step(1e-4, 10)
    for k, o in enumerate(cpus):
        for _ in range(10):
            o.step(1e-4)
        ids, x, v, rho = batch.state(k)
        assert np.array_equal(ids, o.ids())
        assert np.array_equal(x, o.positions())
        assert np.array_equal(v, o.velocities())
        assert np.array_equal(rho, o.densities())


def test_scenes_with_different_substep_counts():
    # a fast scene takes several substeps per step while the others take one
    s = scenes.settings_dam_small()
    blocks = small_batch(4, seed=3)
    print("case: mixed substeps", flush=True)

    def build():
        ws = [make(s, b, k) for k, b in enumerate(blocks)]
        fast = FluidWorld(s)
        fast.add_box_container((0, 0, 0), scenes.BOX)
        pts = np.array([[0.3 + 0.02 * i, 0.3, 0.3] for i in range(20)], np.float32)
        vel = np.tile(np.array([[30.0, 5.0, 0.0]], np.float32), (20, 1))
        fast.add_particles(pts, vel)
        ws.insert(1, fast)
        return ws

    worlds, alone = build(), build()
    batch = FluidBatch(worlds)
    reports = batch.step(1e-3, 6)
    assert reports[1].substeps > reports[0].substeps >= 1
    for w in alone:
        for _ in range(6):
            r = w.step(1e-3)
    for k, w in enumerate(alone):
        same_scene(batch, k, w)


def test_a_scene_without_container_grows_the_cell_table():
    s = scenes.settings_dam_small()
    print("case: table growth", flush=True)

    def build():
        a = make(s, (0.3, 0.3, 0.3), 0)
        b = FluidWorld(s)
        pts = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [10.0, 1.0, -5.0], [10.01, 1.0, -5.0]], np.float32)
        b.add_particles(pts)
        return [a, b]

    worlds, alone = build(), build()
    batch = FluidBatch(worlds)
    batch.step(1e-4, 5)
    for w in alone:
        for _ in range(5):
            w.step(1e-4)
    for k, w in enumerate(alone):
        same_scene(batch, k, w)


def test_graph_and_plain_launches_agree_and_tuning_keeps_the_state():
    s = scenes.settings_dam_small()
    blocks = small_batch(5, seed=21)
    print("case: graph vs plain", flush=True)
    one = FluidBatch([make(s, b, k) for k, b in enumerate(blocks)])
    two = FluidBatch([make(s, b, k) for k, b in enumerate(blocks)])
    two.use_graph(False)
    before = [one.state_hash(k) for k in range(len(blocks))]
    t = one.autotune(1e-4, repeats=2)
    assert t["sort_block"] in t["blocks"] and t["force_block"] in t["blocks"]
    assert all(ms > 0 for ms in t["force_ms"])
    assert [one.state_hash(k) for k in range(len(blocks))] == before
    r1 = one.step(1e-4, 12)
    r2 = two.step(1e-4, 12)
    assert r1 == r2
    assert one.graph_nodes() > 0
    for k in range(len(blocks)):
        a, b = one.state(k), two.state(k)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    # another shape gives the same bits
    two.set_tuning(64, 1024, 128)
    two.use_graph(True)
    one.step(1e-4, 3)
    two.step(1e-4, 3)
    for k in range(len(blocks)):
        assert one.state_hash(k) == two.state_hash(k)


def test_store_hands_scenes_back_to_worlds():
    s = scenes.settings_dam_small()
    blocks = small_batch(3, seed=8)
    print("case: store", flush=True)
    worlds = [make(s, b, k) for k, b in enumerate(blocks)]
    alone = [make(s, b, k) for k, b in enumerate(blocks)]
    batch = FluidBatch(worlds)
    batch.step(1e-4, 4)
    batch.store(worlds)
    for w in worlds:
        for _ in range(3):
            w.step(1e-4)
    for w in alone:
        for _ in range(7):
            w.step(1e-4)
    for w, a in zip(worlds, alone):
        assert w.state_hash() == a.state_hash()


def test_a_non_finite_scene_is_named_and_the_others_finish():
    s = FluidSettings(spacing=0.02, max_substeps=1)
    print("case: non-finite", flush=True)

    def build():
        calm = FluidWorld(s)
        calm.add_particle((0.0, 1.0, 0.0))
        wild = FluidWorld(s)
        wild.add_particle((0.0, 0.0, 0.0), (1e38, 0.0, 0.0))
        return [calm, wild]

    worlds, alone = build(), build()
    batch = FluidBatch(worlds)
    with pytest.raises(PhysicsError) as e:
        batch.step(10.0, 1)
    assert e.value.name == "NonFinite" and e.value.detail == "fluid particle"
    assert not batch.scene_failed(0) and batch.scene_failed(1)
    alone[0].step(10.0)
    same_scene(batch, 0, alone[0])


def test_batch_refuses_mixed_settings():
    a = FluidWorld(scenes.settings_dam_small())
    b = FluidWorld(FluidSettings(spacing=0.03))
    with pytest.raises(PhysicsError) as e:
        FluidBatch([a, b])
    assert e.value.name == "InvalidSettings"
    # flows and a domain run in World::step, which a batch does not call
    c = FluidWorld(scenes.settings_dam_small())
    c.add_drain((0, 0, 0), (0.1, 0.1, 0.1))
    with pytest.raises(PhysicsError) as e:
        FluidBatch([a, c])
    assert e.value.name == "Unsupported"
