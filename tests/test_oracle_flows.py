"""Flows (emitters and drains) and the domain in the CPU oracle.

The reference's own tests of src/fluid/flow.rs (:225-339) and of the domain
(src/fluid/save.rs:237-256), run on the oracle's restatement with the same
scenes and the same assertions, plus the bookkeeping the step report adds up.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd.types import FluidSettings, PhysicsError

TICK = 1.0 / 60.0


def channel(**kw):
    w = OracleWorld(FluidSettings(**kw), threads=8)
    w.add_box_container((0, 0, 0), (2.0, 0.6, 0.3))
    return w


def pool(**kw):
    # src/fluid/save.rs:207-229
    w = OracleWorld(FluidSettings(**kw), threads=8)
    w.add_box_container((0, 0, 0), (1.0, 1.0, 0.6))
    w.add_wall((0.45, 0.0, 0.0), (0.55, 0.3, 0.6))
    w.fill_box((0, 0, 0), (0.4, 0.5, 0.6))
    w.add_emitter((0.05, 0.7, 0.2), (0.2, 0.8, 0.4), (1.0, 0.0, 0.0))
    w.add_drain((0.8, 0.0, 0.0), (1.0, 0.1, 0.6))
    return w


def test_an_emitter_fills_its_box_once_and_tops_it_up():
    w = channel()
    w.add_emitter((0, 0, 0), (0.2, 0.3, 0.3), (0, 0, 0))
    r = w.step(TICK)
    assert r.emitted == 4 * 6 * 6
    assert len(w) == 4 * 6 * 6
    total = r.emitted
    for _ in range(30):
        total += w.step(TICK).emitted
    assert len(w) == total
    assert total > 4 * 6 * 6
    # emitted particles take the next ids in slot order
    assert sorted(w.ids().tolist()) == list(range(total))


def test_a_river_from_an_emitter_to_a_drain_keeps_its_count_and_books():
    w = channel()
    w.add_emitter((0, 0, 0), (0.15, 0.3, 0.3), (1.0, 0.0, 0.0))
    w.add_drain((1.8, 0.0, 0.0), (2.0, 0.6, 0.3))
    emitted = drained = 0
    counts = []
    for _ in range(240):
        r = w.step(TICK)
        emitted += r.emitted
        drained += r.drained
        assert r.held_back == 0
        counts.append(len(w))
    assert len(w) == emitted - drained
    assert drained > 0, "water reached the drain"
    late = counts[180:]
    spread = (max(late) - min(late)) / max(late)
    assert spread < 0.15, (min(late), max(late))
    assert (w.positions()[:, 0] < 1.8 + 0.1).all()


def test_an_emitter_never_puts_water_on_top_of_water():
    w = channel()
    spacing = np.float32(w.settings.spacing)
    w.add_particles([(float(spacing * np.float32(1.05)), float(spacing * np.float32(0.5)),
                      float(spacing * np.float32(0.5)))], [(0, 0, 0)])
    w.add_emitter((0, 0, 0), (0.1, 0.05, 0.05), (0, 0, 0))
    w.run_flows()
    p = w.positions().astype(np.float64)
    for a in range(len(p)):
        for b in range(a + 1, len(p)):
            assert np.linalg.norm(p[a] - p[b]) > 0.5 * spacing, (p[a], p[b])


def test_flows_refuse_bad_boxes_and_go_away():
    w = channel()
    with pytest.raises(PhysicsError):
        w.add_drain((1, 1, 1), (0, 0, 0))
    nan = (math.nan, 0.0, 0.0)
    with pytest.raises(PhysicsError):
        w.add_emitter((0, 0, 0), (1, 1, 1), nan)
    with pytest.raises(PhysicsError):
        w.add_drain(nan, (1, 1, 1))
    drain = w.add_drain((0, 0, 0), (1, 1, 1))
    assert w.remove_flow(drain)
    assert not w.remove_flow(drain)


def test_a_full_world_stops_the_emitter_with_an_error():
    w = OracleWorld(FluidSettings(max_particles=10))
    w.add_emitter((0, 0, 0), (0.2, 0.2, 0.2), (0, 0, 0))
    with pytest.raises(PhysicsError) as e:
        w.step(TICK)
    assert e.value.code == 16  # FluidFull
    assert len(w) == 10


def test_water_past_the_domain_is_taken_out_and_the_domain_kept():
    w = pool()
    low, high = (-1.0, -1.0, -1.0), (2.0, 0.3, 2.0)
    w.set_domain((low, high))
    before = len(w)
    r = w.step(TICK)
    assert r.lost > 0
    assert (w.positions()[:, 1] <= np.float32(0.3)).all()
    assert len(w) + r.lost <= before + r.emitted
    assert len(w) == before + r.emitted - r.drained - r.lost
    with pytest.raises(PhysicsError):
        w.set_domain((high, low))
    with pytest.raises(PhysicsError):
        w.set_domain((low, (math.nan, math.nan, math.nan)))
    w.set_domain(None)
    assert w.step(TICK).lost == 0


def test_a_zero_step_runs_no_flows_and_keeps_the_domain():
    # step returns before the flows and the domain when dt is 0
    # (src/fluid/step.rs:66-68)
    w = pool()
    w.set_domain(((-1.0, -1.0, -1.0), (2.0, 0.2, 2.0)))
    before = len(w)
    r = w.step(0.0)
    assert (r.emitted, r.drained, r.lost) == (0, 0, 0)
    assert len(w) == before
    r = w.step(TICK)
    assert r.lost > 0 and r.emitted > 0
    assert (w.positions()[:, 1] <= np.float32(0.2)).all()
