"""The allocation-free step (sfl_step_no_alloc) against sfl_step, through the
C ABI.

sfl_step_no_alloc has no counterpart in the reference; it is sfl_step with
the control block and the sort's cell table in a scratch array that every
such call on the GPU shares. The bar is sfl_step's own: reports, ids,
positions, velocities, densities and the state hash equal bit for bit, here
on calls in a row over worlds of different sizes (each call finds the last
one's scratch contents), and the refusals and limits it documents.
"""
from dataclasses import replace

import numpy as np
import pytest

from solunar_amd import scenes
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, OUT_OF_RANGE, UNSUPPORTED, PhysicsError

pytestmark = pytest.mark.gpu


def twins(settings, build):
    print(f"case: {settings}", flush=True)
    a, b = FluidWorld(settings), FluidWorld(settings)
    build(a)
    build(b)
    return a, b


def assert_same(a, b):
    assert np.array_equal(a.ids(), b.ids())
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    assert np.array_equal(a.densities(), b.densities())
    assert a.state_hash() == b.state_hash()


def small_dam(w):
    scenes.dam_break(w, block=(0.2, 0.2, 0.2), box=(0.6, 0.5, 0.4))


def test_calls_in_a_row_equal_the_step_and_a_smaller_call_follows_a_larger_one():
    big = twins(scenes.settings_dam_small(), scenes.dam_break)  # 16,000 + 54,464 walls
    small = twins(scenes.settings_dam_small(), small_dam)
    assert len(small[0]) < len(big[0])
    for k, dt in enumerate((1e-4, 1e-4, 4e-3, 1e-4)):  # the third has several substeps
        for ref, cut in (big, small):
            rr, rc = ref.step(dt), cut.step_no_alloc(dt)
            assert rr == rc, (k, rr, rc)
            assert_same(ref, cut)
    assert big[0].step(4e-3).substeps > 1


def test_the_last_grid_is_kept_when_the_world_holds_a_table():
    ref, cut = twins(scenes.settings_dam_small(), scenes.dam_break)
    ref.step(1e-4)
    cut.step(1e-4)  # the world's own cell table now exists
    ref.step(1e-4)
    cut.step_no_alloc(1e-4)
    for w in (ref, cut):
        w.compute_densities()
    assert_same(ref, cut)
    for r, c in ((ref.neighbors(), cut.neighbors()), (ref.wall_neighbors(), cut.wall_neighbors())):
        assert np.array_equal(r[0], c[0]) and np.array_equal(r[1], c[1])


def test_a_world_without_its_own_table_drops_the_lists():
    ref, cut = twins(scenes.settings_dam_small(), small_dam)
    ref.step(1e-4)
    cut.step_no_alloc(1e-4)
    with pytest.raises(PhysicsError) as e:
        cut.compute_densities()
    assert e.value.code == OUT_OF_RANGE
    for w in (ref, cut):
        w.update_neighbors()
        w.compute_densities()
    assert_same(ref, cut)


def test_refusals_leave_the_world_as_it_was():
    w = FluidWorld(replace(scenes.settings_dam_small(), solver=DFSPH))
    small_dam(w)
    before = w.state_hash()
    with pytest.raises(PhysicsError) as e:
        w.step_no_alloc(1e-4)
    assert e.value.code == UNSUPPORTED
    assert w.state_hash() == before
    w = FluidWorld(scenes.settings_dam_small())
    small_dam(w)
    w.add_drain((0.5, 0.0, 0.0), (0.6, 0.1, 0.1))
    before = w.state_hash()
    with pytest.raises(PhysicsError) as e:
        w.step_no_alloc(1e-4)
    assert e.value.code == UNSUPPORTED
    assert w.state_hash() == before


def test_a_sort_past_the_scratch_is_refused_and_the_state_kept():
    # two particles far apart at radius 0.04: a box of 1,001 x 1,001 x 6
    # cells, more than the 32 MiB scratch holds (about 4.19M); sfl_step
    # grows its own table instead
    ref, cut = twins(scenes.settings_dam_small(),
                     lambda w: w.add_particles(np.array([[0.0, 0.0, 0.0], [40.0, 40.0, 0.2]], np.float32)))
    before = cut.state_hash()
    with pytest.raises(PhysicsError) as e:
        cut.step_no_alloc(1e-4)
    assert e.value.code == OUT_OF_RANGE
    assert cut.state_hash() == before
    ref.step(1e-4)
    # and the next call on a world that fits is unaffected
    a, b = twins(scenes.settings_dam_small(), small_dam)
    assert a.step(1e-4) == b.step_no_alloc(1e-4)
    assert_same(a, b)
