"""The flow solve (the tiled Jacobi passes waiting tile by tile on their
neighbours, SOLUNAR_SOLVE_FLOW, on by default) against the same solve with
its passes meeting at grid barriers (SOLUNAR_SOLVE_FLOW=0) and against the
CPU oracle in the device's stop-sum mode, bit for bit.

The cases are the ones its bookkeeping has to get right:
- every iteration cap from 0 to 7 (the least number of iterations at most
  2, as settings require), so solves stop on odd and even
  iterations (the velocities of an odd stop live in the second buffer and
  are copied back) and at the cap (a forced stop);
- a world of a few tiles, so most XCD groups have no tiles at all;
- a tolerance no solve reaches before the cap, and one every solve meets at
  its first allowed test;
- the stop tests' averages and counts (the solve log) equal to the barrier
  solve's, so the decisions were taken on the same sums in the same order.
"""
from dataclasses import replace

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, FluidSettings

pytestmark = pytest.mark.gpu


def world(settings, flow, monkeypatch):
    monkeypatch.setenv("SOLUNAR_SOLVE_FLOW", "1" if flow else "0")
    return FluidWorld(settings)


def assert_same(a, b, what):
    assert np.array_equal(a.ids(), b.ids()), what
    assert np.array_equal(a.positions(), b.positions()), what
    assert np.array_equal(a.velocities(), b.velocities()), what
    assert np.array_equal(a.densities(), b.densities()), what


def run_three(settings, build, steps, dt, monkeypatch):
    flow, barrier = world(settings, True, monkeypatch), world(settings, False, monkeypatch)
    cpu = OracleWorld(settings, residual_mode=1, threads=8)
    for w in (flow, barrier, cpu):
        build(w)
    for k in range(steps):
        rf, rb, rc = flow.step(dt), barrier.step(dt), cpu.step(dt)
        assert rf == rb == rc, (k, rf, rb, rc)
        af, lf = flow.solve_log()
        ab, lb = barrier.solve_log()
        assert np.array_equal(af, ab) and np.array_equal(lf, lb), k
        assert_same(flow, barrier, f"flow and barrier solves differ at step {k}")
        assert_same(flow, cpu, f"flow solve and oracle differ at step {k}")
    # the flow world holds the per-tile pass counters and tile list the
    # barrier world does without: the flow solve is the one that ran
    assert flow.memory_bytes() > barrier.memory_bytes()
    return flow


@pytest.mark.parametrize("cap", list(range(8)))
def test_every_iteration_cap_stops_as_the_barrier_solve_and_the_oracle(cap, monkeypatch):
    # settings ask for max_iterations >= min_iterations
    s = replace(scenes.settings_dam_1m(), spacing=0.02, min_iterations=min(2, cap), max_iterations=cap)
    print(f"case: dam break at 0.02, iterations {s.min_iterations}..{cap}", flush=True)
    flow = run_three(s, scenes.dam_break, 4, 5e-4, monkeypatch)
    r = flow.step(5e-4)
    assert r.iterations_density <= cap * r.substeps and r.iterations_divergence <= cap * r.substeps


def test_a_world_of_a_few_tiles_leaves_most_groups_empty(monkeypatch):
    # 2 x 2 x 2 cells of water: one or two solve tiles for eight groups
    s = FluidSettings(solver=DFSPH, spacing=0.02)
    print("case: a cube of 64 particles in a small container", flush=True)

    def build(w):
        w.add_box_container((0.0, 0.0, 0.0), (0.2, 0.2, 0.2))
        w.fill_box_jittered((0.0, 0.0, 0.0), (0.08, 0.08, 0.08), 0.05, 3)

    run_three(s, build, 6, 5e-4, monkeypatch)


@pytest.mark.parametrize("tolerances", [(1e-9, 1e-9), (1.0, 1.0)], ids=["never-met", "met-at-once"])
def test_solves_that_never_settle_or_settle_at_once(tolerances, monkeypatch):
    s = replace(scenes.settings_dam_1m(), spacing=0.02, max_iterations=12,
                divergence_tolerance=tolerances[0], density_tolerance=tolerances[1])
    print(f"case: dam break at 0.02, tolerances {tolerances}", flush=True)
    flow = run_three(s, scenes.dam_break, 3, 5e-4, monkeypatch)
    r = flow.step(5e-4)
    if tolerances[0] < 1e-6:
        assert r.iterations_density == 12 * r.substeps
    else:
        assert r.iterations_density == s.min_iterations * r.substeps
