"""Parity of the gfx950 fluid step with the CPU oracle, through the C ABI.

Bars: the sort permutation and the neighbour lists bit-exact; per-particle
density and force within 1e-5 relative (denominator max(|a|, |b|)) after one
step; rollouts within 1e-6 m in position. The device path computes every pair
term and every neighbour sum in the reference's order with IEEE f32
arithmetic, so the tests also report (and for the neighbour build require)
exact equality. DFSPH is compared with the oracle in residual mode 1 (the
device's fixed-shape FP64 stop sum); mode 0 (the reference's f32 fold) is
compared on iteration counts in test_oracle_dfsph.py.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.fluid import FluidWorld
from solunar_amd.types import DFSPH, WCSPH, FluidSettings, PhysicsError

pytestmark = pytest.mark.gpu

REL = 1e-5           # density / force tolerance, relative
POS_TOL = 1e-6       # rollout position tolerance, metres


def pair(settings, residual_mode=1, threads=8):
    print(f"case: {settings}", flush=True)
    return FluidWorld(settings), OracleWorld(settings, residual_mode=residual_mode, threads=threads)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(a), np.abs(b))
    den[den == 0] = 1.0
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def same_lists(g, c):
    gs, gi = g
    cs, ci = c
    assert np.array_equal(gs, cs), "list starts differ"
    assert np.array_equal(gi, ci), "list items differ"


@pytest.mark.parametrize("count", [900, 3000])
def test_neighbour_build_of_shaken_lattices_is_bit_exact(count):
    gpu, cpu = pair(FluidSettings())
    scenes.shaken(gpu, count)
    scenes.shaken(cpu, count)
    assert np.array_equal(gpu.positions(), cpu.positions())
    gpu.update_neighbors()
    cpu.update_neighbors()
    assert np.array_equal(gpu.ids(), cpu.ids()), "sort permutation differs"
    assert np.array_equal(gpu.positions(), cpu.positions())
    same_lists(gpu.neighbors(), cpu.neighbors())
    gpu.compute_densities()
    cpu.compute_densities()
    assert rel_err(gpu.densities(), cpu.densities()) <= REL
    assert np.array_equal(gpu.densities(), cpu.densities())
    assert gpu.state_hash() == cpu.state_hash()


def test_neighbour_build_of_the_dam_break_with_walls_is_bit_exact():
    gpu, cpu = pair(scenes.settings_dam_small())
    for w in (gpu, cpu):
        scenes.dam_break(w)
    assert np.array_equal(gpu.wall_positions(), cpu.wall_positions())
    gpu.update_neighbors()
    cpu.update_neighbors()
    assert np.array_equal(gpu.ids(), cpu.ids())
    same_lists(gpu.neighbors(), cpu.neighbors())
    same_lists(gpu.wall_neighbors(), cpu.wall_neighbors())
    gpu.compute_densities()
    cpu.compute_densities()
    assert rel_err(gpu.densities(), cpu.densities()) <= REL
    assert np.array_equal(gpu.densities(), cpu.densities())


def test_one_wcsph_step_density_and_force_match():
    gpu, cpu = pair(scenes.settings_dam_small())
    for w in (gpu, cpu):
        scenes.dam_break(w)
    gpu.keep_forces(True)
    rg, rc = gpu.step(1e-4), cpu.step(1e-4)
    assert rg == rc, (rg, rc)
    assert np.array_equal(gpu.ids(), cpu.ids())
    assert rel_err(gpu.densities(), cpu.densities()) <= REL
    assert rel_err(gpu.forces(), cpu.forces()) <= REL
    assert np.array_equal(gpu.forces(), cpu.forces())
    assert np.array_equal(gpu.positions(), cpu.positions())
    assert np.array_equal(gpu.velocities(), cpu.velocities())


def test_a_large_wcsph_dam_break_steps_as_the_oracle_does():
    # ~300k particles: past 262,144 the force pass takes its two-phase walk
    s = replace(scenes.settings_dam_small(), spacing=0.0075)
    gpu, cpu = pair(s)
    for w in (gpu, cpu):
        n = scenes.dam_break(w)
    assert n > 262144
    gpu.keep_forces(True)
    cpu.keep_forces(True)
    for k in range(3):
        rg, rc = gpu.step(1e-4), cpu.step(1e-4)
        assert rg == rc, (k, rg, rc)
        assert np.array_equal(gpu.forces(), cpu.forces()), k
    assert np.array_equal(gpu.ids(), cpu.ids())
    assert np.array_equal(gpu.positions(), cpu.positions())
    assert np.array_equal(gpu.velocities(), cpu.velocities())
    assert np.array_equal(gpu.densities(), cpu.densities())


def test_wcsph_dam_break_100_steps_roll_out_identically():
    gpu, cpu = pair(scenes.settings_dam_small())
    for w in (gpu, cpu):
        scenes.dam_break(w)
    for k in range(100):
        rg, rc = gpu.step(1e-4), cpu.step(1e-4)
        assert rg == rc, (k, rg, rc)
    assert np.array_equal(gpu.ids(), cpu.ids())
    assert float(np.max(np.abs(gpu.positions() - cpu.positions()))) <= POS_TOL
    assert gpu.state_hash() == cpu.state_hash()


def test_wcsph_ticks_with_substeps_match():
    s = FluidSettings()
    gpu, cpu = pair(s)
    for w in (gpu, cpu):
        w.add_box_container((0, 0, 0), (0.9, 0.5, 0.2))
        w.fill_box((0, 0, 0), (0.3, 0.3, 0.2))
    for _ in range(6):
        rg, rc = gpu.step(1 / 60), cpu.step(1 / 60)
        assert rg == rc and rg.substeps > 1
    assert gpu.state_hash() == cpu.state_hash()


@pytest.mark.parametrize("ticks", [4])
def test_dfsph_ticks_match_with_the_fixed_shape_residual(ticks):
    s = FluidSettings(solver=DFSPH)
    gpu, cpu = pair(s, residual_mode=1)
    for w in (gpu, cpu):
        w.add_box_container((0, 0, 0), (0.9, 0.5, 0.2))
        w.fill_box_jittered((0, 0, 0), (0.3, 0.3, 0.2), 0.05, 11)
    gpu.keep_forces(True)
    for k in range(ticks):
        rg, rc = gpu.step(1 / 60), cpu.step(1 / 60)
        assert rg == rc, (k, rg, rc)
        assert np.array_equal(gpu.positions(), cpu.positions()), k
    assert rel_err(gpu.densities(), cpu.densities()) <= REL
    assert rel_err(gpu.forces(), cpu.forces()) <= REL
    assert gpu.state_hash() == cpu.state_hash()


def test_dfsph_dam_break_at_fine_spacing_matches():
    s = replace_settings(scenes.settings_dam_1m(), spacing=0.02)
    gpu, cpu = pair(s, residual_mode=1)
    for w in (gpu, cpu):
        scenes.dam_break(w)
    for k in range(5):
        rg, rc = gpu.step(5e-4), cpu.step(5e-4)
        assert rg == rc, (k, rg, rc)
    assert gpu.state_hash() == cpu.state_hash()


def replace_settings(s, **kw):
    from dataclasses import replace

    return replace(s, **kw)


def test_an_empty_world_steps():
    gpu = FluidWorld(FluidSettings())
    r = gpu.step(1 / 60)
    assert r.substeps == 0 and len(gpu) == 0
    gpu.update_neighbors()
    gpu.compute_densities()


def test_free_fall_without_walls_grows_the_cell_table():
    # two clumps far apart: the cell box is large and sparse, the table grows
    s = FluidSettings()
    gpu, cpu = pair(s)
    pts = np.concatenate([scenes.shaken_positions(200), scenes.shaken_positions(200) + np.float32(40.0)])
    for w in (gpu, cpu):
        w.add_particles(pts)
    for _ in range(3):
        rg, rc = gpu.step(1 / 60), cpu.step(1 / 60)
        assert rg == rc
    assert gpu.state_hash() == cpu.state_hash()


def test_particles_started_outside_the_container_are_held_back():
    s = FluidSettings()
    gpu, cpu = pair(s)
    for w in (gpu, cpu):
        w.add_box_container((0, 0, 0), (0.5, 0.5, 0.5))
        w.fill_box((0, 0, 0), (0.2, 0.2, 0.2))
        w.add_particles([[0.6, 0.2, 0.2], [-0.2, 0.1, 0.1]], [[1.0, 0, 0], [-1.0, 0, 0]])
    rg, rc = gpu.step(1 / 60), cpu.step(1 / 60)
    assert rg == rc and rg.held_back > 0
    assert gpu.state_hash() == cpu.state_hash()


def test_a_gate_taken_away_matches():
    s = FluidSettings()
    gpu, cpu = pair(s)
    gates = []
    for w in (gpu, cpu):
        w.add_box_container((0, 0, 0), (1.2, 0.6, 0.2))
        w.fill_box((0, 0, 0), (0.4, 0.4, 0.2))
        gates.append(w.add_wall((0.4, 0.0, 0.0), (0.5, 0.6, 0.2)))
    for _ in range(3):
        assert gpu.step(1 / 60) == cpu.step(1 / 60)
    assert gpu.remove_boundary(gates[0]) and cpu.remove_boundary(gates[1])
    assert not gpu.remove_boundary(gates[0])
    for _ in range(3):
        assert gpu.step(1 / 60) == cpu.step(1 / 60)
    assert gpu.state_hash() == cpu.state_hash()


def test_bad_input_is_refused_and_changes_nothing():
    gpu = FluidWorld(FluidSettings(max_particles=100))
    with pytest.raises(PhysicsError) as e:
        gpu.add_particle((math.nan, 0, 0))
    assert e.value.name == "NonFinite" and e.value.detail == "fluid particle"
    with pytest.raises(PhysicsError) as e:
        gpu.fill_box((0, 0, 0), (1, 1, 1))
    assert e.value.name == "FluidFull"
    assert gpu.is_empty()
    assert gpu.fill_box((0, 0, 0), (0.2, 0.2, 0.2)) == 64
    with pytest.raises(PhysicsError) as e:
        gpu.compute_densities()
    assert e.value.name == "OutOfRange"
    with pytest.raises(PhysicsError) as e:
        gpu.step(-1.0)
    assert e.value.name == "NonFinite"
    with pytest.raises(PhysicsError) as e:
        gpu.set_domain(((1, 1, 1), (0, 0, 0)))
    assert e.value.name == "OutOfRange"
    with pytest.raises(PhysicsError) as e:
        gpu.add_emitter((0, 0, 0), (1, 1, 1), (float("nan"), 0, 0))
    assert e.value.name == "NonFinite"
    for bad in (dict(spacing=0.0), dict(support=9.0), dict(max_particles=0), dict(courant=1.5)):
        with pytest.raises(PhysicsError) as e:
            FluidWorld(FluidSettings(**bad))
        assert e.value.name == "InvalidSettings"


def test_the_device_step_is_deterministic():
    hashes = []
    for _ in range(2):
        w = FluidWorld(scenes.settings_dam_small())
        scenes.dam_break(w)
        for _ in range(10):
            w.step(1e-4)
        hashes.append(w.state_hash())
    assert hashes[0] == hashes[1]


def test_a_capped_step_reports_it():
    gpu, cpu = pair(FluidSettings(max_substeps=2))
    for w in (gpu, cpu):
        w.fill_box((0, 0, 0), (0.1, 0.1, 0.1))
    rg, rc = gpu.step(1 / 60), cpu.step(1 / 60)
    assert rg.capped and rg.substeps == 2 and rg == rc
    assert gpu.state_hash() == cpu.state_hash()


@pytest.mark.parametrize("spacing, steps", [(0.01, 3), (0.005, 1)])
def test_dfsph_dam_break_at_scale_matches(spacing, steps):
    # 128,000 and 1,024,000 particles: hundreds of blocks meet at the solve's
    # grid barriers
    s = replace_settings(scenes.settings_dam_1m(), spacing=spacing)
    gpu, cpu = pair(s, residual_mode=1, threads=8)
    for w in (gpu, cpu):
        scenes.dam_break(w)
    for k in range(steps):
        rg, rc = gpu.step(5e-4), cpu.step(5e-4)
        print(f"step {k}: {rg}", flush=True)
        assert rg == rc, (k, rg, rc)
    assert np.array_equal(gpu.ids(), cpu.ids())
    assert np.array_equal(gpu.positions(), cpu.positions())
    assert np.array_equal(gpu.velocities(), cpu.velocities())
    assert np.array_equal(gpu.densities(), cpu.densities())


@pytest.mark.parametrize("solver", [WCSPH, DFSPH])
def test_sphere_obstacles_as_boundary_points_match(solver):
    # the 16M configuration's obstacles at a coarse spacing: 200 spheres added
    # through sfl_add_boundary_points, both solvers
    base = scenes.settings_dam_small() if solver == WCSPH else scenes.settings_dam_1m()
    s = replace_settings(base, spacing=0.02)
    gpu, cpu = pair(s, residual_mode=1)
    for w in (gpu, cpu):
        scenes.dam_with_obstacles(w, obstacles=200, seed=3)
    assert np.array_equal(gpu.wall_positions(), cpu.wall_positions())
    dt = 1e-4 if solver == WCSPH else 5e-4
    for k in range(10):
        rg, rc = gpu.step(dt), cpu.step(dt)
        assert rg == rc, (k, rg, rc)
    assert gpu.state_hash() == cpu.state_hash()


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_dfsph_matches_in_every_neighbour_format(fmt, monkeypatch):
    # 16-bit masks, 32-bit masks, lists, compact lists: one arithmetic, one
    # result
    monkeypatch.setenv("SOLUNAR_NEIGHBOUR_FORMAT", str(fmt))
    s = replace(scenes.settings_dam_1m(), spacing=0.02)
    gpu, cpu = pair(s, residual_mode=1)
    for w in (gpu, cpu):
        scenes.dam_break(w)
    for k in range(6):
        rg, rc = gpu.step(5e-4), cpu.step(5e-4)
        assert rg == rc, (fmt, k, rg, rc)
    assert gpu.state_hash() == cpu.state_hash()


@pytest.mark.parametrize("knobs", [{"SOLUNAR_SOLVE_TILES": "0"}, {"SOLUNAR_SOLVE_TILE_PARTICLES": "640"}],
                         ids=["global-gathers", "small-tiles"])
def test_dfsph_solve_matches_with_and_without_lds_tiles(knobs, monkeypatch):
    # the list solve stages each tile's neighbourhood in LDS; a tile that
    # does not fit gathers from global memory; without tiles every particle
    # does. A cap of 640 staged particles sends the dense tiles of this scene
    # down the global path and keeps the sparse ones in LDS.
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    s = replace(scenes.settings_dam_1m(), spacing=0.02)
    gpu, cpu = pair(s, residual_mode=1)
    for w in (gpu, cpu):
        scenes.dam_break(w)
    for k in range(6):
        rg, rc = gpu.step(5e-4), cpu.step(5e-4)
        assert rg == rc, (knobs, k, rg, rc)
    assert gpu.state_hash() == cpu.state_hash()


def test_crowded_cells_widen_the_neighbour_format():
    # particles at half the spacing put 64 in one cell: past 16- and 32-bit
    # masks, so the world falls back to explicit lists and still matches
    s = FluidSettings(solver=DFSPH, spacing=0.02)
    gpu, cpu = pair(s, residual_mode=1)
    g = np.arange(0.0, 0.04, 0.01, dtype=np.float32) + np.float32(0.321)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    for w in (gpu, cpu):
        w.add_box_container((0, 0, 0), (0.8, 0.8, 0.8))
        w.add_particles(pts)
    for k in range(2):
        rg, rc = gpu.step(1e-4), cpu.step(1e-4)
        assert rg == rc, (k, rg, rc)
    assert gpu.state_hash() == cpu.state_hash()


def test_compact_lists_widen_past_their_offsets():
    # 343 particles a cell in a column 12 cells tall: the column's staged
    # range holds 4,116 particles, past the compact format's 12-bit offsets,
    # so the world widens to full lists and still matches
    s = FluidSettings(solver=DFSPH, spacing=0.02)
    gpu, cpu = pair(s, residual_mode=1)
    step = np.float32(0.04) / np.float32(7)
    xy = np.arange(7, dtype=np.float32) * step + np.float32(0.3205)
    z = np.arange(84, dtype=np.float32) * step + np.float32(0.2005)
    pts = np.stack(np.meshgrid(xy, xy, z, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    for w in (gpu, cpu):
        w.add_box_container((0, 0, 0), (0.8, 0.8, 0.8))
        w.add_particles(pts)
    rg, rc = gpu.step(1e-4), cpu.step(1e-4)
    assert rg == rc, (rg, rc)
    assert gpu.state_hash() == cpu.state_hash()


def test_the_kernel_domain_and_one_particles_list_read_back_as_the_oracle_has_them():
    # FluidWorld::kernel with value / slope / gradient (mod.rs:140-164,
    # 376-378), domain (mod.rs:446-448), neighbors_of (mod.rs:612-614)
    s = scenes.settings_dam_small()
    gpu, cpu = pair(s)
    for d in np.linspace(0.0, 0.05, 101, dtype=np.float32):
        assert np.float32(gpu.kernel_value(float(d))) == np.float32(cpu.kernel_value(float(d)))
        assert np.float32(gpu.kernel_slope(float(d))) == np.float32(cpu.kernel_slope(float(d)))
    for off in [(0.0, 0.0, 0.0), (0.01, -0.02, 0.005), (0.03, 0.0, 0.0), (0.05, 0.0, 0.0)]:
        o = np.asarray(off, np.float32)
        d = np.sqrt(np.float32(o[0] * o[0] + o[1] * o[1]) + o[2] * o[2])
        want = o * (np.float32(cpu.kernel_slope(float(d))) / d) if d > 0 else np.zeros(3, np.float32)
        assert np.array_equal(np.asarray(gpu.kernel_gradient(off), np.float32), want.astype(np.float32))
    assert gpu.domain() is None
    box = ((0.0, -1.0, 0.0), (1.6, 2.0, 0.8))
    gpu.set_domain(box)
    got = gpu.domain()
    assert [tuple(np.float32(c) for c in v) for v in got] == [tuple(np.float32(c) for c in v) for v in box]
    gpu.set_domain(None)
    assert gpu.domain() is None
    for w in (gpu, cpu):
        scenes.dam_break(w)
        w.update_neighbors()
    for i in (0, 1, 777, len(gpu) - 1):
        assert np.array_equal(gpu.neighbors_of(i), cpu.neighbors_of(i)), i
    with pytest.raises(PhysicsError):
        gpu.neighbors_of(len(gpu))
