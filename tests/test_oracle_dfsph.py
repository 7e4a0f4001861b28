"""DFSPH and precision checks of the CPU oracle.

The reference has no known answer for DFSPH (src/fluid/mod.rs:49-52 marks it
"Not ready"), so its restatement (src/fluid/dfsph.rs:138-275) is pinned from
outside: closed forms and conservation laws for the density rate, factors
and push_apart, hand-worked FP64 pair values, the free-surface rule and the
clamp of the divergence solve, and planted mistakes that each fail one of
those checks; then thread-count invariance (test_oracle_reference.py),
agreement of the f32 step with the same algorithm in FP64, the solves ending
under their tolerance, and agreement of the reference's f32 fold of the
errors (src/fluid/dfsph.rs:263-267) with the device's fixed-shape FP64 sum.

Errors stated here (measured on these scenes, asserted with margin):
  WCSPH dam_small, 100 steps of 1e-4 s: f32 vs FP64 position 1.0e-6 m,
      velocity 1.0e-4 m/s, density 3.1e-5 relative.
  DFSPH dam break at spacing 0.02, 40 steps of 5e-4 s: position 6.6e-7 m,
      velocity 3.9e-5 m/s, density 1.8e-5 relative; the two stop sums take
      the same iterations and their averages differ by at most 4.1e-6
      relative.
"""
from dataclasses import replace

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.types import DFSPH, WCSPH, FluidSettings

POS_TOL = 1e-5     # m
VEL_TOL = 1e-3     # m/s
RHO_TOL = 1e-4     # relative


def run(settings, precision, mode, steps, dt, log=False):
    w = OracleWorld(settings, precision=precision, threads=8, residual_mode=mode)
    scenes.dam_break(w)
    w.log_solves(log)
    reports = [w.step(dt) for _ in range(steps)]
    order = np.argsort(w.ids())
    return w, reports, w.positions()[order], w.velocities()[order], w.densities()[order]


@pytest.mark.parametrize("solver, steps, dt", [(WCSPH, 100, 1e-4), (DFSPH, 40, 5e-4)])
def test_the_f32_step_stays_close_to_fp64(solver, steps, dt):
    s = scenes.settings_dam_small() if solver == WCSPH else replace(scenes.settings_dam_1m(), spacing=0.02)
    a = run(s, 32, 0, steps, dt)
    b = run(s, 64, 0, steps, dt)
    assert [r.substeps for r in a[1]] == [r.substeps for r in b[1]]
    assert np.abs(a[2] - b[2]).max() < POS_TOL
    assert np.abs(a[3] - b[3]).max() < VEL_TOL
    assert (np.abs(a[4] - b[4]) / b[4]).max() < RHO_TOL


def test_the_two_stop_sums_agree_on_the_dam_break():
    s = replace(scenes.settings_dam_1m(), spacing=0.02)
    ref = run(s, 32, 0, 40, 5e-4, log=True)
    dev = run(s, 32, 1, 40, 5e-4)
    assert [r.iterations for r in ref[1]] == [r.iterations for r in dev[1]]
    assert np.array_equal(ref[2], dev[2])
    log = ref[0].solve_log()
    rel = np.abs(log["average_f32"] - log["average_f64"]) / np.maximum(log["average_f64"], 1e-30)
    assert rel.max() < 1e-4


def test_every_solve_ends_under_its_tolerance_or_at_the_cap():
    s = replace(scenes.settings_dam_1m(), spacing=0.02)
    w, reports, *_ = run(s, 32, 0, 20, 5e-4, log=True)
    log = w.solve_log()
    ends = [k for k in range(len(log["solve"])) if k + 1 == len(log["solve"]) or log["iteration"][k + 1] == 0]
    assert ends
    for k in ends:
        tol = s.divergence_tolerance if log["solve"][k] == 0 else s.density_tolerance
        it = log["iteration"][k]
        assert log["average_f32"][k] <= np.float32(tol) or it >= s.max_iterations
        assert it >= s.min_iterations


def test_the_density_solve_squeezes_less_than_wcsph():
    # same scene, same time: DFSPH holds the density to its 0.01 % tolerance
    # on average, the weakly compressible solver squeezes about (v / c)^2
    s = replace(scenes.settings_dam_1m(), spacing=0.02)
    w, reports, _, _, rho = run(s, 32, 0, 20, 5e-4)
    over = np.maximum(rho / 1000.0 - 1.0, 0.0)
    assert over[over > 0].mean() < 1e-3


# ---- pins of the DFSPH restatement (src/fluid/dfsph.rs:138-275) ----
#
# The probe gives one substep's quantities at the current state: factors,
# density rates, both solves' first stiffness and error, and push_apart's
# velocity changes for a given stiffness. The checks below are closed forms
# and conservation laws, each chosen so that a plausible slip fails one:
#   rigid motion           rate exactly 0 (translation) or 0 to FP64
#                          rounding (rotation); a dilation v = a (x - c)
#                          gives the continuity equation's -3 a rho inside
#   momentum               push_apart's pair terms are antisymmetric, so the
#                          fluid's momentum and angular momentum do not change
#   hand-worked pairs      alpha, rate, stiffness and changes for two and three
#                          particles and a particle by one wall point, from
#                          the Wendland closed form in FP64
#   free surface, clamp    a divergence term is 0 below 20 neighbours (walls
#                          counted) and for a negative rate
# Planted mistakes (oracle/fluid_oracle.hpp, planted_mistake 4-8) each fail
# one of them.

DX = 0.02


def probe_world(positions, velocities, walls=None, precision=64):
    w = OracleWorld(FluidSettings(solver=DFSPH, spacing=DX), precision=precision)
    if walls is not None:
        w.add_boundary_points(walls)
    if precision == 64:
        w.add_particles_f64(positions, velocities)  # exact in the FP64 world
    else:
        w.add_particles(positions, velocities)
    w.update_neighbors()
    return w


def lattice(n, jitter=0.05, seed=1):
    g = (np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * DX
    return g + np.random.default_rng(seed).uniform(-jitter, jitter, g.shape) * DX


def wendland(d):
    """Gradient size W'(d) of the Wendland C2 kernel at the f32 radius, FP64."""
    r = float(np.float32(DX) * np.float32(2.0))
    q = d / r
    return 21.0 / (2.0 * np.pi * r ** 3) * -20.0 * q * (1.0 - q) ** 3 / r if q < 1 else 0.0


def grad(off):
    d = float(np.linalg.norm(off))
    return np.asarray(off, float) * (wendland(d) / d)


def by_id(w, values):
    return np.asarray(values)[np.argsort(w.ids())]


def test_the_density_rate_vanishes_under_rigid_motion():
    x = lattice(10)
    c = x.mean(0)
    # translation: every (v_i - v_j) is exactly zero, in f32 too
    for precision in (32, 64):
        w = probe_world(x, np.tile([0.3, -0.2, 0.1], (len(x), 1)), precision=precision)
        assert np.all(w.dfsph_probe(1e-3)["rate"] == 0.0)
    # rotation: (w x off) . off = 0 for every pair
    omega = np.array([0.4, -1.1, 0.7])
    w = probe_world(x, np.cross(omega, x - c))
    spin = w.dfsph_probe(1e-3)["rate"]
    # dilation of the same speeds as the yardstick
    w = probe_world(x, 1.3 * (x - c))
    grow = w.dfsph_probe(1e-3)["rate"]
    assert np.abs(spin).max() < 1e-12 * np.abs(grow).max()


def test_a_dilation_thins_the_water_as_the_continuity_equation_says():
    n, a = 12, 0.3
    x = lattice(n)
    c = x.mean(0)
    for sign in (1.0, -1.0):
        w = probe_world(x, sign * a * (x - c))
        rate, rho, p = w.dfsph_probe(1e-3)["rate"], w.densities(), w.positions()
        inside = np.all((p > 4 * DX) & (p < (n - 4) * DX), axis=1)
        assert inside.sum() == 64
        # D rho / Dt = -rho div v = -3 a rho for v = a (x - c); the lattice
        # sum is within 10 % of the integral at support 2
        ratio = rate[inside] / (-3.0 * sign * a * rho[inside])
        assert 0.9 < ratio.min() and ratio.max() < 1.0


def test_push_apart_keeps_momentum_and_angular_momentum():
    x = lattice(9, jitter=0.2, seed=4)
    rng = np.random.default_rng(7)
    w = probe_world(x, rng.normal(0, 0.1, x.shape))
    kappa = rng.lognormal(0.0, 1.0, len(x)) * 1e3
    ch = w.dfsph_probe(1e-3, kappa)["change"]
    p = w.positions()
    assert np.abs(ch.sum(0)).max() < 1e-12 * np.abs(ch).sum()
    assert np.abs(np.cross(p, ch).sum(0)).max() < 1e-12 * (np.linalg.norm(p, axis=1) * np.linalg.norm(ch, axis=1)).sum()
    assert np.abs(ch).max() > 0


def test_a_hand_worked_pair_and_triangle():
    dt = 2.0 ** -9  # exact in f32, as the probe takes it
    r = 2 * DX
    # a pair at 0.7 r along a slanted direction, closing on each other
    e = np.array([2.0, -1.0, 0.5]) / np.linalg.norm([2.0, -1.0, 0.5])
    x = np.array([[0.3, 0.3, 0.3], [0.3, 0.3, 0.3] + 0.7 * r * e])
    v = np.array([[0.2, 0.1, -0.3], [-0.4, 0.0, 0.1]])
    kappa = np.array([300.0, 700.0])
    w = probe_world(x, v)
    m = w.particle_mass()
    order = np.argsort(w.ids())  # storage -> id order
    ks = np.empty(2)
    ks[order] = kappa  # stiffness in storage order
    pr = w.dfsph_probe(dt, ks)
    g = grad(x[0] - x[1])  # gradient at particle 0, toward particle 1
    alpha = 1.0 / (2.0 * m * m * (g @ g))
    rate0 = m * ((v[0] - v[1]) @ g)
    assert rate0 > 0  # closing in
    np.testing.assert_allclose(pr["alpha"][order], [alpha, alpha], rtol=1e-12)
    np.testing.assert_allclose(pr["rate"][order], [rate0, rate0], rtol=1e-12)
    dv0 = -dt * m * g * (kappa[0] + kappa[1])
    np.testing.assert_allclose(pr["change"][order], [dv0, -dv0], rtol=1e-12)
    # two particles are a free surface and far below rest density: no load
    for k in ("div_stiffness", "div_error", "den_stiffness", "den_error"):
        assert np.all(pr[k] == 0.0)

    # a triangle with sides 0.5 r, 0.6 r, 0.8 r
    x = np.array([[0.3, 0.3, 0.3], [0.3 + 0.5 * r, 0.3, 0.3], [0.3, 0.3 + 0.6 * r, 0.3]])
    w = probe_world(x, np.zeros((3, 3)))
    pr = w.dfsph_probe(dt)
    order = np.argsort(w.ids())
    for i in range(3):
        gs = [grad(x[i] - x[j]) for j in range(3) if j != i]
        total = m * m * (np.sum(gs, axis=0) @ np.sum(gs, axis=0) + sum(gg @ gg for gg in gs))
        assert pr["alpha"][order][i] == pytest.approx(1.0 / total, rel=1e-12)


def test_a_particle_by_a_wall_point():
    dt = 2.0 ** -9  # exact in f32, as the probe takes it
    r = 2 * DX
    wall = np.array([[0.5, 0.5, 0.5]])
    x = np.array([[0.5 + 0.4 * r, 0.5 + 0.3 * r, 0.5]])
    v = np.array([[-0.3, 0.1, 0.2]])
    w = probe_world(x, v, walls=wall)
    m = w.particle_mass()
    pr = w.dfsph_probe(dt, np.array([500.0]))
    g = grad(x[0] - wall[0])
    # walls add to the summed gradient, not to the squares
    assert pr["alpha"][0] == pytest.approx(1.0 / (m * m * (g @ g)), rel=1e-12)
    rate = m * (v[0] @ g)
    assert rate > 0  # moving into the wall
    assert pr["rate"][0] == pytest.approx(rate, rel=1e-12)
    np.testing.assert_allclose(pr["change"][0], -dt * m * 500.0 * g, rtol=1e-12)


def sphere_points(k, radius):
    """k points spread over a sphere (Fibonacci lattice)."""
    i = np.arange(k) + 0.5
    phi = np.arccos(1 - 2 * i / k)
    theta = np.pi * (1 + 5 ** 0.5) * i
    return radius * np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], 1)


def centre_terms(k, sign, walls=0):
    """A particle at rest with k others at half the radius around it, moving
    in (sign +1) or out (-1), and `walls` wall points beside it; returns the
    centre's divergence stiffness, error and rate."""
    c = np.array([0.5, 0.5, 0.5])
    x = np.vstack([c, c + sphere_points(k, DX)])
    v = np.vstack([np.zeros(3), -sign * 0.5 * sphere_points(k, 1.0)])
    wp = c + sphere_points(walls, 1.5 * DX) if walls else None
    w = probe_world(x, v, walls=wp)
    centre = int(np.argmin(np.linalg.norm(w.positions() - c, axis=1)))
    pr = w.dfsph_probe(1e-3)
    return pr["div_stiffness"][centre], pr["div_error"][centre], pr["rate"][centre]


def test_the_divergence_solve_leaves_the_free_surface_alone():
    # crowd = itself + k fluid + walls; SURFACE_NEIGHBORS = 20 (dfsph.rs:16)
    k19, e19, rate19 = centre_terms(18, +1)
    assert rate19 > 0 and k19 == 0.0 and e19 == 0.0
    k20, e20, rate20 = centre_terms(19, +1)
    assert rate20 > 0 and k20 > 0.0 and e20 > 0.0
    kw, ew, _ = centre_terms(18, +1, walls=1)  # a wall neighbour counts
    assert kw > 0.0 and ew > 0.0


def test_the_divergence_solve_does_not_pull_water_together():
    k, e, rate = centre_terms(30, -1)
    assert rate < 0 and k == 0.0 and e == 0.0


def test_a_density_solve_ends_under_its_tolerance_not_at_the_cap():
    s = replace(scenes.settings_dam_1m(), spacing=0.02)
    w, reports, *_ = run(s, 32, 0, 20, 5e-4, log=True)
    log = w.solve_log()
    ends = [k for k in range(len(log["solve"])) if k + 1 == len(log["solve"]) or log["iteration"][k + 1] == 0]
    for k in ends:
        assert log["iteration"][k] < s.max_iterations
        starts = k - log["iteration"][k]
        if log["iteration"][k] > 2:  # the solve brought its error down
            assert log["average_f32"][k] < log["average_f32"][starts]


@pytest.mark.parametrize("mistake, check", [
    (4, test_push_apart_keeps_momentum_and_angular_momentum),
    (5, test_a_hand_worked_pair_and_triangle),
    (6, test_the_divergence_solve_does_not_pull_water_together),
    (7, test_the_divergence_solve_leaves_the_free_surface_alone),
    (8, test_a_particle_by_a_wall_point),
])
def test_a_planted_dfsph_mistake_fails_its_check(mistake, check):
    from oracle.oracle import plant_mistake

    plant_mistake(mistake)
    try:
        with pytest.raises(AssertionError):
            check()
    finally:
        plant_mistake(0)
    check()
