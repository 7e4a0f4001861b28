"""Long runs, skipped unless SOLUNAR_LONG_TESTS=1 (minutes of CPU oracle
time each; every step prints a line).

- dam_1m's whole configured rollout: 200 steps of 5e-4 s (1,024,000 fluid
  + 832,064 wall particles, DFSPH, divergence tolerance 1e-3, density
  tolerance 1e-4), which include the collapse's steps of 60-70 density
  iterations, bit-exact with the CPU oracle in the device's stop-sum mode at
  every step: the reports, every stop test's average and count, and the
  ids, positions, velocities and densities.
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleWorld
from solunar_amd import scenes
from solunar_amd.fluid import FluidWorld

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(os.environ.get("SOLUNAR_LONG_TESTS") != "1", reason="long run: set SOLUNAR_LONG_TESTS=1"),
]

THREADS = max(1, min(16, os.cpu_count() or 1))


def test_dam_1m_whole_rollout_bit_exact():
    s = scenes.settings_dam_1m()
    print(f"case: dam_1m x200, {THREADS} oracle threads", flush=True)
    gpu, cpu = FluidWorld(s), OracleWorld(s, residual_mode=1, threads=THREADS)
    for w in (gpu, cpu):
        scenes.dam_break(w)
    for k in range(200):
        cpu.log_solves(True)  # clears the oracle's log
        rg, rc = gpu.step(5e-4), cpu.step(5e-4)
        print(f"step {k}: {rg.substeps} substeps, {rg.iterations_divergence}+{rg.iterations_density} iterations",
              flush=True)
        assert rg == rc, (k, rg, rc)
        avg, loaded = gpu.solve_log()
        log = cpu.solve_log()
        want = np.where(log["solve"] == 0, log["average_f64"], -log["average_f64"])
        assert np.array_equal(avg, want), k
        assert np.array_equal(loaded, log["loaded"]), k
        assert np.array_equal(gpu.ids(), cpu.ids()), k
        assert np.array_equal(gpu.positions(), cpu.positions()), k
        assert np.array_equal(gpu.velocities(), cpu.velocities()), k
        assert np.array_equal(gpu.densities(), cpu.densities()), k
    assert gpu.state_hash() == cpu.state_hash()
