"""Throughput of the gfx950 fluid step on the 1M-particle DFSPH dam break.

Workload (solunar_amd/scenes.py, dam_1m): DFSPH, spacing 0.005 m, Wendland C2
with support 2, 0.4 x 0.8 x 0.4 m water block (1,024,000 particles, lattice
plus seeded +-0.05 spacing jitter) in a 1.6 x 1.6 x 0.8 m container
(832,064 wall particles), divergence solve to 0.1 %, density solve to 0.01 %,
at most 100 Jacobi iterations each, `step(5e-4)`. A step is one
FluidWorld::step over all particles, state resident in HBM.

Printed: one JSON line with particle-steps/s, the HBM roofline of the
dominant kernel (algorithmic bytes from HIP events on the world's stream over
the timed region) and the CPU oracle timed on this host on a bounded sample.
After the timed steps (outside `value`), the same line carries the whole
200-step rollout of a fresh dam_1m world (`rollout_200`) and the 64k WCSPH
adjoint's reverse:forward time ratio (`adjoint_64k`); --no-extras skips
them.

With --workload batch_256: 256 WCSPH dam breaks (spacing 0.02 m, block sides
uniform in [0.2, 0.8] m, 2,000-64,000 particles each, 3.86M in all) stepped
as one batch, `steps` steps of 1e-4 s in one graph launch, next to the same
256 worlds stepped one by one on the single-world path.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu] [--no-extras]
                       [--workload dam_1m|dam_16m|batch_256|adjoint_64k|flows_1m|snapshot_1m]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DT = 5e-4

# Algorithmic HBM bytes per fluid particle (FP32 state; neighbour re-reads
# are served from cache and not counted):
#   rate pass  : x 12 + v 12 + alpha 4 + rho 4 read, kappa 4 written  = 36
#   push pass  : x 12 + kappa 4 read, v 12 read + 12 written          = 40
#   substep    : neighbour build 108, density/factors 24, viscosity
#                40 + 12 ... totalled as 292 outside the solve iterations
RATE_BYTES = 36
PUSH_BYTES = 40
SUBSTEP_FIXED_BYTES = 292
VISCOUS_BYTES = 40  # viscosity and gravity: x, v, rho 28 read, v 12 written


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(steps: int, warmup: int):
    """The CPU oracle (a C++ restatement of the reference step, with the
    reference's f32 stop sum) on the same workload: `warmup` untimed steps
    from the initial state, as the device run, then `steps` timed steps, on
    all host threads (at most 16)."""
    from oracle.oracle import OracleWorld
    from solunar_amd import scenes

    threads = min(16, os.cpu_count() or 1)
    w = OracleWorld(scenes.settings_dam_1m(), threads=threads)
    n = scenes.dam_break(w)
    for _ in range(warmup):
        w.step(DT)
    t0 = time.perf_counter()
    substeps = 0
    for _ in range(steps):
        substeps += w.step(DT).substeps
    dt = time.perf_counter() - t0
    return {
        "value": n * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"steps {warmup}..{warmup + steps - 1} of dam_1m ({n} particles, {substeps} substeps), "
                  f"{dt:.1f} s wall on {threads} host threads",
    }


# WCSPH force pass, algorithmic bytes per fluid particle and launch (float4
# state): x 16, v (with rho) 16 and p/rho^2 4 read, x 16 and v 16 written;
# with terrain the slope (float4) 16 more read. Neighbour re-reads are not
# counted: the cell order keeps them in L2.
FORCE_BYTES = 68
FORCE_SLOPE_BYTES = 16


def force_roofline(phases, particles: int, per_particle: int):
    """The WCSPH force pass against HBM: algorithmic bytes per launch over
    its average launch time (HIP events round each launch on the world's
    stream)."""
    ms, launches = phases.get("force", (0.0, 0))
    if not launches or ms <= 0:
        return None
    seconds = ms * 1e-3 / launches
    achieved = particles * per_particle / seconds / 1e9
    return {"kernel": "k_wcsph_force", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "launch_us": seconds * 1e6, "bytes_per_particle": per_particle}


def cpu_sample(name: str, build, step, steps: int):
    """The CPU oracle timed on `steps` steps of a workload from its initial
    state: `build(world)` makes the scene on an OracleWorld and returns its
    particle count, `step(world, k)` takes step k and returns its report."""
    from oracle.oracle import OracleWorld

    threads = min(16, os.cpu_count() or 1)
    w, n = build(lambda settings: OracleWorld(settings, threads=threads))
    t0 = time.perf_counter()
    substeps = 0
    for k in range(steps):
        substeps += step(w, k).substeps
    dt = time.perf_counter() - t0
    return {
        "value": n * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"steps 0..{steps - 1} of {name} ({n} particles, {substeps} substeps), "
                  f"{dt:.1f} s wall on {threads} host threads",
    }


# Batched WCSPH: algorithmic bytes per fluid particle and substep of each
# kernel group (FP32 state in float4 pairs (x, id) and (v, rho))
#   sort    : bounds 16, count 16 + 8, scatter 8 + 16 + 8, gather 8 + 32 + 32 = 144
#   density : x 16 read, rho 4 + p/rho^2 4 written                            = 24
#   force   : x 16 + v 16 + p/rho^2 4 read, x 16 + v 16 written                = 68
BATCH_BYTES = {"sort": 144, "density": 24, "force": 68}
BATCH_DT = 1e-4


def run_batch(args) -> dict:
    from solunar_amd import scenes
    from solunar_amd.fluid import FluidBatch, FluidWorld

    worlds = scenes.batch_worlds(FluidWorld, 256)
    sizes = [len(w) for w in worlds]
    n = sum(sizes)
    log(f"batch_256: {len(worlds)} scenes, {n} fluid particles ({min(sizes)}..{max(sizes)} per scene)")
    batch = FluidBatch(worlds)
    tuning = batch.autotune(BATCH_DT, repeats=3)
    log(f"tuned: sort {tuning['sort_block']}, density {tuning['density_block']}, force {tuning['force_block']}")
    if args.warmup:
        batch.step(BATCH_DT, args.warmup)
    batch.prepare(args.steps)
    batch.synchronize()
    log(f"timing {args.steps} batched steps ({batch.graph_nodes()} graph nodes)")
    t0 = time.perf_counter()
    reports = batch.step(BATCH_DT, args.steps)
    elapsed = time.perf_counter() - t0
    substeps = sum(r.substeps for r in reports)
    nodes = batch.graph_nodes()
    memory = batch.memory_bytes()
    tuning_after = batch.autotune(BATCH_DT, repeats=3)
    batch.close()
    value = n * args.steps / elapsed
    log(f"batched: {elapsed:.3f} s, {value:.3e} particle-steps/s")
    # the same scenes on the single-world path, one world after another
    per_steps = max(1, min(args.steps, args.per_scene_steps))
    for w in worlds:
        for _ in range(args.warmup):
            w.step(BATCH_DT)
        w.synchronize()
    log(f"timing {per_steps} steps of each of the {len(worlds)} worlds alone")
    t0 = time.perf_counter()
    for k, w in enumerate(worlds):
        for _ in range(per_steps):
            w.step(BATCH_DT)
        if k % 64 == 63:
            log(f"world {k + 1}/{len(worlds)}")
    for w in worlds:
        w.synchronize()
    alone = time.perf_counter() - t0
    alone_value = n * per_steps / alone
    for w in worlds:
        w.close()
    group_ms = {g: tuning_after[f"{g}_ms"][tuning_after["blocks"].index(tuning_after[f"{g}_block"])]
                for g in BATCH_BYTES}
    dom = max(group_ms, key=group_ms.get)
    dom_bytes = BATCH_BYTES[dom] * n
    achieved = dom_bytes / (group_ms[dom] * 1e-3) / 1e9
    return {
        "metric": "particle-steps/s, 256 batched WCSPH scenes",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch_256: 256 WCSPH dam breaks, spacing 0.02 m, blocks U[0.2, 0.8] m per axis "
                               "with 2,000-64,000 particles, 1.6 x 1.6 x 0.8 m containers, step 1e-4 s",
                   "fluid_particles": n, "scenes": len(worlds), "substeps": substeps, "graph_nodes": nodes,
                   "device_memory_bytes": memory, "tuning": tuning},
        "per_scene_launches": {"value": alone_value, "unit": "particle-steps/s", "steps": per_steps,
                               "seconds": alone, "speedup": value / alone_value},
        "roofline": {"bound": "hbm", "kernel": f"batch {dom}", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": group_ms[dom], "algorithmic_bytes_per_launch": dom_bytes,
                     "group_ms": group_ms},
        "cpu_baseline": None,
    }


def adjoint_times(repeats: int) -> dict:
    """The 64,000-particle WCSPH rollout of 200 steps of 1e-4 s, loss =
    mean |x_T - x_T*|^2 against the viscosity-0.08 rollout: seconds of the
    plain forward rollout, of the recording forward (checkpoints every 20
    steps) and of the reverse pass (segment replay + adjoint kernels, best of
    `repeats` after one untimed), with the tape's footprint."""
    import numpy as np

    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld, Tape

    steps, dt, every = 200, 1e-4, 20
    x0 = scenes.adjoint_positions()
    v0 = scenes.adjoint_velocities(len(x0))
    n = len(x0)

    def world(alpha):
        w = FluidWorld(scenes.settings_adjoint_64k(alpha))
        scenes.container(w)
        w.add_particles(x0, v0)
        return w

    log(f"adjoint_64k: {n} particles, {steps} steps; forward rollout (viscosity 0.08)")
    target_world = world(0.08)
    target_world.step(dt)  # first-call set-up outside the timing
    target_world.close()
    target_world = world(0.08)
    target_world.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        target_world.step(dt)
    target_world.synchronize()
    forward = time.perf_counter() - t0
    target = np.empty_like(x0)
    target[target_world.ids()] = target_world.positions()
    target_world.close()
    log(f"forward {forward * 1e3:.1f} ms; recording rollout (viscosity 0.05)")
    w = world(0.05)
    w.synchronize()
    t0 = time.perf_counter()
    tape = Tape(w, dt, steps, every)
    w.synchronize()
    record = time.perf_counter() - t0
    xT = np.empty_like(x0)
    xT[w.ids()] = w.positions()
    seed = (2.0 * (xT.astype(np.float64) - target) / n).astype(np.float32)
    loss = float(np.sum((xT.astype(np.float64) - target) ** 2) / n)
    log(f"record {record * 1e3:.1f} ms; reverse pass (warm-up, then {max(repeats, 1)} timed)")
    tape.adjoint(seed)
    runs = []
    for _ in range(max(repeats, 1)):
        t0 = time.perf_counter()
        gx, gv, ga = tape.adjoint(seed)
        runs.append(time.perf_counter() - t0)
    reverse = min(runs)
    memory = tape.memory_bytes()
    tape.close()
    w.close()
    log(f"reverse {reverse * 1e3:.1f} ms (best of {len(runs)})")
    return {"particles": n, "steps": steps, "checkpoint_every": every, "forward_s": forward, "record_s": record,
            "reverse_s": reverse, "reverse_runs": len(runs), "loss": loss, "dloss_dviscosity": ga,
            "tape_bytes": memory}


def run_adjoint(args) -> dict:
    """Differentiable WCSPH: 64,000 particles, 200 steps of 1e-4 s, loss =
    mean |x_T - x_T*|^2 against the viscosity-0.08 rollout. Times the plain
    forward rollout, the recording forward (checkpoints every 20 steps) and
    the reverse pass (segment replay + adjoint kernels); the metric is
    reverse / forward."""
    a = adjoint_times(args.steps)
    forward, record, reverse = a["forward_s"], a["record_s"], a["reverse_s"]
    n, steps = a["particles"], a["steps"]
    return {
        "metric": "adjoint:forward time ratio, 64k WCSPH rollout",
        "value": reverse / forward,
        "unit": "x",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": reverse * 1e3 / steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "adjoint_64k: WCSPH, spacing 0.01 m, 0.4 m cube of 64,000 particles in a 1.6 x 1.6 x "
                               "0.8 m container, v0 ~ N(0, 0.1 m/s), 200 steps of 1e-4 s, checkpoints every 20 "
                               "steps, loss = mean |x_T - x_T*|^2 against viscosity 0.08, gradients in v0 and alpha",
                   "forward_ms": forward * 1e3, "record_ms": record * 1e3, "reverse_ms": reverse * 1e3,
                   "reverse_runs": a["reverse_runs"],
                   "record_plus_reverse_over_forward": (record + reverse) / forward,
                   "forward_particle_steps_per_s": n * steps / forward,
                   "reverse_particle_steps_per_s": n * steps / reverse,
                   "loss": a["loss"], "dL_dalpha": a["dloss_dviscosity"], "tape_device_bytes": a["tape_bytes"]},
        "roofline": None,
        "cpu_baseline": None,
    }


# flows_1m: the dam_1m scene with an emitter above the water (40 x 10 x 40
# lattice slots, new water moving down at 1 m/s), a drain on the far floor
# and a domain around the container. After each step a drain or the domain
# reads every particle's position (16 B) to flag it, an emitter reads them
# once more to mark its taken slots; the flows' algorithmic bytes per step
# are those reads.
FLOW_EMITTER = ((0.1, 0.85, 0.1), (0.3, 0.9, 0.3), (0.0, -1.0, 0.0))
FLOW_DRAIN = ((1.2, 0.0, 0.0), (1.6, 0.1, 0.8))
FLOW_DOMAIN = ((-0.05, -0.05, -0.05), (1.65, 1.65, 0.85))


def run_flows(args) -> dict:
    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld

    world = FluidWorld(scenes.settings_dam_1m())
    n0 = scenes.dam_break(world)
    world.add_emitter(*FLOW_EMITTER)
    world.add_drain(*FLOW_DRAIN)
    world.set_domain(FLOW_DOMAIN)
    log(f"flows_1m: {n0} fluid particles, emitter, drain, domain; warmup {args.warmup} steps")
    for k in range(args.warmup):
        r = world.step(DT)
        log(f"warmup step {k}: {r.substeps} substeps, emitted {r.emitted}, drained {r.drained}, lost {r.lost}")
    world.synchronize()
    world.profile(True)
    world.profile_read(reset=True)
    world.synchronize()
    particle_steps = 0
    emitted = drained = lost = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        particle_steps += len(world)
        r = world.step(DT)
        emitted += r.emitted
        drained += r.drained
        lost += r.lost
    world.synchronize()
    elapsed = time.perf_counter() - t0
    phases = world.profile_read(reset=True)
    world.profile(False)
    n_end = len(world)
    world.close()
    flows_ms, flows_launches = phases.get("flows", (0.0, 0))
    flow_bytes = 16 * 3 * particle_steps  # drain, domain and emitter each read every position
    flows_s = flows_ms * 1e-3
    return {
        "metric": "particle-steps/s at 1M DFSPH with flows",
        "value": particle_steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "flows_1m: dam_1m (DFSPH, 1,024,000 fluid particles, spacing 0.005 m, step 5e-4 s) "
                               "with an emitter of 40 x 10 x 40 slots above the water (new water at 1 m/s "
                               "downwards), a drain on the far floor and a domain around the container",
                   "fluid_particles_start": n0, "fluid_particles_end": n_end,
                   "emitted": emitted, "drained": drained, "lost": lost},
        "flows": {"ms_per_step": flows_ms / max(args.steps, 1), "share_of_step": flows_s / elapsed,
                  "launches": flows_launches,
                  "achieved_GBps": flow_bytes / flows_s / 1e9 if flows_s > 0 else None,
                  "algorithmic_bytes": flow_bytes},
        "phase_ms": {p: round(v[0], 3) for p, v in phases.items()},
        "roofline": None,
        "cpu_baseline": None,
    }


def run_bodies(args) -> dict:
    """bodies_1m: the dam break under WCSPH with 200 coupled spheres, each
    step taking the bodies' poses (step_with_bodies) and handing back every
    body's lift and damping impulses; then the same world stepped without
    poses (plain WCSPH) for the cost of the coupling."""
    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld
    from solunar_amd.types import Shape

    world = FluidWorld(scenes.settings_bodies_1m())
    n = scenes.dam_break(world)
    field = scenes.body_field()
    for handle, r, *_ in field:
        world.couple(handle, Shape.sphere(r))
    log(f"bodies_1m: {n} fluid particles, {len(field)} coupled spheres; warmup {args.warmup} steps")
    t = 0.0
    for k in range(args.warmup):
        r, _ = world.step_with_bodies(DT, scenes.body_poses_array(field, t))
        t += DT
        log(f"warmup step {k}: {r.substeps} substeps")
    world.synchronize()
    world.profile(True)
    world.profile_read(reset=True)
    particle_steps = substeps = 0
    lifted = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        particle_steps += len(world)
        r, pushes = world.step_with_bodies(DT, scenes.body_poses_array(field, t))
        t += DT
        substeps += r.substeps
        lifted += int(pushes["lifted"].sum())
    world.synchronize()
    elapsed = time.perf_counter() - t0
    phases = world.profile_read(reset=True)
    world.profile(False)
    log(f"coupled: {elapsed * 1e3 / args.steps:.2f} ms a step; timing plain steps")
    world.synchronize()
    plain_steps = 0
    t1 = time.perf_counter()
    for k in range(args.steps):
        plain_steps += len(world)
        world.step(DT)
    world.synchronize()
    plain = time.perf_counter() - t1
    points = sum(len(scenes.sphere_points((0, 0, 0), r, 0.005, 0.01)) for _, r, *_ in field)
    world.close()
    bodies_ms, bodies_launches = phases.get("bodies", (0.0, 0))
    return {
        "metric": "particle-steps/s at 1M WCSPH with 200 coupled bodies",
        "value": particle_steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "bodies_1m: dam_1m's 1,024,000 fluid particles under WCSPH (spacing 0.005 m, "
                               "step 5e-4 s) with 200 coupled spheres of radius 0.02 to 0.08 m, 100 settling "
                               "onto the water's top at 0.1 m/s and 100 on the dry floor downstream, drifting "
                               "and turning about y, every other one dynamic",
                   "fluid_particles": n, "bodies": len(field), "body_points_approx": points,
                   "substeps_per_step": substeps / max(args.steps, 1),
                   "lifts_per_step": lifted / max(args.steps, 1)},
        "bodies": {"ms_per_step": bodies_ms / max(args.steps, 1), "share_of_step": bodies_ms * 1e-3 / elapsed,
                   "launches": bodies_launches},
        "plain_wcsph": {"value": plain_steps / plain, "ms_per_step": plain * 1e3 / args.steps},
        "phase_ms": {p: round(v[0], 3) for p, v in phases.items()},
        "roofline": force_roofline(phases, n, FORCE_BYTES),
        "cpu_baseline": None if args.no_cpu else bodies_cpu(args.cpu_steps),
    }


def bodies_cpu(steps: int):
    from solunar_amd import scenes
    from solunar_amd.types import Shape

    log("cpu baseline (bodies_1m)")
    field = scenes.body_field()

    def build(make):
        w = make(scenes.settings_bodies_1m())
        n = scenes.dam_break(w)
        for handle, r, *_ in field:
            w.couple(handle, Shape.sphere(r))
        return w, n

    return cpu_sample("bodies_1m", build, lambda w, k: w.step_with_bodies(DT, scenes.body_poses(field, k * DT))[0],
                      steps)


def run_trench(args) -> dict:
    """trench_720k: a WCSPH dam break down a trench in terrain ground: the
    ground's density sites, slope, viscosity and contact in every substep."""
    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld

    world = FluidWorld(scenes.settings_trench_720k())
    n = scenes.trench_dam(world)
    log(f"trench_720k: {n} fluid particles; warmup {args.warmup} steps")
    for k in range(args.warmup):
        r = world.step(DT)
        log(f"warmup step {k}: {r.substeps} substeps")
    world.synchronize()
    world.profile(True)
    world.profile_read(reset=True)
    particle_steps = substeps = held = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        particle_steps += len(world)
        r = world.step(DT)
        substeps += r.substeps
        held += r.held_back
    world.synchronize()
    elapsed = time.perf_counter() - t0
    phases = world.profile_read(reset=True)
    world.profile(False)
    world.close()
    return {
        "metric": "particle-steps/s of a WCSPH dam break in terrain",
        "value": particle_steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "trench_720k: WCSPH, spacing 0.005 m, step 5e-4 s; 720,000 particles (1 x 0.3 x "
                               "0.3 m) released down a 2 m trench 0.3 m deep and wide in flat ground, one terrain "
                               "region of 833 x 137 x 137 nodes",
                   "fluid_particles": n, "substeps_per_step": substeps / max(args.steps, 1),
                   "held_back": held},
        "phase_ms": {p: round(v[0], 3) for p, v in phases.items()},
        "roofline": force_roofline(phases, n, FORCE_BYTES + FORCE_SLOPE_BYTES),
        "cpu_baseline": None if args.no_cpu else trench_cpu(args.cpu_steps),
    }


def trench_cpu(steps: int):
    from solunar_amd import scenes

    log("cpu baseline (trench_720k)")

    def build(make):
        w = make(scenes.settings_trench_720k())
        return w, scenes.trench_dam(w)

    return cpu_sample("trench_720k", build, lambda w, k: w.step(DT), steps)


def run_snapshot(args) -> dict:
    """snapshot_1m: save and load of the dam_1m world after `warmup` steps
    (1,024,000 particles at 28 B and 832,064 walls at 16 B in the SLFL
    bytes). Both cross PCIe: save reads x and v back, load uploads them and
    settles the walls."""
    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld

    world = FluidWorld(scenes.settings_dam_1m())
    n = scenes.dam_break(world)
    for _ in range(args.warmup):
        world.step(DT)
    world.synchronize()
    reps = max(1, min(args.steps, 5))
    t0 = time.perf_counter()
    for _ in range(reps):
        data = world.save_snapshot()
    save_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        loaded = FluidWorld.load_snapshot(data)
        loaded.synchronize()
        same = loaded.state_hash() == world.state_hash()
        loaded.close()
    load_s = (time.perf_counter() - t0) / reps
    world.close()
    return {
        "metric": "snapshot bytes/s, 1M DFSPH world (save)",
        "value": len(data) / save_s,
        "unit": "B/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": save_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "snapshot_1m: SLFL v3 save and load of the dam_1m world (1,024,000 fluid, "
                               "832,064 wall particles) after the warmup steps",
                   "bytes": len(data), "fluid_particles": n},
        "save_ms": save_s * 1e3,
        "load_ms": load_s * 1e3,
        "load_bytes_per_s": len(data) / load_s,
        "loaded_state_hash_equal": same,
        "roofline": None,
        "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="dam_1m: skip the 200-step rollout and the adjoint ratio after the timed steps")
    ap.add_argument("--workload", choices=["dam_1m", "dam_16m", "batch_256", "adjoint_64k", "flows_1m",
                                           "snapshot_1m", "bodies_1m", "trench_720k"],
                    default="dam_1m",
                    help="dam_16m: 16,000,000 fluid particles at spacing 0.002 with 200 sphere obstacles; "
                         "batch_256: 256 WCSPH scenes in one batch; flows_1m: dam_1m with an emitter, a "
                         "drain and a domain")
    ap.add_argument("--per-scene-steps", type=int, default=10,
                    help="batch_256: steps of each world timed on the single-world path")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 256 MiB device copy, the known-size kernel that scales the rocprofv3 byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the fluid step here is built and timed on a single MI355X "
              "(multi-GPU sharding is a design note, not a code path)", file=sys.stderr)
        return 2

    if args.workload == "batch_256":
        print(json.dumps(run_batch(args)), flush=True)
        return 0
    if args.workload == "adjoint_64k":
        print(json.dumps(run_adjoint(args)), flush=True)
        return 0
    if args.workload == "flows_1m":
        print(json.dumps(run_flows(args)), flush=True)
        return 0
    if args.workload == "snapshot_1m":
        print(json.dumps(run_snapshot(args)), flush=True)
        return 0
    if args.workload == "bodies_1m":
        print(json.dumps(run_bodies(args)), flush=True)
        return 0
    if args.workload == "trench_720k":
        print(json.dumps(run_trench(args)), flush=True)
        return 0

    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld

    large = args.workload == "dam_16m"
    settings = scenes.settings_dam_16m() if large else scenes.settings_dam_1m()
    world = FluidWorld(settings)
    n = scenes.dam_with_obstacles(world) if large else scenes.dam_break(world)
    walls = len(world.wall_positions())
    log(f"{args.workload}: {n} fluid, {walls} wall particles; warmup {args.warmup} steps")
    for k in range(args.warmup):
        r = world.step(DT)
        log(f"warmup step {k}: {r.substeps} substeps, {r.iterations} iterations")
    world.synchronize()
    world.profile(True)
    world.profile_read(reset=True)
    world.work_read(reset=True)
    log(f"timing {args.steps} steps")
    world.synchronize()
    t0 = time.perf_counter()
    substeps = iterations = 0
    for k in range(args.steps):
        r = world.step(DT)
        substeps += r.substeps
        iterations += r.iterations
        if k % 5 == 4:
            log(f"step {k + 1}/{args.steps}")
    world.synchronize()
    elapsed = time.perf_counter() - t0
    phases = world.profile_read(reset=True)
    work = world.work_read(reset=True)
    world.profile(False)
    solve_blocks = world.solve_grid()
    memory = world.memory_bytes()
    log(f"timed: {elapsed:.3f} s, {substeps} substeps, {iterations} iterations")

    value = n * args.steps / elapsed
    # dominant kernel by event time
    dom = max(phases, key=lambda p: phases[p][0])
    dom_ms, dom_launches = phases[dom]
    if dom == "solve":
        dom_bytes = RATE_BYTES * work["particle_rates"] + PUSH_BYTES * work["particle_pushes"]
        # the tiled density solve runs the viscosity pass in its own launch
        # (no viscous launches of their own then)
        if phases.get("viscous", (0.0, 0))[1] == 0:
            dom_bytes += VISCOUS_BYTES * work["particle_substeps"]
    else:
        per = {"bounds": 12, "sort": 40, "reorder": 64, "density": 24, "viscous": 52, "advance": 48, "force": 52}
        dom_bytes = per.get(dom, 0) * work["particle_substeps"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    step_bytes = (SUBSTEP_FIXED_BYTES * work["particle_substeps"] + RATE_BYTES * work["particle_rates"]
                  + PUSH_BYTES * work["particle_pushes"])
    # HBM bytes per launch from the committed rocprofv3 counter summary
    # (FETCH_SIZE/WRITE_SIZE per Jacobi iteration, scaled by a plain copy),
    # times this run's iterations per launch
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_dam_1m.json")
    if large:
        tpath = ""
    if os.path.exists(tpath) and dom == "solve" and dom_launches:
        with open(tpath) as f:
            per_iteration = json.load(f)["bytes_per_iteration"]
        traffic = per_iteration * (work["particle_pushes"] / n) / dom_launches
    out = {
        "metric": "particle-steps/s at 1M DFSPH" if not large else "particle-steps/s at 16M DFSPH with obstacles",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("dam_1m: DFSPH dam break, 1,024,000 fluid + 832,064 wall particles, spacing "
                                "0.005 m, step 5e-4 s, divergence tol 1e-3, density tol 1e-4") if not large else
                               ("dam_16m: DFSPH dam break, 16,000,000 fluid particles, spacing 0.002 m, container "
                                "walls + 200 sphere obstacles (radius 0.02-0.08 m) as boundary particles, step 5e-4 s"),
                   "device_memory_bytes": memory, "device_memory_fraction_of_288GB": memory / 288e9,
                   "fluid_particles": n, "wall_particles": walls, "substeps": substeps,
                   "jacobi_iterations": iterations, "solve_blocks": solve_blocks},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "launches": dom_launches, "avg_launch_ms": dom_ms / max(dom_launches, 1),
                     "algorithmic_bytes_per_launch": dom_bytes / max(dom_launches, 1)},
        "step_roofline": {"achieved": step_bytes / elapsed / 1e9, "frac": step_bytes / elapsed / 1e9 / HBM_PEAK_GBS,
                          "unit": "GB/s", "bytes": step_bytes},
        "phase_ms": {p: round(v[0], 3) for p, v in phases.items()},
    }
    if not large and not args.no_extras:
        out["rollout_200"] = full_rollout(settings)
        a = adjoint_times(2)
        out["adjoint_64k"] = {
            "reverse_over_forward": a["reverse_s"] / a["forward_s"],
            "record_and_reverse_over_forward": (a["record_s"] + a["reverse_s"]) / a["forward_s"],
            "forward_ms": a["forward_s"] * 1e3, "record_ms": a["record_s"] * 1e3, "reverse_ms": a["reverse_s"] * 1e3,
            "particles": a["particles"], "steps": a["steps"], "checkpoint_every": a["checkpoint_every"],
        }
    if not args.no_cpu and not large:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, args.warmup)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


def full_rollout(settings) -> dict:
    """The whole 200-step dam_1m rollout from a fresh world, steps 0-199
    timed once (the first step's buffer set-up included), and the rate
    without step 0."""
    from solunar_amd import scenes
    from solunar_amd.fluid import FluidWorld

    w = FluidWorld(settings)
    n = scenes.dam_break(w)
    log("rollout: 200 steps of a fresh dam_1m world")
    w.synchronize()
    t0 = time.perf_counter()
    w.step(DT)
    w.synchronize()
    t1 = time.perf_counter()
    substeps = iterations = 0
    for k in range(1, 200):
        r = w.step(DT)
        substeps += r.substeps
        iterations += r.iterations
        if k % 50 == 49:
            log(f"rollout step {k + 1}/200")
    w.synchronize()
    t2 = time.perf_counter()
    w.close()
    return {"steps": 200, "seconds": t2 - t0, "value": n * 200 / (t2 - t0), "unit": "particle-steps/s",
            "first_step_ms": (t1 - t0) * 1e3, "value_steps_1_to_199": n * 199 / (t2 - t1),
            "substeps_1_to_199": substeps, "jacobi_iterations_1_to_199": iterations}


if __name__ == "__main__":
    sys.exit(main())
