"""TEST INFRASTRUCTURE: ctypes loader of the CPU oracle (liboracle_fluid.so).

`OracleWorld` has the same methods as `solunar_amd.fluid.FluidWorld`, so a
test can run one scenario on both and compare. Only tests/, the smoke check
and the CPU baseline leg of bench.py use it.

Extras the reference's test dispatchers give: `threads` (jobs of 512
particles on that many workers, run backwards like dispatch::testing::Threads,
src/dispatch.rs:28-48), `skip_job` (dispatch::testing::SkipsJobs,
src/dispatch.rs:50-65), `precision` (32 = reference arithmetic, 64 = FP64
yardstick) and `residual_mode` (0 = the reference's f32 fold of DFSPH errors,
1 = the device's fixed-shape FP64 sum over its solve tiles).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

from solunar_amd.types import (CBodyPose, CBodyPush, CReport, CSettings, CShape, FluidReport, FluidSettings,
                               PhysicsError, Shape, SnapshotError, pose_array, terrain_samples)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_fluid.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    f = ctypes.c_float
    sz = ctypes.c_size_t
    i32 = ctypes.c_int32
    u32p = ctypes.POINTER(ctypes.c_uint32)
    fp = ctypes.POINTER(ctypes.c_float)
    box = [P, f, f, f, f, f, f]
    sigs = {
        "orc_settings_size": (sz, []),
        "orc_report_size": (sz, []),
        "orc_create": (i32, [ctypes.POINTER(CSettings), i32, ctypes.POINTER(P)]),
        "orc_destroy": (None, [P]),
        "orc_set_threads": (None, [P, i32]),
        "orc_set_skip_job": (None, [P, ctypes.c_int64]),
        "orc_set_residual_mode": (None, [P, i32]),
        "orc_set_solve_log": (None, [P, i32]),
        "orc_read_solve_log": (sz, [P, sz, ctypes.POINTER(i32), u32p, u32p, fp, ctypes.POINTER(ctypes.c_double)]),
        "orc_len": (sz, [P]),
        "orc_particle_mass": (ctypes.c_double, [P]),
        "orc_kernel_value": (ctypes.c_double, [P, ctypes.c_double]),
        "orc_kernel_slope": (ctypes.c_double, [P, ctypes.c_double]),
        "orc_add_particles": (i32, [P, sz, fp, fp, u32p]),
        "orc_fill_box": (i32, box + [ctypes.POINTER(sz)]),
        "orc_fill_box_jittered": (i32, box + [f, ctypes.c_uint32, ctypes.POINTER(sz)]),
        "orc_add_box_container": (i32, box + [u32p]),
        "orc_add_wall": (i32, box + [u32p]),
        "orc_add_boundary_points": (i32, [P, sz, fp, u32p]),
        "orc_remove_boundary": (i32, [P, ctypes.c_uint32, ctypes.POINTER(i32)]),
        "orc_set_domain": (i32, [P, fp, fp]),
        "orc_add_emitter": (i32, [P, fp, fp, fp, u32p]),
        "orc_add_drain": (i32, [P, fp, fp, u32p]),
        "orc_remove_flow": (i32, [P, ctypes.c_uint32]),
        "orc_run_flows": (i32, [P, u32p, u32p]),
        "orc_save_snapshot": (i32, [P, ctypes.POINTER(ctypes.c_uint8), sz, ctypes.POINTER(sz)]),
        "orc_load_snapshot": (i32, [ctypes.POINTER(ctypes.c_uint8), sz, ctypes.POINTER(CSettings), i32,
                                    ctypes.POINTER(P), ctypes.POINTER(i32), u32p, ctypes.POINTER(ctypes.c_char_p)]),
        "orc_wall_len": (sz, [P]),
        "orc_read_wall_positions": (i32, [P, sz, fp]),
        "orc_step": (i32, [P, f, ctypes.POINTER(CReport)]),
        "orc_update_neighbors": (i32, [P]),
        "orc_compute_densities": (i32, [P]),
        "orc_read_ids": (i32, [P, sz, u32p]),
        "orc_read_vectors": (i32, [P, i32, sz, ctypes.c_void_p]),
        "orc_add_particles_f64": (i32, [P, sz, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "orc_adjoint_record": (i32, [P, ctypes.c_float, ctypes.c_uint32, i32]),
        "orc_adjoint_reverse": (i32, [P] + [ctypes.POINTER(ctypes.c_double)] * 5),
        "orc_read_densities": (i32, [P, sz, ctypes.c_void_p]),
        "orc_set_particles": (i32, [P, sz, u32p, fp, fp]),
        "orc_read_neighbors": (i32, [P, i32, sz, u32p, sz, u32p, ctypes.POINTER(sz)]),
        "orc_state_hash": (ctypes.c_uint64, [P]),
        "orc_dfsph_probe": (i32, [P, f] + [ctypes.POINTER(ctypes.c_double)] * 8),
        "orc_error_detail": (ctypes.c_char_p, [P]),
        "orc_fnv1a": (ctypes.c_uint64, [ctypes.c_char_p, sz]),
        "orc_tile_tree_sum": (ctypes.c_double, [fp, ctypes.POINTER(i32), sz]),
        "orc_grid_selftest": (i32, []),
        "orc_plant_mistake": (None, [i32]),
        "orc_terrain_nodes": (i32, [P, fp, fp, ctypes.POINTER(sz)]),
        "orc_add_terrain": (i32, [P, fp, fp, fp, sz, u32p]),
        "orc_terrain_len": (sz, [P]),
        "orc_terrain_sample": (i32, [P, sz, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]),
        "orc_ground_near": (i32, [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]),
        "orc_couple": (i32, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(CShape)]),
        "orc_uncouple": (i32, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(i32)]),
        "orc_coupled_len": (sz, [P]),
        "orc_coupled_info": (i32, [P, sz, u32p, u32p, ctypes.POINTER(sz), ctypes.POINTER(sz), fp]),
        "orc_read_coupled_points": (i32, [P, sz, i32, sz, fp]),
        "orc_place_movers": (i32, [P, ctypes.POINTER(CBodyPose), sz]),
        "orc_mover_len": (sz, [P]),
        "orc_read_movers": (i32, [P, sz, fp, fp, u32p]),
        "orc_lifts": (i32, [P, f, ctypes.POINTER(CBodyPose), sz, ctypes.POINTER(CBodyPush), sz, ctypes.POINTER(sz)]),
        "orc_step_with_bodies": (i32, [P, f, ctypes.POINTER(CBodyPose), sz, ctypes.POINTER(CReport),
                                       ctypes.POINTER(CBodyPush), sz, ctypes.POINTER(sz)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    assert lib.orc_settings_size() == ctypes.sizeof(CSettings), "settings layout differs"
    assert lib.orc_report_size() == ctypes.sizeof(CReport), "report layout differs"
    _lib = lib
    return lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, 3))


def settings_from_snapshot(data: bytes, base: Optional[FluidSettings]) -> FluidSettings:
    """The settings fields an accepted snapshot holds (save.rs:20-40) over `base`."""
    import struct
    from dataclasses import replace

    solver = data[8]
    spacing, rest, support, sound, viscosity, courant = struct.unpack_from("<6f", data, 9)
    (max_particles,) = struct.unpack_from("<Q", data, 33)
    gravity = struct.unpack_from("<3f", data, 41)
    (max_substeps,) = struct.unpack_from("<I", data, 53)
    return replace(base or FluidSettings(), solver=solver, spacing=spacing, rest_density=rest, support=support,
                   sound_speed=sound, viscosity=viscosity, courant=courant, max_particles=max_particles,
                   gravity=tuple(gravity), max_substeps=max_substeps)


def _v3(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3))


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class OracleWorld:
    def __init__(self, settings: Optional[FluidSettings] = None, precision: int = 32, threads: int = 1,
                 residual_mode: int = 0):
        self._lib = library()
        self._settings = settings or FluidSettings()
        self.precision = 64 if precision == 64 else 32
        h = ctypes.c_void_p()
        st = self._lib.orc_create(ctypes.byref(CSettings.of(self._settings)), self.precision, ctypes.byref(h))
        if st:
            raise PhysicsError(st)
        self._h = h
        self.set_threads(threads)
        self.set_residual_mode(residual_mode)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def _check(self, st):
        if st:
            raise PhysicsError(st, self._lib.orc_error_detail(self._h).decode())

    # controls
    def set_threads(self, n: int) -> None:
        self._lib.orc_set_threads(self._h, n)

    def skip_job(self, job: int) -> None:
        self._lib.orc_set_skip_job(self._h, job)

    def set_residual_mode(self, mode: int) -> None:
        self._lib.orc_set_residual_mode(self._h, mode)

    def log_solves(self, enable: bool = True) -> None:
        self._lib.orc_set_solve_log(self._h, 1 if enable else 0)

    def solve_log(self):
        n = self._lib.orc_read_solve_log(self._h, 0, None, None, None, None, None)
        s = np.zeros(n, np.int32)
        it = np.zeros(n, np.uint32)
        ld = np.zeros(n, np.uint32)
        a32 = np.zeros(n, np.float32)
        a64 = np.zeros(n, np.float64)
        self._lib.orc_read_solve_log(self._h, n, _p(s, ctypes.c_int32), _p(it, ctypes.c_uint32),
                                     _p(ld, ctypes.c_uint32), _p(a32, ctypes.c_float), _p(a64, ctypes.c_double))
        return {"solve": s, "iteration": it, "loaded": ld, "average_f32": a32, "average_f64": a64}

    def kernel_value(self, d: float) -> float:
        return self._lib.orc_kernel_value(self._h, d)

    def kernel_slope(self, d: float) -> float:
        return self._lib.orc_kernel_slope(self._h, d)

    # API mirror
    @property
    def settings(self) -> FluidSettings:
        return self._settings

    def particle_mass(self) -> float:
        return float(self._lib.orc_particle_mass(self._h))

    def __len__(self) -> int:
        return int(self._lib.orc_len(self._h))

    def is_empty(self) -> bool:
        return len(self) == 0

    def add_particle(self, position, velocity=(0.0, 0.0, 0.0)) -> int:
        return int(self.add_particles([position], [velocity])[0])

    def add_particles(self, positions, velocities=None) -> np.ndarray:
        p = _f32(positions)
        v = None if velocities is None else _f32(velocities)
        ids = np.zeros(p.shape[0], np.uint32)
        self._check(self._lib.orc_add_particles(self._h, p.shape[0], _p(p, ctypes.c_float),
                                                None if v is None else _p(v, ctypes.c_float),
                                                _p(ids, ctypes.c_uint32)))
        return ids

    def fill_box(self, low, high) -> int:
        c = ctypes.c_size_t()
        self._check(self._lib.orc_fill_box(self._h, *map(float, low), *map(float, high), ctypes.byref(c)))
        return c.value

    def fill_box_jittered(self, low, high, jitter: float, seed: int) -> int:
        c = ctypes.c_size_t()
        self._check(self._lib.orc_fill_box_jittered(self._h, *map(float, low), *map(float, high), jitter, seed,
                                                    ctypes.byref(c)))
        return c.value

    def add_box_container(self, low, high) -> int:
        o = ctypes.c_uint32()
        self._check(self._lib.orc_add_box_container(self._h, *map(float, low), *map(float, high), ctypes.byref(o)))
        return o.value

    def add_wall(self, low, high) -> int:
        o = ctypes.c_uint32()
        self._check(self._lib.orc_add_wall(self._h, *map(float, low), *map(float, high), ctypes.byref(o)))
        return o.value

    def set_domain(self, domain) -> None:
        """Water outside domain = (low, high) is taken out after each step; None clears it."""
        if domain is None:
            self._check(self._lib.orc_set_domain(self._h, None, None))
            return
        lo, hi = _v3(domain[0]), _v3(domain[1])
        self._check(self._lib.orc_set_domain(self._h, _p(lo, ctypes.c_float), _p(hi, ctypes.c_float)))

    def add_emitter(self, low, high, velocity) -> int:
        lo, hi, v = _v3(low), _v3(high), _v3(velocity)
        o = ctypes.c_uint32()
        self._check(self._lib.orc_add_emitter(self._h, _p(lo, ctypes.c_float), _p(hi, ctypes.c_float),
                                              _p(v, ctypes.c_float), ctypes.byref(o)))
        return o.value

    def add_drain(self, low, high) -> int:
        lo, hi = _v3(low), _v3(high)
        o = ctypes.c_uint32()
        self._check(self._lib.orc_add_drain(self._h, _p(lo, ctypes.c_float), _p(hi, ctypes.c_float), ctypes.byref(o)))
        return o.value

    def save_snapshot(self) -> bytes:
        """The SLFL version 3 bytes (src/fluid/save.rs:17-66)."""
        n = ctypes.c_size_t()
        self._check(self._lib.orc_save_snapshot(self._h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_uint8 * n.value)()
        self._check(self._lib.orc_save_snapshot(self._h, buf, n.value, ctypes.byref(n)))
        return bytes(buf)

    @classmethod
    def load_snapshot(cls, data: bytes, settings: Optional[FluidSettings] = None, precision: int = 32,
                      threads: int = 1, residual_mode: int = 0) -> "OracleWorld":
        """A world from snapshot bytes (src/fluid/save.rs:68-195); raises
        SnapshotError for refused bytes, PhysicsError(101) for coupled bodies
        or terrain."""
        lib = library()
        data = bytes(data)
        buf = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data + b"\0")
        base = CSettings.of(settings) if settings is not None else None
        h = ctypes.c_void_p()
        kind, version, what = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_char_p()
        st = lib.orc_load_snapshot(buf, len(data), ctypes.byref(base) if base is not None else None,
                                   64 if precision == 64 else 32, ctypes.byref(h), ctypes.byref(kind),
                                   ctypes.byref(version), ctypes.byref(what))
        if st == 101:
            raise PhysicsError(101, (what.value or b"").decode())
        if st:
            raise SnapshotError(kind.value, version.value, (what.value or b"").decode())
        w = cls.__new__(cls)
        w._lib = lib
        w._h = h
        w.precision = 64 if precision == 64 else 32
        w._settings = settings_from_snapshot(data, settings)
        w.set_threads(threads)
        w.set_residual_mode(residual_mode)
        return w

    # ---- terrain ground (src/fluid/terrain.rs) ----

    def terrain_nodes(self, low, high):
        """The lattice add_terrain samples over the box: nodes per axis (x, y, z)."""
        lo, hi = _v3(low), _v3(high)
        n = (ctypes.c_size_t * 3)()
        self._check(self._lib.orc_terrain_nodes(self._h, _p(lo, ctypes.c_float), _p(hi, ctypes.c_float), n))
        return tuple(n)

    def add_terrain(self, low, high, distance) -> int:
        """add_terrain (terrain.rs:262-305): `distance` is the terrain's signed
        distance, a function of (N, 3) float32 points, or the sampled array."""
        lo, hi = _v3(low), _v3(high)
        d = terrain_samples(self, lo, hi, distance)
        o = ctypes.c_uint32()
        self._check(self._lib.orc_add_terrain(self._h, _p(lo, ctypes.c_float), _p(hi, ctypes.c_float),
                                              _p(d, ctypes.c_float), d.size, ctypes.byref(o)))
        return o.value

    def terrain_len(self) -> int:
        return int(self._lib.orc_terrain_len(self._h))

    def terrain_sample(self, k: int, point):
        """(distance, normal) of field k at `point`, or None outside it."""
        p = np.asarray(point, np.float64)
        d = ctypes.c_double()
        n = (ctypes.c_double * 3)()
        st = self._lib.orc_terrain_sample(self._h, k, _p(p, ctypes.c_double), ctypes.byref(d), n)
        if st == 10:
            return None
        self._check(st)
        return d.value, np.array(n[:])

    def ground_near(self, point):
        """(density, gradient) the ground adds at `point` (terrain.rs:144-170, 307-313)."""
        p = np.asarray(point, np.float64)
        d = ctypes.c_double()
        g = (ctypes.c_double * 3)()
        self._check(self._lib.orc_ground_near(self._h, _p(p, ctypes.c_double), ctypes.byref(d), g))
        return d.value, np.array(g[:])

    # ---- coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) ----

    def couple(self, body, shape: Shape) -> None:
        """Lets water and `body` (index, generation) push each other (couple.rs:165-198)."""
        c = CShape.of(shape)
        self._check(self._lib.orc_couple(self._h, int(body[0]), int(body[1]), ctypes.byref(c)))

    def uncouple(self, body) -> bool:
        o = ctypes.c_int32()
        self._check(self._lib.orc_uncouple(self._h, int(body[0]), int(body[1]), ctypes.byref(o)))
        return bool(o.value)

    def coupled(self):
        """[(body, points, probes, probe_spacing)] in handle order, points in body axes."""
        out = []
        for k in range(int(self._lib.orc_coupled_len(self._h))):
            i, g = ctypes.c_uint32(), ctypes.c_uint32()
            n, m, sp = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_float()
            self._check(self._lib.orc_coupled_info(self._h, k, ctypes.byref(i), ctypes.byref(g), ctypes.byref(n),
                                                   ctypes.byref(m), ctypes.byref(sp)))
            pts = np.zeros((n.value, 3), np.float32)
            prb = np.zeros((m.value, 3), np.float32)
            self._check(self._lib.orc_read_coupled_points(self._h, k, 0, n.value, _p(pts, ctypes.c_float)))
            self._check(self._lib.orc_read_coupled_points(self._h, k, 1, m.value, _p(prb, ctypes.c_float)))
            out.append(((i.value, g.value), pts, prb, sp.value))
        return out

    def place_movers(self, poses) -> None:
        arr = pose_array(poses)
        self._check(self._lib.orc_place_movers(self._h, arr, len(poses)))

    def movers(self):
        """(positions, velocities, owners) of the placed body points, in sorted order."""
        n = int(self._lib.orc_mover_len(self._h))
        x, v = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        o = np.zeros(n, np.uint32)
        self._check(self._lib.orc_read_movers(self._h, n, _p(x, ctypes.c_float), _p(v, ctypes.c_float),
                                              _p(o, ctypes.c_uint32)))
        return x, v, o

    def lifts(self, dt: float, poses):
        """lift.rs:25-152 over the present state, given back to no one."""
        arr = pose_array(poses)
        cap = int(self._lib.orc_coupled_len(self._h))
        out = (CBodyPush * max(cap, 1))()
        n = ctypes.c_size_t()
        self._check(self._lib.orc_lifts(self._h, dt, arr, len(poses), out, cap, ctypes.byref(n)))
        return [out[k].to_push() for k in range(n.value)]

    def step_with_bodies(self, dt: float, poses):
        """step_with_bodies (couple.rs:207-252): (report, [BodyPush] in handle order)."""
        arr = pose_array(poses)
        cap = int(self._lib.orc_coupled_len(self._h))
        out = (CBodyPush * max(cap, 1))()
        n = ctypes.c_size_t()
        r = CReport()
        self._check(self._lib.orc_step_with_bodies(self._h, dt, arr, len(poses), ctypes.byref(r), out, cap,
                                                   ctypes.byref(n)))
        return r.to_report(), [out[k].to_push() for k in range(n.value)]

    def run_flows(self):
        """Drains then emitters once, without a step: (emitted, drained)."""
        e, d = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.orc_run_flows(self._h, ctypes.byref(e), ctypes.byref(d)))
        return e.value, d.value

    def remove_flow(self, flow_id: int) -> bool:
        return bool(self._lib.orc_remove_flow(self._h, flow_id))

    def add_boundary_points(self, points) -> int:
        p = _f32(points)
        o = ctypes.c_uint32()
        self._check(self._lib.orc_add_boundary_points(self._h, p.shape[0], _p(p, ctypes.c_float), ctypes.byref(o)))
        return o.value

    def remove_boundary(self, boundary: int) -> bool:
        o = ctypes.c_int32()
        self._check(self._lib.orc_remove_boundary(self._h, boundary, ctypes.byref(o)))
        return bool(o.value)

    def wall_positions(self) -> np.ndarray:
        n = int(self._lib.orc_wall_len(self._h))
        out = np.zeros((n, 3), np.float32)
        self._check(self._lib.orc_read_wall_positions(self._h, n, _p(out, ctypes.c_float)))
        return out

    def step(self, dt: float) -> FluidReport:
        r = CReport()
        self._check(self._lib.orc_step(self._h, dt, ctypes.byref(r)))
        return r.to_report()

    def update_neighbors(self) -> None:
        self._check(self._lib.orc_update_neighbors(self._h))

    def compute_densities(self) -> None:
        self._check(self._lib.orc_compute_densities(self._h))

    def ids(self) -> np.ndarray:
        out = np.zeros(len(self), np.uint32)
        self._check(self._lib.orc_read_ids(self._h, out.shape[0], _p(out, ctypes.c_uint32)))
        return out

    def _vec(self, field: int) -> np.ndarray:
        dt = np.float64 if self.precision == 64 else np.float32
        out = np.zeros((len(self), 3), dt)
        self._check(self._lib.orc_read_vectors(self._h, field, out.shape[0], out.ctypes.data))
        return out

    def positions(self) -> np.ndarray:
        return self._vec(0)

    def velocities(self) -> np.ndarray:
        return self._vec(1)

    def forces(self) -> np.ndarray:
        return self._vec(2)

    def keep_forces(self, keep: bool = True) -> None:
        pass  # the oracle always keeps them

    def densities(self) -> np.ndarray:
        dt = np.float64 if self.precision == 64 else np.float32
        out = np.zeros(len(self), dt)
        self._check(self._lib.orc_read_densities(self._h, out.shape[0], out.ctypes.data))
        return out

    def set_particles(self, ids, positions, velocities) -> None:
        ids = np.ascontiguousarray(ids, np.uint32)
        p, v = _f32(positions), _f32(velocities)
        self._check(self._lib.orc_set_particles(self._h, ids.shape[0], _p(ids, ctypes.c_uint32),
                                                _p(p, ctypes.c_float), _p(v, ctypes.c_float)))

    def _lists(self, kind: int):
        n = len(self)
        start = np.zeros(n + 1, np.uint32)
        cnt = ctypes.c_size_t()
        self._lib.orc_read_neighbors(self._h, kind, 0, None, 0, None, ctypes.byref(cnt))
        items = np.zeros(max(cnt.value, 1), np.uint32)
        self._check(self._lib.orc_read_neighbors(self._h, kind, n + 1, _p(start, ctypes.c_uint32), items.shape[0],
                                                 _p(items, ctypes.c_uint32), ctypes.byref(cnt)))
        return start, items[: cnt.value]

    def neighbors(self):
        return self._lists(0)

    def wall_neighbors(self):
        return self._lists(1)

    def mover_neighbors(self):
        return self._lists(2)

    def neighbors_of(self, index: int) -> np.ndarray:
        s, it = self.neighbors()
        return it[s[index]:s[index + 1]]

    def dfsph_probe(self, dt: float, stiffness=None) -> dict:
        """The DFSPH quantities of a substep's first solve iteration at the
        current state, in storage order (call update_neighbors first):
        factors `alpha`, density `rate`, `div_stiffness`, `div_error`,
        `den_stiffness`, `den_error`, and, for a given stiffness, the velocity
        `change` push_apart makes (src/fluid/dfsph.rs:138-256)."""
        n = len(self)
        d = ctypes.POINTER(ctypes.c_double)
        out = {k: np.zeros(n) for k in ("alpha", "rate", "div_stiffness", "div_error", "den_stiffness", "den_error")}
        change = np.zeros((n, 3))
        k = None if stiffness is None else np.ascontiguousarray(np.asarray(stiffness, np.float64).reshape(n))
        self._check(self._lib.orc_dfsph_probe(
            self._h, dt, None if k is None else _p(k, ctypes.c_double),
            *[_p(out[name], ctypes.c_double) for name in ("alpha", "rate", "div_stiffness", "div_error",
                                                          "den_stiffness", "den_error")],
            _p(change, ctypes.c_double) if k is not None else ctypes.cast(None, d)))
        if k is not None:
            out["change"] = change
        return out

    def state_hash(self) -> int:
        return int(self._lib.orc_state_hash(self._h))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))


def _add_particles_f64(self, positions, velocities) -> None:
    """Particles from float64 inputs (exact in the FP64 world)."""
    x, v = _f64(positions), _f64(velocities)
    self._check(self._lib.orc_add_particles_f64(self._h, len(x), _p(x, ctypes.c_double), _p(v, ctypes.c_double)))


def _record(self, dt: float, steps: int, reverse_precision: Optional[int] = None) -> None:
    """`steps` WCSPH steps of dt, recorded for `adjoint`; the reverse pass
    runs in the world's precision, or in FP64 with reverse_precision=64."""
    rp = self.precision if reverse_precision is None else reverse_precision
    self._check(self._lib.orc_adjoint_record(self._h, dt, steps, rp))


def _adjoint(self, xbar, vbar=None):
    """Per-id adjoints of the final x, v -> per-id adjoints of the recorded
    initial x, v and dL/d(viscosity), in the world's precision."""
    xb = _f64(xbar)
    vb = _f64(np.zeros_like(xb) if vbar is None else vbar)
    x0 = np.zeros_like(xb)
    v0 = np.zeros_like(xb)
    a = ctypes.c_double()
    self._check(self._lib.orc_adjoint_reverse(self._h, _p(xb, ctypes.c_double), _p(vb, ctypes.c_double),
                                              _p(x0, ctypes.c_double), _p(v0, ctypes.c_double), ctypes.byref(a)))
    return x0, v0, a.value


OracleWorld.add_particles_f64 = _add_particles_f64
OracleWorld.record = _record
OracleWorld.adjoint = _adjoint


def plant_mistake(which: int) -> None:
    """Switches a deliberate mistake into the oracle (0 = none); CPU only."""
    library().orc_plant_mistake(which)


def fnv1a(data: bytes) -> int:
    return int(library().orc_fnv1a(data, len(data)))


def tile_tree_sum(values, cells) -> float:
    """The device's FP64 stop sum of `values` whose cells (n x 3 int32, in
    sorted order: by cell x, y, z) are `cells`."""
    v = np.ascontiguousarray(values, np.float32)
    c = np.ascontiguousarray(cells, np.int32).reshape(-1, 3)
    assert c.shape[0] == v.shape[0]
    return float(library().orc_tile_tree_sum(_p(v, ctypes.c_float), _p(c, ctypes.c_int32), v.shape[0]))
