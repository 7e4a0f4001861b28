"""`FluidWorld` over the gfx950 library (libsolunar_fluid.so) through its C ABI.

The methods mirror the reference's `FluidWorld` (src/fluid/mod.rs:341-692,
src/fluid/step.rs:51-90, src/fluid/boundary.rs:35-120): same names, same
argument meaning, errors raised as `PhysicsError` carrying the reference
variant name. The reference's `dispatch` argument has no counterpart: the
device has one schedule.

There is no CPU fallback: if the library is missing this module raises on
first use.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import sys
import weakref
from typing import Optional, Sequence, Tuple

import numpy as np

from .types import (POSE_DTYPE, PUSH_DTYPE, CBatchTuning, CBodyPose, CBodyPush, CReport, CSettings, CShape,
                    CSnapshotError, CVec3, FluidReport, FluidSettings, PhysicsError, Shape, pose_array,
                    terrain_samples)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsolunar_fluid.so")

_lib = None

PHASES = ("bounds", "sort", "reorder", "density", "force", "solve", "viscous", "advance", "flows", "bodies")


SNAPSHOT_REFUSED = 102  # SFL_SNAPSHOT_REFUSED

# Every live handle, closed by an exit hook before the interpreter tears down:
# freeing device memory after the HIP runtime (or a profiler attached to it)
# has finalised crashes the process at exit. Tapes go first (they read their
# world's arrays), then batches, then worlds.
_live = {"tape": weakref.WeakSet(), "batch": weakref.WeakSet(), "world": weakref.WeakSet()}


@atexit.register
def close_all() -> None:
    for kind in ("tape", "batch", "world"):
        for obj in list(_live[kind]):
            try:
                obj.close()
            except Exception:
                pass


def _release(obj, destroy) -> None:
    """Frees obj's handle once; skipped while the interpreter finalises,
    when the exit hook has already run and the runtime may be gone."""
    h = getattr(obj, "_h", None)
    if not h:
        return
    obj._h = None
    if sys.is_finalizing():
        return
    destroy(h)


def settings_of(c: CSettings) -> FluidSettings:
    """FluidSettings from an `sfl_settings`."""
    kw = {}
    for f in FluidSettings.__dataclass_fields__:
        v = getattr(c, f)
        kw[f] = tuple(float(x) for x in v) if f == "gravity" else v
    return FluidSettings(**kw)


def library() -> ctypes.CDLL:
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32p = ctypes.POINTER(ctypes.c_uint32)
    fp = ctypes.POINTER(ctypes.c_float)
    sz = ctypes.c_size_t
    i32 = ctypes.c_int32
    sigs = {
        "sfl_settings_default": (None, [ctypes.POINTER(CSettings)]),
        "sfl_create": (i32, [ctypes.POINTER(CSettings), ctypes.POINTER(P)]),
        "sfl_destroy": (None, [P]),
        "sfl_get_settings": (i32, [P, ctypes.POINTER(CSettings)]),
        "sfl_particle_mass": (ctypes.c_float, [P]),
        "sfl_len": (sz, [P]),
        "sfl_add_particles": (i32, [P, sz, fp, fp, u32p]),
        "sfl_fill_box": (i32, [P, CVec3, CVec3, ctypes.POINTER(sz)]),
        "sfl_fill_box_jittered": (i32, [P, CVec3, CVec3, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(sz)]),
        "sfl_set_domain": (i32, [P, ctypes.POINTER(CVec3), ctypes.POINTER(CVec3)]),
        "sfl_get_domain": (i32, [P, ctypes.POINTER(i32), ctypes.POINTER(CVec3), ctypes.POINTER(CVec3)]),
        "sfl_kernel_value": (ctypes.c_float, [P, ctypes.c_float]),
        "sfl_kernel_slope": (ctypes.c_float, [P, ctypes.c_float]),
        "sfl_kernel_gradient": (i32, [P, CVec3, ctypes.POINTER(CVec3)]),
        "sfl_neighbors_of": (i32, [P, sz, sz, u32p, ctypes.POINTER(sz)]),
        "sfl_add_emitter": (i32, [P, CVec3, CVec3, CVec3, u32p]),
        "sfl_add_drain": (i32, [P, CVec3, CVec3, u32p]),
        "sfl_remove_flow": (i32, [P, ctypes.c_uint32, ctypes.POINTER(i32)]),
        "sfl_save_snapshot": (i32, [P, ctypes.POINTER(ctypes.c_uint8), sz, ctypes.POINTER(sz)]),
        "sfl_load_snapshot": (i32, [ctypes.POINTER(ctypes.c_uint8), sz, ctypes.POINTER(CSettings), ctypes.POINTER(P),
                                    ctypes.POINTER(CSnapshotError)]),
        "sfl_add_box_container": (i32, [P, CVec3, CVec3, u32p]),
        "sfl_add_wall": (i32, [P, CVec3, CVec3, u32p]),
        "sfl_add_boundary_points": (i32, [P, sz, fp, u32p]),
        "sfl_remove_boundary": (i32, [P, ctypes.c_uint32, ctypes.POINTER(i32)]),
        "sfl_wall_len": (sz, [P]),
        "sfl_read_wall_positions": (i32, [P, sz, fp]),
        "sfl_step": (i32, [P, ctypes.c_float, ctypes.POINTER(CReport)]),
        "sfl_step_no_alloc": (i32, [P, ctypes.c_float, ctypes.POINTER(CReport)]),
        "sfl_terrain_nodes": (i32, [P, CVec3, CVec3, ctypes.POINTER(ctypes.c_uint64)]),
        "sfl_add_terrain": (i32, [P, CVec3, CVec3, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t,
                                  ctypes.POINTER(ctypes.c_uint32)]),
        "sfl_terrain_len": (ctypes.c_size_t, [P]),
        "sfl_couple": (i32, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(CShape)]),
        "sfl_uncouple": (i32, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(i32)]),
        "sfl_coupled_len": (ctypes.c_size_t, [P]),
        "sfl_step_with_bodies": (i32, [P, ctypes.c_float, ctypes.POINTER(CBodyPose), ctypes.c_size_t,
                                       ctypes.POINTER(CReport), ctypes.POINTER(CBodyPush), ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
        "sfl_update_neighbors": (i32, [P]),
        "sfl_compute_densities": (i32, [P]),
        "sfl_read_ids": (i32, [P, sz, u32p]),
        "sfl_read_positions": (i32, [P, sz, fp]),
        "sfl_read_velocities": (i32, [P, sz, fp]),
        "sfl_read_densities": (i32, [P, sz, fp]),
        "sfl_keep_forces": (i32, [P, i32]),
        "sfl_read_forces": (i32, [P, sz, fp]),
        "sfl_read_neighbors": (i32, [P, i32, sz, u32p, sz, u32p, ctypes.POINTER(sz)]),
        "sfl_state_hash": (i32, [P, ctypes.POINTER(ctypes.c_uint64)]),
        "sfl_error_detail": (ctypes.c_char_p, [P]),
        "sfl_status_text": (ctypes.c_char_p, [i32]),
        "sfl_profile": (i32, [P, i32]),
        "sfl_profile_read": (i32, [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64), i32]),
        "sfl_work_read": (i32, [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                ctypes.POINTER(ctypes.c_uint64), i32]),
        "sfl_synchronize": (i32, [P]),
        "sfl_solve_grid": (i32, [P]),
        "sfl_memory_bytes": (sz, [P]),
        "sfl_read_solve_log": (i32, [P, sz, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.POINTER(sz)]),
        "sfl_batch_create": (i32, [ctypes.POINTER(P), sz, ctypes.POINTER(P)]),
        "sfl_batch_destroy": (None, [P]),
        "sfl_batch_scenes": (sz, [P]),
        "sfl_batch_len": (sz, [P, sz]),
        "sfl_batch_step": (i32, [P, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(CReport)]),
        "sfl_batch_scene_failed": (i32, [P, sz, ctypes.POINTER(i32)]),
        "sfl_batch_prepare": (i32, [P, ctypes.c_uint32]),
        "sfl_batch_autotune": (i32, [P, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(CBatchTuning)]),
        "sfl_batch_set_tuning": (i32, [P, ctypes.POINTER(CBatchTuning)]),
        "sfl_batch_get_tuning": (i32, [P, ctypes.POINTER(CBatchTuning)]),
        "sfl_batch_use_graph": (i32, [P, i32]),
        "sfl_batch_read_state": (i32, [P, sz, sz, u32p, fp, fp, fp]),
        "sfl_batch_state_hash": (i32, [P, sz, ctypes.POINTER(ctypes.c_uint64)]),
        "sfl_batch_store": (i32, [P, ctypes.POINTER(P), sz]),
        "sfl_batch_error_detail": (ctypes.c_char_p, [P]),
        "sfl_batch_memory_bytes": (sz, [P]),
        "sfl_batch_graph_nodes": (ctypes.c_uint64, [P]),
        "sfl_batch_synchronize": (i32, [P]),
        "sfl_wcsph_record": (i32, [P, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(P)]),
        "sfl_wcsph_adjoint": (i32, [P, fp, fp, fp, fp, fp]),
        "sfl_tape_destroy": (None, [P]),
        "sfl_tape_substeps": (ctypes.c_uint64, [P]),
        "sfl_tape_memory_bytes": (sz, [P]),
        "sfl_tape_error_detail": (ctypes.c_char_p, [P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _f32(a, cols=3) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, cols))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class FluidWorld:
    """The reference's FluidWorld with its particle state in HBM."""

    def __init__(self, settings: Optional[FluidSettings] = None):
        self._lib = library()
        self._settings = settings or FluidSettings()
        handle = ctypes.c_void_p()
        st = self._lib.sfl_create(ctypes.byref(CSettings.of(self._settings)), ctypes.byref(handle))
        if st:
            raise PhysicsError(st, self._lib.sfl_error_detail(None).decode())
        self._h = handle
        _live["world"].add(self)

    def close(self) -> None:
        _release(self, self._lib.sfl_destroy)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st:
            raise PhysicsError(st, self._lib.sfl_error_detail(self._h).decode())

    # ---- accessors (mod.rs:372-408) ----
    @property
    def settings(self) -> FluidSettings:
        return self._settings

    def particle_mass(self) -> float:
        return float(self._lib.sfl_particle_mass(self._h))

    def __len__(self) -> int:
        return int(self._lib.sfl_len(self._h))

    def len(self) -> int:
        return len(self)

    def is_empty(self) -> bool:
        return len(self) == 0

    # ---- particles ----
    def add_particle(self, position: Sequence[float], velocity: Sequence[float] = (0.0, 0.0, 0.0)) -> int:
        ids = self.add_particles([position], [velocity])
        return int(ids[0])

    def add_particles(self, positions, velocities=None) -> np.ndarray:
        p = _f32(positions)
        v = None if velocities is None else _f32(velocities)
        n = p.shape[0]
        ids = np.zeros(n, dtype=np.uint32)
        st = self._lib.sfl_add_particles(self._h, n, _ptr(p, ctypes.c_float),
                                         None if v is None else _ptr(v, ctypes.c_float),
                                         _ptr(ids, ctypes.c_uint32))
        self._check(st)
        return ids

    def fill_box(self, low, high) -> int:
        count = ctypes.c_size_t()
        self._check(self._lib.sfl_fill_box(self._h, CVec3.of(low), CVec3.of(high), ctypes.byref(count)))
        return count.value

    def fill_box_jittered(self, low, high, jitter: float, seed: int) -> int:
        count = ctypes.c_size_t()
        self._check(self._lib.sfl_fill_box_jittered(self._h, CVec3.of(low), CVec3.of(high), jitter, seed,
                                                    ctypes.byref(count)))
        return count.value

    def set_domain(self, domain: Optional[Tuple[Sequence[float], Sequence[float]]]) -> None:
        if domain is None:
            self._check(self._lib.sfl_set_domain(self._h, None, None))
        else:
            lo, hi = CVec3.of(domain[0]), CVec3.of(domain[1])
            self._check(self._lib.sfl_set_domain(self._h, ctypes.byref(lo), ctypes.byref(hi)))

    def domain(self):
        """FluidWorld::domain (mod.rs:446-448): (low, high) or None."""
        has, lo, hi = ctypes.c_int32(), CVec3(), CVec3()
        self._check(self._lib.sfl_get_domain(self._h, ctypes.byref(has), ctypes.byref(lo), ctypes.byref(hi)))
        return ((lo.x, lo.y, lo.z), (hi.x, hi.y, hi.z)) if has.value else None

    # ---- the kernel (mod.rs:126-165, 376-378) ----
    def kernel_value(self, distance: float) -> float:
        return float(self._lib.sfl_kernel_value(self._h, distance))

    def kernel_slope(self, distance: float) -> float:
        return float(self._lib.sfl_kernel_slope(self._h, distance))

    def kernel_gradient(self, offset) -> Tuple[float, float, float]:
        out = CVec3()
        self._check(self._lib.sfl_kernel_gradient(self._h, CVec3.of(offset), ctypes.byref(out)))
        return (out.x, out.y, out.z)

    # ---- flows (flow.rs) ----
    def add_emitter(self, low, high, velocity) -> int:
        out = ctypes.c_uint32()
        self._check(self._lib.sfl_add_emitter(self._h, CVec3.of(low), CVec3.of(high), CVec3.of(velocity),
                                              ctypes.byref(out)))
        return out.value

    def add_drain(self, low, high) -> int:
        out = ctypes.c_uint32()
        self._check(self._lib.sfl_add_drain(self._h, CVec3.of(low), CVec3.of(high), ctypes.byref(out)))
        return out.value

    def remove_flow(self, flow_id: int) -> bool:
        out = ctypes.c_int32()
        self._check(self._lib.sfl_remove_flow(self._h, flow_id, ctypes.byref(out)))
        return out.value != 0

    # ---- snapshots (save.rs) ----
    def save_snapshot(self) -> bytes:
        """The SLFL version 3 bytes of the world (src/fluid/save.rs:17-66)."""
        n = ctypes.c_size_t()
        self._check(self._lib.sfl_save_snapshot(self._h, None, 0, ctypes.byref(n)))
        out = bytearray(n.value)
        buf = (ctypes.c_uint8 * n.value).from_buffer(out) if n.value else None
        self._check(self._lib.sfl_save_snapshot(self._h, buf, n.value, ctypes.byref(n)))
        del buf
        return bytes(out)

    @classmethod
    def load_snapshot(cls, data: bytes, settings: Optional[FluidSettings] = None) -> "FluidWorld":
        """A new world from snapshot bytes (src/fluid/save.rs:68-195); raises
        SnapshotError for refused bytes. `settings` gives what the snapshot
        does not hold (the extensions and the device)."""
        lib = library()
        data = bytes(data)
        buf = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data + b"\0")
        base = CSettings.of(settings) if settings is not None else None
        handle = ctypes.c_void_p()
        err = CSnapshotError()
        st = lib.sfl_load_snapshot(buf, len(data), ctypes.byref(base) if base is not None else None,
                                   ctypes.byref(handle), ctypes.byref(err))
        if st == SNAPSHOT_REFUSED:
            raise err.error()
        if st:
            raise PhysicsError(st, lib.sfl_error_detail(None).decode())
        w = cls.__new__(cls)
        w._lib = lib
        w._h = handle
        _live["world"].add(w)
        c = CSettings()
        lib.sfl_get_settings(handle, ctypes.byref(c))
        w._settings = settings_of(c)
        return w

    # ---- walls (boundary.rs) ----
    def add_box_container(self, low, high) -> int:
        out = ctypes.c_uint32()
        self._check(self._lib.sfl_add_box_container(self._h, CVec3.of(low), CVec3.of(high), ctypes.byref(out)))
        return out.value

    def add_wall(self, low, high) -> int:
        out = ctypes.c_uint32()
        self._check(self._lib.sfl_add_wall(self._h, CVec3.of(low), CVec3.of(high), ctypes.byref(out)))
        return out.value

    def add_boundary_points(self, points) -> int:
        p = _f32(points)
        out = ctypes.c_uint32()
        self._check(self._lib.sfl_add_boundary_points(self._h, p.shape[0], _ptr(p, ctypes.c_float),
                                                      ctypes.byref(out)))
        return out.value

    def remove_boundary(self, boundary: int) -> bool:
        out = ctypes.c_int32()
        self._check(self._lib.sfl_remove_boundary(self._h, boundary, ctypes.byref(out)))
        return bool(out.value)

    def wall_len(self) -> int:
        return int(self._lib.sfl_wall_len(self._h))

    def wall_positions(self) -> np.ndarray:
        n = int(self._lib.sfl_wall_len(self._h))
        out = np.zeros((n, 3), dtype=np.float32)
        self._check(self._lib.sfl_read_wall_positions(self._h, n, _ptr(out, ctypes.c_float)))
        return out

    # ---- stepping ----
    def step(self, dt: float) -> FluidReport:
        rep = CReport()
        self._check(self._lib.sfl_step(self._h, dt, ctypes.byref(rep)))
        return rep.to_report()

    def step_no_alloc(self, dt: float) -> FluidReport:
        """`step` without device allocation (sfl_step_no_alloc): WCSPH
        worlds, scratch shared by every such call on the GPU."""
        rep = CReport()
        self._check(self._lib.sfl_step_no_alloc(self._h, dt, ctypes.byref(rep)))
        return rep.to_report()

    # ---- terrain ground (terrain.rs) ----
    def terrain_nodes(self, low, high):
        """The lattice add_terrain samples over the box: nodes per axis (x, y, z)."""
        n = (ctypes.c_uint64 * 3)()
        self._check(self._lib.sfl_terrain_nodes(self._h, CVec3.of(low), CVec3.of(high), n))
        return tuple(int(v) for v in n)

    def add_terrain(self, low, high, distance) -> int:
        """add_terrain (src/fluid/terrain.rs:262-305): `distance` is the
        terrain's signed distance (positive above it), a function of (N, 3)
        float32 points or the sampled array."""
        d = terrain_samples(self, low, high, distance)
        out = ctypes.c_uint32()
        self._check(self._lib.sfl_add_terrain(self._h, CVec3.of(low), CVec3.of(high), _ptr(d, ctypes.c_float),
                                              d.size, ctypes.byref(out)))
        return out.value

    def terrain_len(self) -> int:
        return int(self._lib.sfl_terrain_len(self._h))

    # ---- coupled bodies (couple.rs, lift.rs) ----
    def couple(self, body, shape: Shape) -> None:
        """Lets water and `body` (index, generation) push each other in
        step_with_bodies (src/fluid/couple.rs:165-198)."""
        c = CShape.of(shape)
        self._check(self._lib.sfl_couple(self._h, int(body[0]), int(body[1]), ctypes.byref(c)))

    def uncouple(self, body) -> bool:
        out = ctypes.c_int32()
        self._check(self._lib.sfl_uncouple(self._h, int(body[0]), int(body[1]), ctypes.byref(out)))
        return out.value != 0

    def coupled_len(self) -> int:
        return int(self._lib.sfl_coupled_len(self._h))

    def step_with_bodies(self, dt: float, poses):
        """step_with_bodies (src/fluid/couple.rs:207-252) with the bodies
        where `poses` has them: (FluidReport, pushes) with each coupled body's
        lift and damping impulses, in handle order. Poses as BodyPose values
        give BodyPush values back; poses as a numpy array of POSE_DTYPE give a
        numpy array of PUSH_DTYPE back, with no per-body Python objects."""
        cap = self.coupled_len()
        n = ctypes.c_size_t()
        rep = CReport()
        if isinstance(poses, np.ndarray):
            if poses.dtype != POSE_DTYPE:
                raise TypeError("poses as an array need POSE_DTYPE")
            arr = np.ascontiguousarray(poses)
            out = np.zeros(max(cap, 1), PUSH_DTYPE)
            self._check(self._lib.sfl_step_with_bodies(
                self._h, dt, arr.ctypes.data_as(ctypes.POINTER(CBodyPose)), len(arr), ctypes.byref(rep),
                out.ctypes.data_as(ctypes.POINTER(CBodyPush)), cap, ctypes.byref(n)))
            return rep.to_report(), out[: n.value]
        arr = pose_array(poses)
        out = (CBodyPush * max(cap, 1))()
        self._check(self._lib.sfl_step_with_bodies(self._h, dt, arr, len(poses), ctypes.byref(rep), out, cap,
                                                   ctypes.byref(n)))
        return rep.to_report(), [out[k].to_push() for k in range(n.value)]

    def update_neighbors(self) -> None:
        self._check(self._lib.sfl_update_neighbors(self._h))

    def compute_densities(self) -> None:
        self._check(self._lib.sfl_compute_densities(self._h))

    # ---- readback ----
    def ids(self) -> np.ndarray:
        n = len(self)
        out = np.zeros(n, dtype=np.uint32)
        self._check(self._lib.sfl_read_ids(self._h, n, _ptr(out, ctypes.c_uint32)))
        return out

    def _vectors(self, fn) -> np.ndarray:
        n = len(self)
        out = np.zeros((n, 3), dtype=np.float32)
        self._check(fn(self._h, n, _ptr(out, ctypes.c_float)))
        return out

    def positions(self) -> np.ndarray:
        return self._vectors(self._lib.sfl_read_positions)

    def velocities(self) -> np.ndarray:
        return self._vectors(self._lib.sfl_read_velocities)

    def densities(self) -> np.ndarray:
        n = len(self)
        out = np.zeros(n, dtype=np.float32)
        self._check(self._lib.sfl_read_densities(self._h, n, _ptr(out, ctypes.c_float)))
        return out

    def keep_forces(self, keep: bool = True) -> None:
        self._check(self._lib.sfl_keep_forces(self._h, 1 if keep else 0))

    def forces(self) -> np.ndarray:
        return self._vectors(self._lib.sfl_read_forces)

    def _lists(self, kind: int) -> Tuple[np.ndarray, np.ndarray]:
        n = len(self)
        start = np.zeros(n + 1, dtype=np.uint32)
        count = ctypes.c_size_t()
        st = self._lib.sfl_read_neighbors(self._h, kind, n + 1, _ptr(start, ctypes.c_uint32), 0, None,
                                          ctypes.byref(count))
        if st not in (0, 10):
            self._check(st)
        items = np.zeros(max(count.value, 1), dtype=np.uint32)
        self._check(self._lib.sfl_read_neighbors(self._h, kind, n + 1, _ptr(start, ctypes.c_uint32),
                                                 items.shape[0], _ptr(items, ctypes.c_uint32),
                                                 ctypes.byref(count)))
        return start, items[: count.value]

    def neighbors(self) -> Tuple[np.ndarray, np.ndarray]:
        """CSR (start, items) of the fluid lists of the last build."""
        return self._lists(0)

    def wall_neighbors(self) -> Tuple[np.ndarray, np.ndarray]:
        return self._lists(1)

    def neighbors_of(self, index: int) -> np.ndarray:
        """neighbors_of (mod.rs:612-614): one particle's fluid list."""
        n = ctypes.c_size_t()
        st = self._lib.sfl_neighbors_of(self._h, index, 0, None, ctypes.byref(n))
        if st and n.value == 0:
            self._check(st)
        out = np.zeros(n.value, np.uint32)
        self._check(self._lib.sfl_neighbors_of(self._h, index, n.value, _ptr(out, ctypes.c_uint32), ctypes.byref(n)))
        return out

    def state_hash(self) -> int:
        out = ctypes.c_uint64()
        self._check(self._lib.sfl_state_hash(self._h, ctypes.byref(out)))
        return out.value

    # ---- measurement ----
    def profile(self, enable: bool) -> None:
        self._check(self._lib.sfl_profile(self._h, 1 if enable else 0))

    def profile_read(self, reset: bool = False):
        ms = (ctypes.c_double * len(PHASES))()
        launches = (ctypes.c_uint64 * len(PHASES))()
        self._check(self._lib.sfl_profile_read(self._h, ms, launches, 1 if reset else 0))
        return {p: (ms[i], launches[i]) for i, p in enumerate(PHASES)}

    def work_read(self, reset: bool = False):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._lib.sfl_work_read(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                            1 if reset else 0))
        return {"particle_substeps": a.value, "particle_pushes": b.value, "particle_rates": c.value}

    def memory_bytes(self) -> int:
        return int(self._lib.sfl_memory_bytes(self._h))

    def solve_log(self):
        """The last step's DFSPH stop tests: signed averages (+ divergence,
        - density) and the particles each averaged over."""
        n = ctypes.c_size_t()
        self._check(self._lib.sfl_read_solve_log(self._h, 0, None, None, ctypes.byref(n)))
        avg = np.zeros(n.value, np.float64)
        loaded = np.zeros(n.value, np.uint32)
        if n.value:
            self._check(self._lib.sfl_read_solve_log(self._h, n.value, avg.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     loaded.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                     ctypes.byref(n)))
        return avg, loaded

    def solve_grid(self) -> int:
        return int(self._lib.sfl_solve_grid(self._h))

    def synchronize(self) -> None:
        self._check(self._lib.sfl_synchronize(self._h))


class Tape:
    """A recorded WCSPH rollout of a world and its reverse pass
    (include/solunar_fluid.h, sfl_wcsph_record / sfl_wcsph_adjoint)."""

    def __init__(self, world: "FluidWorld", dt: float, steps: int, every: int = 20):
        self._lib = library()
        self.world = world  # the tape reads the world's arrays; keep it alive
        h = ctypes.c_void_p()
        st = self._lib.sfl_wcsph_record(world._h, dt, steps, every, ctypes.byref(h))
        if st:
            raise PhysicsError(st, self._lib.sfl_error_detail(world._h).decode())
        self._h = h
        self.n = len(world)
        _live["tape"].add(self)

    def close(self) -> None:
        _release(self, self._lib.sfl_tape_destroy)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def adjoint(self, xbar, vbar=None):
        """dL/dx, dL/dv of the final state (per id, (n, 3)) -> dL/dx0,
        dL/dv0 per id and dL/d(viscosity)."""
        xb = _f32(xbar)
        vb = _f32(np.zeros_like(xb) if vbar is None else vbar)
        if len(xb) != self.n or len(vb) != self.n:
            raise ValueError("adjoints need one row per particle")
        x0 = np.zeros((self.n, 3), np.float32)
        v0 = np.zeros((self.n, 3), np.float32)
        alpha = ctypes.c_float()
        st = self._lib.sfl_wcsph_adjoint(self._h, _ptr(xb, ctypes.c_float), _ptr(vb, ctypes.c_float),
                                         _ptr(x0, ctypes.c_float), _ptr(v0, ctypes.c_float), ctypes.byref(alpha))
        if st:
            raise PhysicsError(st, self._lib.sfl_tape_error_detail(self._h).decode())
        return x0, v0, float(alpha.value)

    def substeps(self) -> int:
        return int(self._lib.sfl_tape_substeps(self._h))

    def memory_bytes(self) -> int:
        return int(self._lib.sfl_tape_memory_bytes(self._h))


class FluidBatch:
    """Many WCSPH worlds stepped together (include/solunar_fluid.h,
    sfl_batch_*). Scene s starts as a copy of worlds[s] and steps bit-equal
    to it; `store` hands the scenes back to worlds."""

    def __init__(self, worlds: Sequence[FluidWorld]):
        self._lib = library()
        handles = (ctypes.c_void_p * len(worlds))(*[w._h for w in worlds])
        h = ctypes.c_void_p()
        st = self._lib.sfl_batch_create(handles, len(worlds), ctypes.byref(h))
        if st:
            raise PhysicsError(st, self._lib.sfl_batch_error_detail(None).decode())
        self._h = h
        _live["batch"].add(self)

    def close(self) -> None:
        _release(self, self._lib.sfl_batch_destroy)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st:
            raise PhysicsError(st, self._lib.sfl_batch_error_detail(self._h).decode())

    def scenes(self) -> int:
        return int(self._lib.sfl_batch_scenes(self._h))

    def len(self, scene: int) -> int:
        return int(self._lib.sfl_batch_len(self._h, scene))

    def step(self, dt: float, steps: int = 1):
        """`steps` steps of every scene; the last step's report of each."""
        n = self.scenes()
        reports = (CReport * n)()
        self._check(self._lib.sfl_batch_step(self._h, dt, steps, reports))
        return [r.to_report() for r in reports]

    def prepare(self, steps: int) -> None:
        """Capture the graph of a `steps`-step call ahead of it."""
        self._check(self._lib.sfl_batch_prepare(self._h, steps))

    def scene_failed(self, scene: int) -> bool:
        out = ctypes.c_int32()
        self._check(self._lib.sfl_batch_scene_failed(self._h, scene, ctypes.byref(out)))
        return bool(out.value)

    def autotune(self, dt: float, repeats: int = 3) -> dict:
        t = CBatchTuning()
        self._check(self._lib.sfl_batch_autotune(self._h, dt, repeats, ctypes.byref(t)))
        return t.to_dict()

    def tuning(self) -> dict:
        t = CBatchTuning()
        self._check(self._lib.sfl_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.to_dict()

    def set_tuning(self, sort_block: int, density_block: int, force_block: int) -> None:
        t = CBatchTuning()
        t.sort_block, t.density_block, t.force_block = sort_block, density_block, force_block
        self._check(self._lib.sfl_batch_set_tuning(self._h, ctypes.byref(t)))

    def use_graph(self, on: bool) -> None:
        self._check(self._lib.sfl_batch_use_graph(self._h, 1 if on else 0))

    def state(self, scene: int):
        """(ids, positions, velocities, densities) of a scene, current order."""
        n = self.len(scene)
        ids = np.zeros(n, np.uint32)
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        self._check(self._lib.sfl_batch_read_state(self._h, scene, n, _ptr(ids, ctypes.c_uint32),
                                                   _ptr(x, ctypes.c_float), _ptr(v, ctypes.c_float),
                                                   _ptr(rho, ctypes.c_float)))
        return ids, x, v, rho

    def state_hash(self, scene: int) -> int:
        out = ctypes.c_uint64()
        self._check(self._lib.sfl_batch_state_hash(self._h, scene, ctypes.byref(out)))
        return int(out.value)

    def store(self, worlds: Sequence[FluidWorld]) -> None:
        handles = (ctypes.c_void_p * len(worlds))(*[w._h for w in worlds])
        self._check(self._lib.sfl_batch_store(self._h, handles, len(worlds)))

    def memory_bytes(self) -> int:
        return int(self._lib.sfl_batch_memory_bytes(self._h))

    def graph_nodes(self) -> int:
        return int(self._lib.sfl_batch_graph_nodes(self._h))

    def synchronize(self) -> None:
        self._check(self._lib.sfl_batch_synchronize(self._h))
