"""Plain data types of the fluid API: settings, report, errors.

Mirrors `FluidSettings` (src/fluid/mod.rs:55-95), `FluidReport`
(src/fluid/step.rs:28-49) and `PhysicsError` (src/error.rs:5-22) of the
reference, plus the DFSPH constants of src/fluid/dfsph.rs:8-16 made settable.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, fields
from typing import Tuple

WCSPH = 0
DFSPH = 1

# status codes of include/solunar_fluid.h (PhysicsError order, offset by one)
STATUS_NAMES = {
    0: "Ok",
    1: "InvalidHandle",
    2: "InvalidFrame",
    3: "FrameTooDeep",
    4: "NonFinite",
    5: "InvalidShape",
    6: "InvalidMaterial",
    7: "InvalidSettings",
    8: "NotDynamic",
    9: "NotKinematic",
    10: "OutOfRange",
    11: "DispatchIncomplete",
    12: "InvalidJoint",
    13: "BodyHasJoints",
    14: "ShapeInUse",
    15: "TerrainBusy",
    16: "FluidFull",
    100: "DeviceError",
    101: "Unsupported",
}
INVALID_HANDLE = 1
NON_FINITE = 4
INVALID_SHAPE = 5
INVALID_SETTINGS = 7
OUT_OF_RANGE = 10
DISPATCH_INCOMPLETE = 11
FLUID_FULL = 16
DEVICE_ERROR = 100
UNSUPPORTED = 101


SNAPSHOT_KINDS = {1: "TooLarge", 2: "Truncated", 3: "BadMagic", 4: "UnsupportedVersion", 5: "BadChecksum",
                  6: "Invalid"}


class SnapshotError(Exception):
    """SnapshotError (src/snapshot/mod.rs:56-63): `name` is the variant,
    `version` set for UnsupportedVersion, `what` for Invalid."""

    def __init__(self, kind: int, version: int = 0, what: str = ""):
        self.kind = kind
        self.name = SNAPSHOT_KINDS.get(kind, f"Kind{kind}")
        self.version = version
        self.what = what
        super().__init__(f"{self.name}" + (f" {version}" if kind == 4 else "") + (f": {what}" if what else ""))

    def __eq__(self, other):
        return (isinstance(other, SnapshotError) and (self.kind, self.version, self.what) ==
                (other.kind, other.version, other.what))

    def __hash__(self):
        return hash((self.kind, self.version, self.what))


class CSnapshotError(ctypes.Structure):
    """`sfl_snapshot_error` of include/solunar_fluid.h."""

    _fields_ = [("kind", ctypes.c_int32), ("version", ctypes.c_uint32), ("what", ctypes.c_char_p)]

    def error(self) -> SnapshotError:
        return SnapshotError(self.kind, self.version, self.what.decode() if self.what else "")


class PhysicsError(Exception):
    """A non-zero status; `name` is the PhysicsError variant."""

    def __init__(self, code: int, detail: str = ""):
        self.code = code
        self.name = STATUS_NAMES.get(code, f"Status{code}")
        self.detail = detail
        super().__init__(f"{self.name}" + (f"({detail})" if detail else ""))


@dataclass
class FluidSettings:
    solver: int = WCSPH
    spacing: float = 0.05
    rest_density: float = 1000.0
    support: float = 2.0
    max_particles: int = 1 << 22
    sound_speed: float = 20.0
    viscosity: float = 0.15
    gravity: Tuple[float, float, float] = (0.0, -9.81, 0.0)
    courant: float = 0.4
    max_substeps: int = 2000
    divergence_tolerance: float = 1e-3
    density_tolerance: float = 5e-4
    min_iterations: int = 2
    max_iterations: int = 100
    device: int = 0

    def radius(self) -> float:
        import numpy as np

        return float(np.float32(self.spacing) * np.float32(self.support))


@dataclass
class FluidReport:
    substeps: int = 0
    capped: bool = False
    max_compression: float = 0.0
    max_speed: float = 0.0
    held_back: int = 0
    iterations: int = 0
    emitted: int = 0
    drained: int = 0
    lost: int = 0
    sound_speed: float = 0.0
    iterations_divergence: int = 0
    iterations_density: int = 0


class CSettings(ctypes.Structure):
    """`sfl_settings` of include/solunar_fluid.h."""

    _fields_ = [
        ("solver", ctypes.c_int32),
        ("spacing", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("support", ctypes.c_float),
        ("max_particles", ctypes.c_uint64),
        ("sound_speed", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("courant", ctypes.c_float),
        ("max_substeps", ctypes.c_uint32),
        ("divergence_tolerance", ctypes.c_float),
        ("density_tolerance", ctypes.c_float),
        ("min_iterations", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("device", ctypes.c_int32),
    ]

    @classmethod
    def of(cls, s: FluidSettings) -> "CSettings":
        c = cls()
        for f in fields(s):
            v = getattr(s, f.name)
            if f.name == "gravity":
                c.gravity = (ctypes.c_float * 3)(*v)
            else:
                setattr(c, f.name, v)
        return c


class CReport(ctypes.Structure):
    """`sfl_report` of include/solunar_fluid.h."""

    _fields_ = [
        ("substeps", ctypes.c_uint32),
        ("capped", ctypes.c_uint32),
        ("max_compression", ctypes.c_float),
        ("max_speed", ctypes.c_float),
        ("held_back", ctypes.c_uint32),
        ("iterations", ctypes.c_uint32),
        ("emitted", ctypes.c_uint32),
        ("drained", ctypes.c_uint32),
        ("lost", ctypes.c_uint32),
        ("sound_speed", ctypes.c_float),
        ("iterations_divergence", ctypes.c_uint32),
        ("iterations_density", ctypes.c_uint32),
    ]

    def to_report(self) -> FluidReport:
        return FluidReport(
            substeps=self.substeps,
            capped=bool(self.capped),
            max_compression=self.max_compression,
            max_speed=self.max_speed,
            held_back=self.held_back,
            iterations=self.iterations,
            emitted=self.emitted,
            drained=self.drained,
            lost=self.lost,
            sound_speed=self.sound_speed,
            iterations_divergence=self.iterations_divergence,
            iterations_density=self.iterations_density,
        )


class CVec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]

    @classmethod
    def of(cls, v) -> "CVec3":
        return cls(float(v[0]), float(v[1]), float(v[2]))


BATCH_CANDIDATES = 5


class CBatchTuning(ctypes.Structure):
    """`sfl_batch_tuning` of include/solunar_fluid.h."""

    _fields_ = [
        ("sort_block", ctypes.c_uint32),
        ("density_block", ctypes.c_uint32),
        ("force_block", ctypes.c_uint32),
        ("candidates", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32 * BATCH_CANDIDATES),
        ("sort_ms", ctypes.c_float * BATCH_CANDIDATES),
        ("density_ms", ctypes.c_float * BATCH_CANDIDATES),
        ("force_ms", ctypes.c_float * BATCH_CANDIDATES),
    ]

    def to_dict(self) -> dict:
        n = self.candidates
        return {
            "sort_block": self.sort_block,
            "density_block": self.density_block,
            "force_block": self.force_block,
            "blocks": list(self.blocks[:n]),
            "sort_ms": [float(x) for x in self.sort_ms[:n]],
            "density_ms": [float(x) for x in self.density_ms[:n]],
            "force_ms": [float(x) for x in self.force_ms[:n]],
        }


# ---- coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) ----

CUBOID = 0
SPHERE = 1
CAPSULE = 2


@dataclass(frozen=True)
class Shape:
    """A shape a body couples with (src/shape.rs:8-49): a cuboid (a, b, c its
    half extents), a sphere (a its radius), a capsule along local y (a its
    radius, b the half height of its cylinder)."""

    kind: int
    a: float
    b: float = 0.0
    c: float = 0.0

    @staticmethod
    def cuboid(half_extents) -> "Shape":
        x, y, z = half_extents
        return Shape(CUBOID, float(x), float(y), float(z))

    @staticmethod
    def sphere(radius: float) -> "Shape":
        return Shape(SPHERE, float(radius))

    @staticmethod
    def capsule(radius: float, half_height: float) -> "Shape":
        return Shape(CAPSULE, float(radius), float(half_height))


class CShape(ctypes.Structure):
    """`sfl_shape` of include/solunar_fluid.h."""

    _fields_ = [("kind", ctypes.c_int32), ("a", ctypes.c_float), ("b", ctypes.c_float), ("c", ctypes.c_float)]

    @classmethod
    def of(cls, s: Shape) -> "CShape":
        return cls(s.kind, s.a, s.b, s.c)


@dataclass
class BodyPose:
    """Where the rigid world has a coupled body for a step (SubstepPose,
    src/substep.rs): `body` is its handle (index, generation), rotation a
    quaternion (x, y, z, w); `dynamic` says whether it takes pushes."""

    body: Tuple[int, int]
    position: Tuple[float, float, float]
    rotation: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 1.0)
    linear_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    angular_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    dynamic: bool = True


class CBodyPose(ctypes.Structure):
    """`sfl_body_pose` of include/solunar_fluid.h."""

    _fields_ = [
        ("index", ctypes.c_uint32),
        ("generation", ctypes.c_uint32),
        ("position", ctypes.c_double * 3),
        ("rotation", ctypes.c_float * 4),
        ("linear_velocity", ctypes.c_float * 3),
        ("angular_velocity", ctypes.c_float * 3),
        ("dynamic", ctypes.c_int32),
    ]


def _pose_dtype():
    import numpy as np

    return np.dtype({"names": ["index", "generation", "position", "rotation", "linear_velocity",
                               "angular_velocity", "dynamic"],
                     "formats": ["<u4", "<u4", ("<f8", 3), ("<f4", 4), ("<f4", 3), ("<f4", 3), "<i4"],
                     "offsets": [0, 4, 8, 32, 48, 60, 72], "itemsize": 80})


def _push_dtype():
    import numpy as np

    return np.dtype({"names": ["index", "generation", "lifted", "lift_impulse", "lift_angular", "impulse", "angular"],
                     "formats": ["<u4", "<u4", "<i4", ("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3)],
                     "offsets": [0, 4, 8, 12, 24, 36, 48], "itemsize": 60})


# numpy record layouts of sfl_body_pose and sfl_body_push, for callers that
# hand poses over and take pushes back as arrays
POSE_DTYPE = _pose_dtype()
PUSH_DTYPE = _push_dtype()


def pose_array(poses):
    """A ctypes array of CBodyPose from BodyPose values."""
    arr = (CBodyPose * max(len(poses), 1))()
    for k, p in enumerate(poses):
        c = arr[k]
        c.index, c.generation = int(p.body[0]), int(p.body[1])
        c.position[:] = [float(v) for v in p.position]
        c.rotation[:] = [float(v) for v in p.rotation]
        c.linear_velocity[:] = [float(v) for v in p.linear_velocity]
        c.angular_velocity[:] = [float(v) for v in p.angular_velocity]
        c.dynamic = 1 if p.dynamic else 0
    return arr


@dataclass
class BodyPush:
    """What the water gave a coupled body over a step, as impulses: the lift
    of src/fluid/lift.rs (set when `lifted`, the body being dynamic) and the
    damping of mover_push (src/fluid/couple.rs:368-408)."""

    body: Tuple[int, int]
    lifted: bool
    lift_impulse: Tuple[float, float, float]
    lift_angular: Tuple[float, float, float]
    impulse: Tuple[float, float, float]
    angular: Tuple[float, float, float]


class CBodyPush(ctypes.Structure):
    """`sfl_body_push` of include/solunar_fluid.h."""

    _fields_ = [
        ("index", ctypes.c_uint32),
        ("generation", ctypes.c_uint32),
        ("lifted", ctypes.c_int32),
        ("lift_impulse", ctypes.c_float * 3),
        ("lift_angular", ctypes.c_float * 3),
        ("impulse", ctypes.c_float * 3),
        ("angular", ctypes.c_float * 3),
    ]

    def to_push(self) -> BodyPush:
        return BodyPush((self.index, self.generation), bool(self.lifted), tuple(self.lift_impulse),
                        tuple(self.lift_angular), tuple(self.impulse), tuple(self.angular))


# ---- terrain ground (src/fluid/terrain.rs) ----

def terrain_samples(world, low, high, distance):
    """The distances add_terrain takes over the box: `distance` (a function
    of (N, 3) float32 points, or the array itself) at low + (x, y, z) step,
    step = spacing / 2 in f32, loops y, z, x (terrain.rs:287-295)."""
    import numpy as np

    nx, ny, nz = world.terrain_nodes(low, high)
    if callable(distance) and (nx * ny * nz > 1 << 24 or not np.isfinite(np.asarray(low, np.float64)).all()):
        return np.zeros(0, np.float32)  # the library refuses the box before it reads any distance
    if not callable(distance):
        return np.ascontiguousarray(np.asarray(distance, np.float32).reshape(-1))
    f = np.float32
    step = f(world.settings.spacing) / f(2.0)
    lo = np.asarray(low, np.float32)
    y, z, x = np.meshgrid(np.arange(ny, dtype=np.float32), np.arange(nz, dtype=np.float32),
                          np.arange(nx, dtype=np.float32), indexing="ij")
    at = np.stack([lo[0] + x.ravel() * step, lo[1] + y.ravel() * step, lo[2] + z.ravel() * step], axis=1)
    return np.ascontiguousarray(np.asarray(distance(at.astype(np.float32)), np.float32).reshape(-1))
