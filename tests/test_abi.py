"""The C ABI library loads without a GPU and exports what include/*.h declares.

No compute call runs here: creating a world needs a device, and without one
the library must refuse with SFL_DEVICE_ERROR and a detail string.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from solunar_amd import fluid
from solunar_amd.types import CBatchTuning, CReport, CSettings, FluidSettings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(sfl_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(fluid.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_header():
    # sfl_settings: int32, 3 floats, u64, 2 floats, vec3, float, u32, 2 floats, 2 u32, int32
    assert ctypes.sizeof(CSettings) == 72
    assert CSettings.max_particles.offset == 16
    assert ctypes.sizeof(CReport) == 48
    # sfl_batch_tuning: 4 u32, u32[5], 3 x float[5]
    assert ctypes.sizeof(CBatchTuning) == 96
    assert CBatchTuning.sort_ms.offset == 36


# FluidSettings::default (src/fluid/mod.rs:80-95, FluidSolver::Wcsph is the
# #[default] at mod.rs:47) and the DFSPH constants (src/fluid/dfsph.rs:8-16),
# as the reference writes them; f32 fields compare as f32 bits
REFERENCE_DEFAULTS = {
    "solver": 0, "spacing": 0.05, "rest_density": 1000.0, "support": 2.0, "max_particles": 1 << 22,
    "sound_speed": 20.0, "viscosity": 0.15, "gravity": (0.0, -9.81, 0.0), "courant": 0.4, "max_substeps": 2000,
    "divergence_tolerance": 1e-3, "density_tolerance": 5e-4, "min_iterations": 2, "max_iterations": 100,
}


def test_defaults_are_the_reference_defaults():
    lib = fluid.library()
    c = CSettings()
    lib.sfl_settings_default(ctypes.byref(c))
    for name, want in REFERENCE_DEFAULTS.items():
        got = getattr(c, name)
        if name == "gravity":
            assert tuple(np.float32(x) for x in got) == tuple(np.float32(w) for w in want)
        elif isinstance(want, float):
            assert np.float32(got) == np.float32(want), name
        else:
            assert got == want, name
    assert c.device == 0
    # and the Python mirror agrees with the C defaults
    d = FluidSettings()
    for name in REFERENCE_DEFAULTS:
        if name != "gravity":
            assert np.float32(getattr(d, name)) == np.float32(getattr(c, name)), name


def test_status_texts_follow_physics_error_display():
    lib = fluid.library()
    assert lib.sfl_status_text(16) == b"fluid has no room for more particles"
    assert lib.sfl_status_text(11) == b"dispatcher did not run every job"


def test_without_a_device_creation_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = fluid.library()
    c = CSettings.of(FluidSettings())
    h = ctypes.c_void_p()
    st = lib.sfl_create(ctypes.byref(c), ctypes.byref(h))
    assert st == 100
    assert lib.sfl_error_detail(None)
