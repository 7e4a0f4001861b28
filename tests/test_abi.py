"""The C ABI library loads without a GPU and exports what include/*.h declares.

No compute call runs here: creating a world needs a device, and without one
the library must refuse with SFL_DEVICE_ERROR and a detail string.
"""
import ctypes
import glob
import os
import re

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


def test_defaults_are_the_reference_defaults():
    lib = fluid.library()
    c = CSettings()
    lib.sfl_settings_default(ctypes.byref(c))
    d = FluidSettings()
    for name, _ in CSettings._fields_:
        got = getattr(c, name)
        want = getattr(d, name)
        if name == "gravity":
            assert tuple(got) == pytest.approx(want)
        elif isinstance(want, float):
            assert got == pytest.approx(want)
        else:
            assert got == want, name


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
