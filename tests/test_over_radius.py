"""The kernels divide by the radius in three operations from rinv =
RN(1 / radius) (`over_radius`, solunar_amd/csrc/device.hpp) instead of an
IEEE division: Markstein's correction, which rounds to the same float. The
reference divides (src/fluid/mod.rs:140-157), so every float the kernels
feed it must come out bit-equal to a / radius.

`over_radius_check.c` states the same three operations in C and compares
them with the C division for every float a with 2^-100 <= |a| <= 2^100 and
for both zeros (the kernels' operands lie between 1e-13 and 1e8), at the
radii of the workloads and at radii whose reciprocals round awkwardly (a
significand of all ones, a power of two just below). The check itself is
pinned by two mistakes that must fail it: the uncorrected product, and the
correction without the zero's sign.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent

RADII = ["0.04", "0.01", "0.004", "0.1", "1.9999999", "0.015624999", "3", "0.7"]


def build(tmp_path, source: str, name: str):
    c = tmp_path / f"{name}.c"
    c.write_text(source)
    exe = tmp_path / name
    subprocess.run(["gcc", "-O2", "-mfma", "-fopenmp", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"],
                   check=True)
    return exe


@pytest.fixture(scope="module")
def source():
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not installed")
    return (HERE / "over_radius_check.c").read_text()


def test_three_operation_division_equals_ieee_division(tmp_path, source):
    exe = build(tmp_path, source, "check")
    out = subprocess.run([str(exe), *RADII], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


@pytest.mark.parametrize("mistake", ["return q0;", "return q;"], ids=["uncorrected", "zero-sign"])
def test_the_check_fails_a_planted_mistake(tmp_path, source, mistake):
    planted = source.replace("return a == 0.0f ? q0 : q;", mistake)
    assert planted != source
    exe = build(tmp_path, planted, "planted")
    out = subprocess.run([str(exe), "0.01"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 1 and out.stdout.startswith("mismatch"), out.stdout
