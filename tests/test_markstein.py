"""CPU: the walk's reciprocal-form divisions equal the IEEE divisions.

The count walk divides the gravity term by the volume through the volume's
reciprocal (div_rcp: water.h:95), and the replay divides the sediment by
1 - evapRate through its reciprocal (evap_sed_rcp: water.h:135), both by
Markstein's correction step.  The bit-exact GPU parity tests meet only the
operands their walks happen to produce; this check compiles the same
functions (simplehydrology_amd/csrc/shy_device.hpp) for the host, whose
fma/fmaf are correctly rounded like the GPU's, and compares them with the
plain division over whole float ranges.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "markstein_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    exe = str(tmp_path_factory.mktemp("markstein") / "markstein_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-ffp-contract=off", "-mfma", "-o", exe, SRC],
                   check=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    checked, bad = r.stdout.split()[-3], r.stdout.split()[-1]
    assert int(bad) == 0
    return int(checked)


def test_evaporation_quotient_every_float(checker):
    # every finite non-negative float, at the default rate, the ensemble's
    # range ends and one more
    assert run(checker, "sed", 0.001, 0.0005, 0.002, 0.0023) == 4 * 0x7F800000


def test_gravity_quotient_every_slope(checker):
    # every float of magnitude 2^-60 .. 2^61 (biased exponents 67..187: the
    # Markstein range), both signs, at volumes of the walk's sequence from
    # its first step to its last
    n = run(checker, "grav", 67, 187, 0.001, 0, 1, 2, 250, 500)
    assert n == 5 * 2 * 121 << 23
    run(checker, "grav", 100, 140, 0.0023, 1, 137, 499)


def test_gravity_quotient_outside_range(checker):
    # tiny and huge slopes and volumes take the plain division
    assert run(checker, "edge") == 14 * 2 * 15
