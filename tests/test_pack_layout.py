"""CPU: the walk records' 2x2-block layout is a bijection from cells to
records (simplehydrology_amd/csrc/shy_device.hpp: pack_rec, pack_records),
with each 2x2 block in one 128-byte line and the pack kernel's record-order
numbering mapping back to the same cells; odd sides leave padding records
that no cell uses.  Compiled for the host with g++ from the same header the
kernels use."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "pack_layout_check.cpp")


def test_pack_layout_bijection(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    exe = str(tmp_path / "pack_layout_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-o", exe, SRC], check=True)
    sizes = [(x, y) for x in range(1, 10) for y in range(1, 10)]
    sizes += [(80, 115), (115, 80), (1024, 1024), (2048, 1), (1, 2047), (129, 257)]
    args = [str(v) for s in sizes for v in s]
    r = subprocess.run([exe, *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    cells = sum(x * y for x, y in sizes)
    assert r.stdout.split() == ["cells", str(cells), "mismatches", "0"]
