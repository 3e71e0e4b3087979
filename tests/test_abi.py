"""The C ABI library: loads, exports every symbol include/shy.h declares, and
reports errors without a GPU (CPU only; no compute call is made)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in ("shy.h", "shy_nn.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(shy_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import simplehydrology_amd as shy
    L = C.CDLL(shy.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_default_params_are_the_reference_statics():
    import simplehydrology_amd as shy
    p = shy.Params()
    ref = dict(evapRate=0.001, depositionRate=0.1, minVol=0.01, maxAge=500.0, entrainment=10.0,
               gravity=1.0, momentumTransfer=1.0, lrate=0.1, maxdiff=0.01, settling=0.8)
    for k, v in ref.items():
        assert getattr(p, k) == pytest.approx(v, rel=1e-7), k


def test_struct_layouts_match_header():
    import simplehydrology_amd as shy
    assert C.sizeof(shy.Params) == 40
    assert C.sizeof(shy.DropStep) == 40
    assert C.sizeof(shy.DropState) == 28
    assert C.sizeof(shy.Stats) == 80
    assert C.sizeof(shy.MapHeader) == 72


def test_map_dump_format_roundtrip(tmp_path):
    """The dump format (header + AoS cells) written and read without a GPU."""
    import numpy as np
    import simplehydrology_amd as shy
    rng = np.random.default_rng(0)
    cells = rng.random((13, 7, 8), dtype=np.float32)
    p = shy.Params(evapRate=0.002)
    path = str(tmp_path / "m.shymap")
    shy.write_map(path, cells, p, seed=99, cycle=17)
    raw = open(path, "rb").read()
    assert raw[:8] == b"SHYMAP01" and len(raw) == 72 + cells.nbytes
    hd, back = shy.read_map(path)
    assert (hd.nx, hd.ny, hd.seed, hd.cycle) == (13, 7, 99, 17)
    assert hd.params.evapRate == np.float32(0.002)
    assert np.array_equal(back, cells)
    open(path, "wb").write(b"NOTAMAP!" + raw[8:])
    with pytest.raises(ValueError):
        shy.read_map(path)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is present")
def test_create_without_gpu_fails_cleanly():
    import simplehydrology_amd as shy
    assert shy.device_count() == 0
    with pytest.raises(shy.ShyError) as e:
        shy.World(8, 8)
    assert e.value.code == shy.SHY_ERR_NODEV
    assert "no GPU" in str(e.value)


def test_product_does_not_reference_the_oracle():
    pkg = os.path.join(ROOT, "simplehydrology_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f
                assert "shy_oracle" not in text, f
