"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
entry point include/par2hip.h declares, and reports no device on a GPU-less
host instead of failing."""
import ctypes as C
import os
import re

import pytest

import par2hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "par2hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(par2hip_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("par2hip_encode", "par2hip_reconstruct", "par2hip_gf_invert", "par2hip_hash_slices",
              "par2hip_gf_muladd_region", "par2hip_gf_muladd", "par2hip_gf_mul_slice", "par2hip_available"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(par2hip.LIB_PATH):
        pytest.skip("libpar2hip.so not built")
    lib = C.CDLL(par2hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_device_is_reported_not_raised():
    if not os.path.exists(par2hip.LIB_PATH):
        pytest.skip("libpar2hip.so not built")
    import torch  # noqa: F401  (same process state as the GPU runs)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = par2hip.lib()
    assert lib.par2hip_available() == 0
    assert lib.par2hip_strategy_name() == b"HIP-gfx950"
    assert lib.par2hip_strategy_priority() == 40
    h = C.c_void_p()
    assert lib.par2hip_ctx_create(0, C.byref(h)) == par2hip.ENODEV
    assert lib.par2hip_encode(None, None, 0, 0, None, 0, None, None, None) == par2hip.EINVAL
    srv = C.c_void_p()
    assert lib.par2hip_server_create(0, 1 << 30, 2, 64, C.byref(srv)) == par2hip.ENODEV
    assert lib.par2hip_server_create(0, 0, 2, 64, C.byref(srv)) == par2hip.EINVAL
    assert lib.par2hip_server_create(0, 1 << 30, 0, 64, C.byref(srv)) == par2hip.EINVAL
    assert not par2hip.HipStrategy().is_available()


def test_autotune_table_text_roundtrip(tmp_path):
    """The tuned-shape table loads from and saves to text without a device;
    unknown shapes and malformed lines are refused."""
    if not os.path.exists(par2hip.LIB_PATH):
        pytest.skip("libpar2hip.so not built")
    good = tmp_path / "tune.txt"
    good.write_text("# key v g d\n8 8 2 16\n10208 4 1 16\n")
    assert par2hip.autotune_load(str(good)) == 2
    out = tmp_path / "saved.txt"
    par2hip.autotune_save(str(out))
    lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert "8 8 2 16" in lines and "10208 4 1 16" in lines
    bad = tmp_path / "bad.txt"
    bad.write_text("9 8 3 16\n")  # no such shape
    with pytest.raises(par2hip.InvalidFormat):
        par2hip.autotune_load(str(bad))
    wide = tmp_path / "wide.txt"
    wide.write_text("4 8 2 16\n")  # wider than the key's alignment allows
    with pytest.raises(par2hip.InvalidFormat):
        par2hip.autotune_load(str(wide))
