"""C-ABI library: builds, loads, exports every symbol include/*.h declares, and
its host-only helpers behave (no device calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("amgt_mbsr.h", "amgt_policy.h")]


def declared_symbols():
    assert sorted(HEADERS) == sorted(os.path.join(ROOT, "include", h)
                                     for h in os.listdir(os.path.join(ROOT, "include")))
    text = "\n".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(amgt_[a-z0-9_]+)\s*\(", text))
    # typedef'd callback types are not functions
    return sorted(n for n in names if not n.endswith("_fn"))


def test_library_is_built_and_exports_header_symbols():
    from amgt_amd import _lib

    assert os.path.exists(_lib.LIB_PATH), "run `python -m amgt_amd.build` first"
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == syms
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}$", out, flags=re.M), s


def test_host_only_helpers():
    from amgt_amd import _lib

    L = _lib.load()
    # get_calculationPrecison, csr_spgemm_device.c:1847-1862
    assert [L.amgt_precision_for_level(v) for v in (-1, 0, 1, 2, 7)] == [0, 0, 1, 2, 2]
    assert L.amgt_status_string(0) == b"ok"
    assert L.amgt_status_string(5) == b"invalid CSR input"
    # argument errors are reported without touching the device
    st = L.amgt_mbsr_create_from_csr(4, 4, 0, None, None, None, None, None)
    assert st == _lib.AMGT_ERR_ARG
    assert b"NULL" in L.amgt_last_error()
    with pytest.raises(_lib.AmgtError):
        _lib.check(st, "create")
    assert L.amgt_set_allocator(ctypes.c_void_p(1), None, None) == _lib.AMGT_ERR_ARG


def test_package_fails_loudly_without_library(tmp_path, monkeypatch):
    from amgt_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.AmgtLibraryMissing):
        _lib.load()


def test_trilinear_interp_rejects_bad_grids_before_device_work():
    """Argument errors return AMGT_ERR_ARG before any HIP call (runs without a GPU)."""
    from amgt_amd import _lib
    L = _lib.load()
    g = (ctypes.c_int32 * 3)(4, 0, 2)
    out = [ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()]
    ptrs = [ctypes.c_void_p() for _ in range(3)]
    rc = L.amgt_interp_trilinear(1, g, ctypes.byref(out[0]), ctypes.byref(out[1]),
                                 ctypes.byref(out[2]), *[ctypes.byref(p) for p in ptrs], None)
    assert rc == _lib.AMGT_ERR_ARG
    assert b"positive" in L.amgt_last_error()
    rc = L.amgt_interp_trilinear(0, g, ctypes.byref(out[0]), ctypes.byref(out[1]),
                                 ctypes.byref(out[2]), *[ctypes.byref(p) for p in ptrs], None)
    assert rc == _lib.AMGT_ERR_ARG


def test_wrapper_length_and_device_checks():
    """The MBSRMatrix wrappers refuse short or foreign tensors before any
    launch (the kernels would index past their ends)."""
    import torch

    from amgt_amd import mbsr as M

    cpu = torch.device("cpu")
    M._need_len(torch.zeros(5), "x", 5, cpu)
    with pytest.raises(ValueError, match="holds 4 entries, 5 needed"):
        M._need_len(torch.zeros(4), "x", 5, cpu)
    with pytest.raises(ValueError, match="is on"):
        M._need_len(torch.zeros(5, device="meta"), "x", 5, cpu)
    M._need_rows_len(torch.zeros(3, 7), "X", 3, 7, cpu)
    with pytest.raises(ValueError):
        M._need_rows_len(torch.zeros(3, 6), "X", 3, 7, cpu)
    with pytest.raises(ValueError):
        M._need_rows_len(torch.zeros(2, 7), "X", 3, 7, cpu)
    with pytest.raises(TypeError):
        M._need_cuda(torch.zeros(3), "x", torch.float64)
