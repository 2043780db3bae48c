"""CPU checks of the C ABI library: it loads without a GPU, exports every
entry point include/fpxg.h declares, and reports errors instead of faulting."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "fpxg.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(fpxg_\w+)\s*\(", text)))


def test_header_lists_match_python_mirror():
    import fastpoly_amd
    assert declared_symbols() == sorted(fastpoly_amd.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    import fastpoly_amd
    if not os.path.exists(fastpoly_amd.lib_path()):
        pytest.skip("libfpxg.so not built")
    L = fastpoly_amd.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_strerror_and_no_gpu_error_path():
    import fastpoly_amd
    if not os.path.exists(fastpoly_amd.lib_path()):
        pytest.skip("libfpxg.so not built")
    L = fastpoly_amd.lib()
    assert L.fpxg_strerror(1).startswith(b"domain")
    assert L.fpxg_strerror(3).startswith(b"device memory")
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = os.path.exists("/dev/kfd")
    if not has_gpu:
        assert L.fpxg_init(0, 0) == fastpoly_amd.E_DEVICE
