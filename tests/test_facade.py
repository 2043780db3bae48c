"""The C++20 drop-in facade used as reference code uses fastpoly (GPU)."""
import ctypes
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "cpp", "libfacade_tests.so")


def test_facade_builds_against_header():
    assert os.path.exists(SO), "run tests/cpp/build.sh (part of __graft_entry__.build)"


@pytest.mark.gpu
def test_facade_on_device():
    lib = ctypes.CDLL(SO)
    lib.fpx_facade_tests.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    buf = ctypes.create_string_buffer(1 << 16)
    fails = lib.fpx_facade_tests(buf, len(buf))
    report = buf.value.decode()
    print(report)
    assert fails == 0, report
