"""CPU: AddressSanitizer + UndefinedBehaviorSanitizer runs (host code only).

* The oracle: built with -fsanitize=address,undefined by `make -C oracle
  sanitize`, its driver (oracle/san_check.c) runs every routine on small and
  edge-case inputs against the O(n^2) product and the series identities.
* The library's host code: tests/cpp/build_san.sh builds libfpxg with the
  host pass sanitized (gfx950 device code as usual) and tests/cpp/san_driver;
  here, without a device, the driver runs the C ABI's error paths (on a GPU
  it runs products on every path, series with graph replay, batches, the
  server and the no-allocation product).  The build takes minutes, so this
  part runs only where it has been built.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan():
    if shutil.which(os.environ.get("CC", "gcc")) is None:
        pytest.skip("no C compiler")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "oracle_check")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ok" in r.stdout


def test_cabi_host_code_under_asan_ubsan():
    drv = os.path.join(ROOT, "tests", "cpp", "_san", "san_driver")
    if not os.path.exists(drv):
        pytest.skip("sanitized library not built (sh tests/cpp/build_san.sh)")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([drv], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert r.stdout.startswith("ok")
