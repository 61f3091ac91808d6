"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code:
the oracle (every entry point on 2D / 3D, periodic and open cases, with
leak checking) and the host layer's CPU paths (parameter.h and config.info
readers, HDF5 I/O, periodic ghost images).  Each build's driver exits
non-zero on a failed check, and the sanitizers abort on the first error."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(path, timeout):
    r = subprocess.run(["make", "-s", "-j8", "-C", path, "sanitize"], capture_output=True, text=True,
                       timeout=timeout)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    return out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan():
    out = _make(os.path.join(ROOT, "oracle"), 300)
    assert out.count("[sanitize]") == 6 and "FAILED" not in out, out


def test_host_layer_under_asan_ubsan():
    host = os.path.join(ROOT, "meshlesshydro_amd", "host")
    if not os.path.exists(os.path.join(ROOT, "meshlesshydro_amd", "libmlh_hip.so")):
        pytest.skip("libmlh_hip.so not built")
    if not os.path.exists("/opt/conda/include/hdf5.h"):
        pytest.skip("HDF5 headers not available")
    out = _make(host, 600)
    assert out.count("[sanitize]") == 4, out
