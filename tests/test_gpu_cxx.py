"""TAL-SH's own C++ wrapper running on libtalsh_amd.so and the GPU.

oracle/_ref/cxx_wrapper_driver is tests/cxx/wrapper_driver.cpp linked with the
reference's talshxx.cpp / talsh_task.cpp compiled unchanged against include/
(oracle/Makefile target cxx, built in the development container).  It runs a
contraction through talsh::Tensor::contractAccumulate against its closed form,
a slice / insert round trip through extractSlice / insertSlice, and
contractAccumulateXL (cut into pieces) against the closed forms of the
reference's own XL test."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "oracle", "_ref", "cxx_wrapper_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="C++ wrapper driver not built (make -C oracle cxx)")
def test_reference_cxx_wrapper_on_gpu():
    print("[case] cxx wrapper driver", flush=True)
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "cxx wrapper driver: OK" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
