"""The reference's C++ wrapper (talsh::Tensor / talsh::TensorTask, talshxx.hpp,
talshxx.cpp, talsh_task.cpp) compiles unchanged against include/ and links
against libtalsh_amd.so with no undefined symbol (oracle/Makefile target cxx).
Needs the reference checkout, so it only runs in the development container."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "talshxx.cpp")), reason="reference checkout absent")
def test_reference_cxx_wrapper_links_against_library():
    out = os.path.join(REPO, "oracle", "_ref", "libtalshxx_check.so")
    r = subprocess.run(["make", "-B", "-C", os.path.join(REPO, "oracle"), "cxx"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert os.path.exists(out)
    syms = subprocess.run(["nm", "-D", "--undefined-only", out], capture_output=True, text=True).stdout
    # the wrapper's calls into the C API are resolved by the product library
    for name in ("talshTensorSlice", "talshTensorInsert", "talshDetermineOptimalDevice", "talshTaskWait"):
        assert name in syms
    # the templates (contractAccumulate, ...) are instantiated by the GPU driver, also linked --no-undefined
    driver = os.path.join(REPO, "oracle", "_ref", "cxx_wrapper_driver")
    assert os.path.exists(driver)
    dsyms = subprocess.run(["nm", "-D", "--undefined-only", driver], capture_output=True, text=True).stdout
    assert "talshTensorContract" in dsyms
