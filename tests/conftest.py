import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from tal_common import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def talsh_gpu():
    """Initialises libtalsh_amd on GPU 0 once per session."""
    import tal_sh_amd as ts

    err = ts.init(host_buf_size=1 << 30, dev_buffer_size=64 << 30)
    assert err == 0, err
    print("talsh_gpu: initialised on GPU 0", flush=True)
    yield ts
    ts.shutdown()
