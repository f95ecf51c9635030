"""__graft_entry__.smoke() (one small contraction and one permutation on the GPU, checked against the oracle),
run in its own process because it initialises and shuts down the library itself."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_smoke_entry_point():
    print("[case] smoke", flush=True)
    r = subprocess.run([sys.executable, "-c", "import __graft_entry__ as g; g.smoke()"], cwd=REPO,
                       capture_output=True, text=True, timeout=150)
    assert r.returncode == 0 and "smoke ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
