"""The storm::gpu::allpairs_sum / allpairs_tiles C++ mirror on the GPU: the
test program (tests/cpp/test_allpairs_gpu.cpp, built next to the libraries)
checks both API forms under every supported policy against brute force."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "stormbitmaps-internal_amd", "test_allpairs_gpu")


def test_cpp_allpairs_mirror():
    assert os.path.exists(BIN), f"{BIN} not built (make -C stormbitmaps-internal_amd)"
    print("case cpp mirror", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert " 0 failures" in r.stdout
