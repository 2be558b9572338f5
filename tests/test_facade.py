"""The header-only C++ facade (include/dmhm_amd/hmat.hpp) used by a
sequential C++ caller written against the reference's class names
(tests/cpp/facade_driver.cpp, following tests/HMat2d/Invert.cpp of the
reference)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "facade_driver")


def _driver():
    if not os.path.exists(DRIVER):
        sys.path.insert(0, ROOT)
        import __graft_entry__
        __graft_entry__.build_facade_driver()
    return DRIVER


def test_facade_map_matches_oracle():
    import oracle as orc
    out = subprocess.run([_driver(), "map"], capture_output=True, text=True, timeout=60, check=True).stdout
    got = np.array(out.split(), dtype=np.int64)
    assert np.array_equal(got, orc.build_map(2, [8, 4], 3))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present; the no-device path is not reachable")
def test_facade_initialize_without_gpu_throws_runtime_error():
    r = subprocess.run([_driver(), "nogpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("runtime_error:")


@pytest.mark.gpu
def test_facade_sequential_caller_on_gpu():
    """Sparse import, H*x, H^T*x, H*Dense, H*H (both forms), CopyFrom,
    UpdateWith, TransposeFrom/Scale/AddConstantToDiagonal, EstimateTwoNorm,
    SchulzInvert, the LowRank constructor, Compress, RoundedUpdate, float
    storage and the exception mapping, each checked inside the driver."""
    print("facade driver: gpu mode", flush=True)
    r = subprocess.run([_driver(), "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
