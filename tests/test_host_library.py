"""CPU tests of libdmhm_amd.so: it loads, exports every entry point of
include/dmhm_amd.h, and its host-side geometry (hierarchical map, block tree,
product recursion) matches the oracle and the reference's counts. No compute
calls are made (no GPU here)."""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import dmhm_amd
from dmhm_amd import _lib
from dmhm_amd import workloads as W
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dmhm_amd.h")
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "tree_counts.json")))


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmhm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "dmhm_hmat_multiply_dense" in names and "dmhm_hmat_schulz_invert" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("sizes,L", [([16, 16], 3), ([20, 12], 3), ([64, 64], 4), ([8, 6, 5], 3),
                                     ([32, 32, 32], 4)])
def test_natural_to_hierarchical_map(sizes, L):
    m = dmhm_amd.build_natural_to_hierarchical_map(sizes, L)
    np.testing.assert_array_equal(m, orc.build_map(len(sizes), sizes, L))
    np.testing.assert_array_equal(m, W.natural_to_hierarchical_map(sizes, L))


@pytest.mark.parametrize("case", GOLDEN["cases"])
def test_geometry_block_counts_match_reference(case):
    d, f = dmhm_amd.geometry_block_counts(case["sizes"], case["numLevels"], case["strong"])
    for lvl, cnt in case["dense"].items():
        assert d[int(lvl)] == cnt
    for lvl, cnt in case["lowRank"].items():
        assert f[int(lvl)] == cnt
    assert d.sum() == sum(case["dense"].values())
    if not case.get("partial"):
        assert f.sum() == case.get("lowRankTotal", sum(case["lowRank"].values()))


@pytest.mark.parametrize("case", GOLDEN["productCounts"]["cases"])
def test_product_recursion_term_counts(case):
    t = dmhm_amd.geometry_product_terms(case["sizes"], case["numLevels"], True)
    assert t["sketch"] == case["sketches"]
    assert t["dd_to_lowrank"] == case["ddToLowRank"]


def test_geometry_matches_oracle_on_ragged_grid():
    op = W.small_operator(12, 3, dim=3, smooth=False)
    H = orc.HMat.from_csr(3, op["sizes"], 3, 4, True, op["rowptr"], op["colind"], op["vals"])
    d0, f0 = H.count_blocks(6)
    d1, f1 = dmhm_amd.geometry_block_counts(op["sizes"], 3, True, 6)
    np.testing.assert_array_equal(d0, d1)
    np.testing.assert_array_equal(f0, f1)


def test_invalid_geometry_is_rejected():
    with pytest.raises(_lib.DmhmError) as e:
        dmhm_amd.geometry_block_counts([16, 16, 16, 16], 3, True)
    assert e.value.code == 1


def test_compute_entry_points_require_init():
    """Without dmhm_init every device entry point reports DMHM_ENODEV (fresh
    process, so an earlier init in this session cannot interfere)."""
    code = r"""
import ctypes, sys
sys.path.insert(0, %r)
from dmhm_amd import _lib
L = _lib.load()
h = ctypes.c_void_p()
sizes = (ctypes.c_int * 2)(4, 4)
rp = (ctypes.c_int * 17)(*range(17)); ci = (ctypes.c_int * 16)(*range(16)); v = (ctypes.c_double * 16)(*([1.0] * 16))
print(L.dmhm_hmat_create_from_csr(2, sizes, 2, 4, 1, 8, rp, ci, v, ctypes.byref(h)))
print(L.dmhm_hmat_multiply_dense(None, b'N', 1.0, None, 16, 1, 0.0, None, 16, 0))
""" % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    codes = [int(x) for x in out.stdout.split()]
    assert codes == [7, 7]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_init_without_gpu_fails_loudly():
    code = "import sys; sys.path.insert(0, %r); from dmhm_amd import _lib; print(_lib.load().dmhm_init(0))" % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0
    assert int(out.stdout.split()[-1]) == 7


def test_workload_operator_matches_reference_stencil():
    """Flux-form 5-point stencil (tests/HMat2d/Invert.cpp:12-67) with mass 1:
    symmetric, diagonal = sum of |off-diagonals| on interior rows + 1 + boundary faces."""
    op = W.small_operator(8, 2)
    D = W.csr_to_dense(op["rowptr"], op["colind"], op["vals"], 64)
    np.testing.assert_allclose(D, D.T, atol=1e-12)
    off = D - np.diag(np.diag(D))
    assert np.all(off <= 0)
    assert np.all(np.diag(D) + off.sum(axis=1) >= 1.0 - 1e-9)
    ev = np.linalg.eigvalsh(D)
    assert ev.min() > 1.0
