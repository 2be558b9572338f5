"""Device operands from torch through the Python mirror.

torch ships its own HIP runtime; libdmhm_amd.so binds the first
libamdhip64.so.7 loaded in the process. The two share one runtime (and device
pointers are interchangeable) only when torch is imported before the library
is loaded, so these checks run in a child process that does exactly that:
column-major device views, X produced by pending torch kernels, Y freshly
allocated by the wrapper (the wrapper orders the library stream after
torch's)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
import torch                      # first: the library then binds torch's HIP runtime
sys.path.insert(0, sys.argv[1])
import dmhm_amd
from dmhm_amd import workloads as W
assert torch.cuda.is_available(), "torch sees no GPU"
dmhm_amd.init(0)
op = W.small_operator(16, 3)
H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, True, op["rowptr"], op["colind"], op["vals"])
def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()
X = np.random.default_rng(0).standard_normal((256, 3))
Xd = torch.from_numpy(X.T.copy()).cuda().t()                 # column-major view
Yd = H.Multiply(1.0, Xd)
torch.cuda.synchronize()
e1 = rel(Yd.cpu().numpy(), H.Multiply(1.0, X))
assert e1 < 1e-14, e1
g = torch.Generator(device="cuda").manual_seed(5)
for k in (1, 3):
    Xd = torch.randn((k, 256), dtype=torch.float64, device="cuda", generator=g).t()
    Xd = Xd * 3.0 + 1.0                                      # pending kernels on torch's stream
    Yd = H.Multiply(1.0, Xd)
    e2 = rel(Yd.cpu().numpy(), H.Multiply(1.0, Xd.cpu().numpy()))
    assert e2 < 1e-14, e2
Y0 = torch.randn((3, 256), dtype=torch.float64, device="cuda", generator=g).t()
Yc = Y0.clone()
H.Multiply(2.0, Xd, -1.0, Yc)                                # beta form, in place
e3 = rel(Yc.cpu().numpy(), 2.0 * H.Multiply(1.0, Xd.cpu().numpy()) - Y0.cpu().numpy())
assert e3 < 1e-13, e3
print("torch interop ok", e1, e2, e3)
"""


@pytest.mark.gpu
def test_torch_device_operands_in_a_torch_first_process():
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch interop ok" in r.stdout
