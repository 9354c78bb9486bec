"""torch.autograd binding of the adjoint (lmtpp_amd.autograd): gradients through ElasticSolve
against the C-ABI adjoint and central finite differences (per element, relative 1e-5 where
|g_e| >= 1e-3 max|g|), and a short identification.

Runs in a child process that imports torch before libfemgpu.so (one HIP runtime per process)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(400)
def test_autograd_gradient_and_identification():
    print("[case] torch autograd (child process)", flush=True)
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_autograd_check.py")
    r = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True, timeout=380)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["grad_vs_capi"] <= 1e-10
    assert len(res["fd"]) >= 4
    for el, fd, g in res["fd"]:   # per element: relative where |g| >= 1e-3 max|g|
        if abs(g) >= 1e-3 * res["gmax"]:
            assert abs(fd - g) <= 1e-5 * abs(g), (el, fd, g)
        else:
            assert abs(fd - g) <= 1e-8 * res["gmax"], (el, fd, g)
    assert res["J_last"] < 0.5 * res["J_first"]
