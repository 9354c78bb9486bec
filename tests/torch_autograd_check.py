"""Child-process check of the torch binding (run by tests/test_gpu_torch.py).

torch is imported before libfemgpu.so is loaded, so both use torch's HIP runtime (a process that
loaded /opt/rocm's runtime first cannot start torch's).  Prints one JSON line of results."""
import json
import os
import sys

import torch  # noqa: F401  (first: one HIP runtime per process)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import lmtpp_amd as L  # noqa: E402
from lmtpp_amd.autograd import elastic_solve, identify_modulus  # noqa: E402
from oracle import oracle as O  # noqa: E402  (the checker: element matrices of the finite differences)
from problems import build, fd_gradient, lognormal_E  # noqa: E402


def main():
    assert torch.cuda.is_available()
    c, e = L.make_rect(L.HEXA, (7, 7, 7))
    E_true = lognormal_E(len(e), 0.3, 8)
    f, *_ = build(L.HEXA, c, e, E_true)
    assert f.solve(rel_tol=1e-11)
    u_true = f.get_result()
    rng = np.random.default_rng(2)
    top = np.where(np.abs(c[:, 2] - 1.0) < 1e-9)[0]
    obs = rng.choice(top, size=12, replace=False)
    odofs = (3 * obs[:, None] + np.arange(3)).ravel()
    meas = torch.tensor(u_true[odofs], dtype=torch.float64, device="cuda")
    obs_t = torch.tensor(odofs, dtype=torch.long, device="cuda")
    # a generic starting field (at a uniform one every interior node holds max|diag K|)
    E_start = 210000.0 * np.exp(0.05 * np.random.default_rng(9).standard_normal(len(e)))
    E0 = torch.tensor(E_start, dtype=torch.float64, device="cuda", requires_grad=True)
    u = elastic_solve(E0, f, 1e-11, 1e-11)
    J = 0.5 * torch.sum((u[obs_t] - meas) ** 2)
    J.backward()
    g = E0.grad.detach().cpu().numpy()
    un = u.detach().cpu().numpy()
    dJdu = np.zeros(3 * len(c))
    dJdu[odofs] = un[odofs] - u_true[odofs]
    g2 = f.adjoint(dJdu, rel_tol=1e-11)
    out = {"grad_vs_capi": float(np.abs(g - g2).max() / np.abs(g).max()), "fd": []}
    res = un[odofs] - u_true[odofs]
    for el in rng.choice(len(e), size=6, replace=False):
        fd = fd_gradient(f, O, E_start, un, res, odofs, int(el))
        if fd is not None:
            out["fd"].append([int(el), fd, float(g[el])])
    out["gmax"] = float(np.abs(g).max())
    _, hist = identify_modulus(f, E_start, odofs, u_true[odofs], steps=8, lr=0.05)
    out["J_first"], out["J_last"] = hist[0], hist[-1]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
