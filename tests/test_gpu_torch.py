"""torch.autograd binding of the adjoint (lmtpp_amd.autograd): gradients through ElasticSolve
against central finite differences and against the C-ABI adjoint, and a short identification."""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from problems import build, lognormal_E

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(400)
def test_autograd_gradient_and_identification():
    print("[case] torch autograd", flush=True)
    import torch
    from lmtpp_amd.autograd import elastic_solve, identify_modulus
    sys.stdout.flush()
    c, e = L.make_rect(L.HEXA, (7, 7, 7))
    E_true = lognormal_E(len(e), 0.3, 8)
    f, *_ = build(L.HEXA, c, e, E_true)
    assert f.solve(rel_tol=1e-13)
    u_true = f.get_result()
    rng = np.random.default_rng(2)
    top = np.where(np.abs(c[:, 2] - 1.0) < 1e-9)[0]
    obs = rng.choice(top, size=12, replace=False)
    odofs = (3 * obs[:, None] + np.arange(3)).ravel()
    meas = torch.tensor(u_true[odofs], dtype=torch.float64, device="cuda")
    obs_t = torch.tensor(odofs, dtype=torch.long, device="cuda")

    E0 = torch.full((len(e),), 210000.0, dtype=torch.float64, device="cuda", requires_grad=True)
    u = elastic_solve(E0, f, 1e-13, 1e-13)
    J = 0.5 * torch.sum((u[obs_t] - meas) ** 2)
    J.backward()
    g = E0.grad.detach().cpu().numpy()
    # same gradient from the C ABI adjoint directly
    dJdu = np.zeros(3 * len(c))
    dJdu[odofs] = (u.detach().cpu().numpy()[odofs] - u_true[odofs])
    g2 = f.adjoint(dJdu, rel_tol=1e-13)
    assert np.abs(g - g2).max() <= 1e-10 * np.abs(g).max()
    # central finite differences on a few elements (step 1e-4 E)
    for el in rng.choice(len(e), size=4, replace=False):
        Jpm = []
        for sgn in (1, -1):
            Ep = torch.full((len(e),), 210000.0, dtype=torch.float64, device="cuda")
            Ep[el] *= 1 + sgn * 1e-4
            up = elastic_solve(Ep, f, 1e-14, 1e-13)
            Jpm.append(float(0.5 * torch.sum((up[obs_t] - meas) ** 2)))
        fd = (Jpm[0] - Jpm[1]) / (2e-4 * 210000.0)
        assert abs(fd - g[el]) <= 1e-5 * np.abs(g).max(), (el, fd, g[el])
    # a few Adam steps decrease the misfit
    E_fit, hist = identify_modulus(f, np.full(len(e), 210000.0), odofs, u_true[odofs], steps=8, lr=0.05)
    assert hist[-1] < 0.5 * hist[0]
    assert E_fit.shape[0] == len(e)
