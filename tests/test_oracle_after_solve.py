"""Pinning of the oracle's after-solve recovery (or_after_solve).

Hexa: against the reference's own generated after_solve (compiled from
std_formulations/definition_probleme_elasticity_dep_static.h:721-789 by oracle/Makefile `ref`,
fixture tests/golden/ref_after_solve_hexa_4_3_5.npz).  Every kind: affine displacements give the
exact constant strain and the Lame stress; a quadratic field on straight-edged Tetra_10 elements
is differentiated exactly at the physical image of xi = (1/2,1/2,1/2) (outside the simplex),
which tells that point apart from the centroid.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def voigt(S):
    return np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


def test_hexa_matches_reference_generated_after_solve(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_after_solve_hexa_4_3_5.npz"))
    c, e = oracle.make_rect(oracle.HEXA, (4, 3, 5))
    assert np.array_equal(e, g["conn"])
    eps, sig, tr = oracle.after_solve(oracle.HEXA, g["conn"], g["coords"], float(g["E"]), float(g["nu"]), g["u"])
    assert np.abs(eps - g["eps"]).max() <= 1e-14 * np.abs(g["eps"]).max()
    assert np.abs(sig - g["sigma"]).max() <= 1e-14 * np.abs(g["sigma"]).max()
    assert np.abs(tr - g["tr"]).max() <= 1e-14 * np.abs(g["tr"]).max()
    # not a trivial field: the strain varies from element to element
    assert np.ptp(g["eps"][:, 0]) > 1e-4


def _mesh(oracle, kind, seed):
    from problems import jitter_interior, tetra10_jittered
    import lmtpp_amd as L
    if kind == oracle.TETRA10:
        return tetra10_jittered(2, seed=seed)
    c, e = L.make_rect(kind, (4, 3, 3))
    return jitter_interior(c, 0.3, 0.2, seed), e


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_affine_field_gives_exact_strain_and_lame_stress(oracle, kind):
    c, e = _mesh(oracle, kind, 5 + kind)
    rng = np.random.default_rng(kind)
    A = 1e-3 * rng.standard_normal((3, 3))
    b = 1e-3 * rng.standard_normal(3)   # same scale as A h, so cancellation stays at rounding level
    u = (c @ A.T + b).ravel()
    E = 1000.0 * (1 + rng.random(len(e)))
    nu = 0.31
    eps, sig, tr = oracle.after_solve(kind, e, c, E, nu, u)
    ex = voigt((A + A.T) / 2)
    assert np.abs(eps - ex).max() <= 1e-13 * np.abs(ex).max()
    assert np.abs(tr - np.trace(A)).max() <= 1e-13 * np.abs(A).max()
    lam, mu = lame(E, nu)
    sx = 2 * mu[:, None] * ex[None, :]
    sx[:, :3] += (lam * np.trace(A))[:, None]
    assert np.abs(sig - sx).max() <= 1e-12 * np.abs(sx).max()


def test_tetra10_point_is_xi_one_half(oracle):
    """On straight-edged Tetra_10 elements a quadratic displacement is represented exactly, so the
    recovered strain is the analytic one at x(xi = 1/2, 1/2, 1/2) = x0 + J^T (1/2, 1/2, 1/2)."""
    import lmtpp_amd as L
    c, e = L.make_rect(L.TETRA10, (3, 3, 3))
    rng = np.random.default_rng(3)
    Q = 1e-3 * rng.standard_normal((3, 3, 3))
    Q = (Q + Q.transpose(0, 2, 1)) / 2          # u_i = x^T Q_i x
    u = np.einsum("nj,ijk,nk->ni", c, Q, c).ravel()
    eps, _, _ = oracle.after_solve(L.TETRA10, e, c, 1.0, 0.3, u)
    for el in range(len(e)):
        x = c[e[el, :4]]
        xp = x[0] + (x[1:] - x[0]).T @ np.array([0.5, 0.5, 0.5])   # affine map of the corners
        grad = 2 * np.einsum("ijk,k->ij", Q, xp)                  # d u_i / d x_j
        assert np.abs(eps[el] - voigt((grad + grad.T) / 2)).max() <= 1e-13
    # the centroid would give a different value
    xc = c[e[0, :4]].mean(axis=0)
    gc = 2 * np.einsum("ijk,k->ij", Q, xc)
    assert np.abs(eps[0] - voigt((gc + gc.T) / 2)).max() > 1e-5
