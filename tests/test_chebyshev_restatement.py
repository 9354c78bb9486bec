"""The numpy restatement of the Chebyshev block-Jacobi PCG (tests/chebyshev_ref.py), pinned on
the CPU: the hash against hand-computed values of the MurmurHash3 finalizer, the polynomial
against the Chebyshev equioscillation it must have, the power bound against the largest
eigenvalue from scipy's symmetric eigensolver, and the PCG against a dense solve."""
import numpy as np
import scipy.linalg as sla

import lmtpp_amd as L
from chebyshev_ref import apply_d, block_jacobi, coefficients, csr, pcg, power_bound, unit_hash
from problems import lognormal_E


def _fmix32(x):
    """The finalizer step by step in Python integers."""
    x = (x * 0x9e3779b1) & 0xffffffff
    x ^= x >> 16
    x = (x * 0x85ebca6b) & 0xffffffff
    x ^= x >> 13
    x = (x * 0xc2b2ae35) & 0xffffffff
    x ^= x >> 16
    return x


def test_hash_matches_integer_finalizer():
    idx = np.array([0, 1, 2, 3, 1000, 123456, 2 ** 31, 2 ** 32 - 1], np.uint64)
    want = np.array([_fmix32(int(i)) * (2.0 / 2 ** 32) - 1.0 for i in idx])
    assert np.array_equal(unit_hash(idx), want)
    assert unit_hash(np.array([0]))[0] == -1.0          # fmix32(0) = 0
    h = unit_hash(np.arange(100000))
    assert -1.0 <= h.min() and h.max() < 1.0 and abs(h.mean()) < 0.01


def test_polynomial_equioscillates_on_the_interval():
    """1 - t p(t) takes +-1/T2((b+a)/(b-a)) at a, (a+b)/2, b: the degree-2 Chebyshev
    residual polynomial; the preconditioned spectrum t p(t) is positive exactly on (0, a + b)."""
    lam = 2.87
    c0, c1 = coefficients(lam)
    b = 1.15 * lam
    a = 0.05 * b
    t2 = 2 * ((b + a) / (b - a)) ** 2 - 1
    R = lambda t: 1 - t * (c0 + c1 * t)  # noqa: E731
    assert np.isclose(R(a), 1 / t2, rtol=1e-12) and np.isclose(R(b), 1 / t2, rtol=1e-12)
    assert np.isclose(R(0.5 * (a + b)), -1 / t2, rtol=1e-12)
    t = np.linspace(1e-6, 0.9999 * (a + b), 10001)
    assert (t * (c0 + c1 * t) > 0).all()
    assert np.isclose(-c0 / c1, a + b, rtol=1e-12)
    assert 1.0001 * (a + b) * (c0 + c1 * 1.0001 * (a + b)) < 0


def _system(n=4, seed=3):
    c, e = L.make_rect(L.HEXA, (n + 1, n, n))
    from oracle import oracle as O
    from problems import clamp_and_traction
    E = lognormal_E(len(e), 0.5, seed)
    faces, tr, dofs = clamp_and_traction(L.HEXA, c, e, (0.0, 0.3, -1.0))
    rp, col, vals, F, rc = O.assemble(O.HEXA, c, e, E, 0.3, faces=faces, face_trac=tr)
    assert rc == 0
    O.penalty(rp, col, vals, F, dofs, None, 1e5)
    return rp, col, vals, F


def test_power_bound_below_largest_eigenvalue():
    rp, col, vals, F = _system()
    A = csr(rp, col, vals)
    Dinv = block_jacobi(rp, col, vals)
    Ad = A.toarray()
    Dd = sla.block_diag(*Dinv)
    # D K is similar to the symmetric D^1/2 K D^1/2: its eigenvalues from the pencil K x = l D^-1 x
    lam = sla.eigh(Ad, np.linalg.inv(Dd), eigvals_only=True)
    rq = power_bound(A, Dinv)
    assert lam.min() > 0
    assert rq <= lam.max() * (1 + 1e-12)
    assert rq >= 0.8 * lam.max()
    assert np.isclose(power_bound(A, Dinv, 400), lam.max(), rtol=1e-3)
    # D applied node by node is the block-diagonal inverse
    v = np.random.default_rng(0).standard_normal(len(F))
    assert np.allclose(apply_d(Dinv, v), Dd @ v, rtol=1e-12, atol=0)


def test_pcg_converges_to_dense_solution_in_fewer_iterations():
    rp, col, vals, F = _system()
    A = csr(rp, col, vals)
    Dinv = block_jacobi(rp, col, vals)
    xs = np.linalg.solve(A.toarray(), F)
    c0, c1 = coefficients(power_bound(A, Dinv))
    x, it = pcg(A, Dinv, F, c0, c1)
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
    # block Jacobi alone: c0 = 1, c1 = 0
    _, it_bj = pcg(A, Dinv, F, 1.0, 0.0)
    assert it < 0.8 * it_bj, (it, it_bj)
