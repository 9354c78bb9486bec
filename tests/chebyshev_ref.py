"""numpy restatement of the Chebyshev block-Jacobi PCG (FEMGPU_CHEBYSHEV_BLOCK_JACOBI), for the
tests: the same start vector, power iteration, bound and polynomial as the library's solver, so
that its iteration counts can be compared with the library's.

The preconditioner has no counterpart in the reference.  With D the inverse of K's 3x3 diagonal
blocks, z = c0 D r + c1 D K D r, where 1 - t (c0 + c1 t) is T2((b + a - 2t) / (b - a)) /
T2((b + a) / (b - a)) (T2(x) = 2 x^2 - 1, the Chebyshev polynomial of degree 2), a = LOWER b and
b = SAFETY x the Rayleigh quotient (K v . D K v) / (v . K v) after POWER_STEPS power steps
v <- D K v / ||v|| from a fixed start.
"""
import numpy as np
import scipy.sparse as sp

POWER_STEPS, SAFETY, LOWER = 20, 1.15, 0.05


def unit_hash(i):
    """MurmurHash3's 32-bit finalizer of i * 0x9e3779b1, mapped to [-1, 1)."""
    m = 0xffffffff
    h = (np.asarray(i, np.uint64) * np.uint64(0x9e3779b1)) & np.uint64(m)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85ebca6b)) & np.uint64(m)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xc2b2ae35)) & np.uint64(m)
    h ^= h >> np.uint64(16)
    return h.astype(np.float64) * (2.0 / 4294967296.0) - 1.0


def csr(rp, col, vals):
    """Scalar CSR of a block-CSR matrix (3x3 blocks)."""
    nn = len(rp) - 1
    rows = np.repeat(np.arange(nn), np.diff(rp))
    R = (3 * rows[:, None, None] + np.arange(3)[None, :, None] + np.zeros(3, int)[None, None, :]).ravel()
    C = (3 * np.asarray(col)[:, None, None] + np.zeros(3, int)[None, :, None] + np.arange(3)[None, None, :]).ravel()
    return sp.csr_matrix((np.asarray(vals).ravel(), (R, C)), shape=(3 * nn, 3 * nn))


def block_jacobi(rp, col, vals):
    """Inverses of the diagonal blocks, shape (n_nodes, 3, 3)."""
    nn = len(rp) - 1
    pos = np.array([rp[i] + np.searchsorted(col[rp[i]:rp[i + 1]], i) for i in range(nn)])
    return np.linalg.inv(np.asarray(vals)[pos])


def apply_d(Dinv, v):
    return np.einsum("nij,nj->ni", Dinv, v.reshape(len(Dinv), 3)).ravel()


def power_bound(A, Dinv, steps=POWER_STEPS):
    """Rayleigh quotient of D K after `steps` power steps from the hashed start vector."""
    v = unit_hash(np.arange(A.shape[0]))
    norm = np.linalg.norm(v)
    rq = 0.0
    for _ in range(steps):
        u = A @ v
        w = apply_d(Dinv, u)
        rq = (u @ w) / (v @ u)
        v = w / norm
        norm = np.linalg.norm(v)
    return rq


def coefficients(lam, safety=SAFETY, lower=LOWER):
    b = safety * lam
    a = lower * b
    th, de = 0.5 * (b + a), 0.5 * (b - a)
    t2 = 2 * (th / de) ** 2 - 1
    return 4 * th / de ** 2 / t2, -2 / de ** 2 / t2


def pcg(A, Dinv, F, c0, c1, rel_tol=1e-12, max_iter=20000):
    """PCG with z = c0 D r + c1 D K D r from x = 0; returns (x, iterations)."""
    def prec(r):
        y = apply_d(Dinv, r)
        return c0 * y + c1 * apply_d(Dinv, A @ y)
    x = np.zeros_like(F)
    r = F.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(F)
    it = 0
    while np.linalg.norm(r) > rel_tol * nb and it < max_iter:
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = prec(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
        it += 1
    return x, it
