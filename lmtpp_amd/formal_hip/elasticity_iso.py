"""The elasticity_iso weak form (formulations/formulation_elasticity_iso.py:45-60) at one Gauss
point, in terms of the physical gradients of the shape functions.

    res = density * dot(acc - f_vol, v) + trace_sym_col(sigma(u), eps(v))   (dV part)
        - dot(f_surf, v)                                                     (dS part)
    sigma = H eps,  H = inverse of the isotropic compliance (formal_lf/material_help.py:116-128)
    eps = grad_sym_col(.), Voigt [xx, yy, zz, xy, xz, yz] with tensor shear
          (formal_lf/element.py:270-296, include/codegen/__init__.py:23-42)
    trace_sym_col(s, e) = sum_normal s_i e_i + 2 sum_shear s_i e_i (include/codegen/__init__.py:82-87)

Stiffness entries are d2 res / d v_{a,i} d u_{b,j} and loads are -d res / d v_{a,i} at u = 0
(formal_lf/calculate_matrix.py:4-61).  With u restricted to node b and v to node a, the block
(a, b) depends only on the gradients g_a, g_b and on H.
"""
import sympy as sp

E, NU = sp.symbols("E nu", positive=True)
GA = sp.symbols("ga0:3")
GB = sp.symbols("gb0:3")
UB = sp.symbols("ub0:3")
VA = sp.symbols("va0:3")
NA = sp.Symbol("Na")
BF = sp.symbols("bf0:3")        # density * f_vol
H0, H1, H2 = sp.symbols("H0 H1 H2")


def compliance():
    s1, s2, s3 = 1 / E, -NU / E, (1 + NU) / E
    S = sp.zeros(6, 6)
    for i in range(3):
        for j in range(3):
            S[i, j] = s1 if i == j else s2
        S[3 + i, 3 + i] = s3
    return S


def hooke():
    """H = S^-1 and its three distinct entries (normal diagonal, normal off-diagonal, shear)."""
    H = sp.simplify(compliance().inv())
    h0, h1, h2 = sp.simplify(H[0, 0]), sp.simplify(H[0, 1]), sp.simplify(H[3, 3])
    # isotropy: every entry is one of h0, h1, h2 or zero
    for i in range(6):
        for j in range(6):
            want = (h0 if i == j else h1) if (i < 3 and j < 3) else (h2 if i == j else 0)
            assert sp.simplify(H[i, j] - want) == 0
    return H, (h0, h1, h2)


def _col(m):
    return [m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[0, 2], m[1, 2]]


def _eps(nodal, grad):
    G = sp.Matrix(3, 3, lambda i, j: nodal[i] * grad[j])   # d(u_i)/dx_j for one node
    return _col((G + G.T) / 2)


def trace_sym_col(s, e):
    return sum(s[i] * e[i] for i in range(3)) + 2 * sum(s[i] * e[i] for i in range(3, 6))


def block_entries():
    """k[3i+j] at one Gauss point (without w detJ) as expressions of GA, GB and H0, H1, H2."""
    Hs = sp.Matrix(6, 6, lambda i, j: (H0 if i == j else H1) if (i < 3 and j < 3) else (H2 if i == j else 0))
    sigma = Hs * sp.Matrix(_eps(UB, GB))
    res = trace_sym_col(list(sigma), _eps(VA, GA))
    return [sp.expand(sp.diff(res, VA[i], UB[j])) for i in range(3) for j in range(3)]


GU = sp.symbols("G0:9")       # G[3i+j] = d u_i / d x_j at the recovery point
EPS = sp.symbols("e0:6")


def strain_entries():
    """grad_sym_col of the displacement from its physical gradient G (formal_lf/element.py:270-296):
    Voigt [xx, yy, zz, xy, xz, yz], tensor shear (apply_on_elements_after_solve,
    formulations/formulation_elasticity_iso.py:64-85)."""
    G = sp.Matrix(3, 3, lambda i, j: GU[3 * i + j])
    return _col((G + G.T) / 2)


def stress_entries():
    """sigma = H eps with the isotropic H (entries H0, H1, H2) and Voigt eps."""
    Hs = sp.Matrix(6, 6, lambda i, j: (H0 if i == j else H1) if (i < 3 and j < 3) else (H2 if i == j else 0))
    return [sp.expand(v) for v in Hs * sp.Matrix(EPS)]


def load_entries():
    """-d res / d v_{a,i} at u = 0 for the volume term density * dot(acc - f_vol, v) (acc = 0,
    static); v = N_a v_a at the Gauss point, density * f_vol passed as BF."""
    res = sum((0 - BF[i]) * NA * VA[i] for i in range(3))
    return [sp.expand(-sp.diff(res, VA[i])) for i in range(3)]
