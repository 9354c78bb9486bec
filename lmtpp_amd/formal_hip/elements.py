"""Element descriptions for the generator: reference-space variables, nodal interpolation and the
Gauss tables of the runtime quadrature.

Interpolations restate formulations/element_Tetra.py:27-30, element_Tetra_10.py:27-41,
element_Hexa.py:61-68 (volume) and element_Triangle.py, element_Triangle_6.py:38-48,
element_Quad.py:22-25 (faces).  Volume Gauss tables are those the runtime loop reads
(include/mesh/element_Tetra.h:8,12-15, element_Hexa.h:12-19) at order 2*degree-2
(formal_lf/formulation.py:30-32,121-123); face tables are the generator-side Python tables of the
face elements at the parent's order (formal_lf/formulation.py:400-442, element_Quad.py:70-75).
Gauss values are kept as the decimal strings of those tables so the emitted constants are the
same doubles.
"""
import sympy as sp

X = sp.symbols("xi0 xi1 xi2")


class Element:
    def __init__(self, name, kind, dim, shapes, gauss):
        self.name, self.kind, self.dim = name, kind, dim
        self.shapes = list(shapes)
        self.gauss = gauss            # list of (weight string, (coordinate strings))
        self.nb_nodes = len(shapes)

    def var(self):
        return X[: self.dim]


def _tetra():
    x, y, z = X
    return Element("Tetra", 0, 3, [1 - x - y - z, x, y, z],
                   [("0.16666666666666666667", ("0.25", "0.25", "0.25"))])


def _tetra10():
    x, y, z = X
    L = [1 - x - y - z, x, y, z]
    shapes = [L[a] * (2 * L[a] - 1) for a in range(4)]
    for i, j in [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]:   # nodes 4..9
        shapes.append(4 * L[i] * L[j])
    a, b = "0.1381966011250105", "0.5854101966249685"
    w = "0.041666666666666666667"
    return Element("Tetra_10", 1, 3, shapes, [(w, (a, a, a)), (w, (b, a, a)), (w, (a, b, a)), (w, (a, a, b))])


def _hexa():
    x, y, z = X
    one = [1 - x, x]
    oy, oz = [1 - y, y], [1 - z, z]
    bits = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    shapes = [one[i] * oy[j] * oz[k] for i, j, k in bits]
    lo, hi = "0.211324865405", "0.788675134595"
    pts = [(lo, lo, lo), (hi, lo, lo), (hi, hi, lo), (lo, hi, lo), (lo, lo, hi), (hi, lo, hi), (hi, hi, hi), (lo, hi, hi)]
    return Element("Hexa", 2, 3, shapes, [("0.125", p) for p in pts])


def _tri3():
    u, v = X[:2]
    return Element("Triangle", 0, 2, [1 - u - v, u, v], [("0.5", ("1/3", "1/3"))])


def _tri6():
    u, v = X[:2]
    L0 = 1 - u - v
    shapes = [(2 * L0 - 1) * L0, (2 * u - 1) * u, (2 * v - 1) * v, 4 * L0 * u, 4 * u * v, 4 * v * L0]
    w = "1/6"
    return Element("Triangle_6", 1, 2, shapes, [(w, ("0.5", "0")), (w, ("0.5", "0.5")), (w, ("0", "0.5"))])


def _quad4():
    u, v = X[:2]
    shapes = [(1 - u) * (1 - v), u * (1 - v), u * v, (1 - u) * v]
    lo, hi = "(1-sqrt(1/3))/2", "(1+sqrt(1/3))/2"
    return Element("Quad", 2, 2, shapes, [("0.25", (hi, hi)), ("0.25", (lo, hi)), ("0.25", (hi, lo)), ("0.25", (lo, lo))])


VOLUME = [_tetra(), _tetra10(), _hexa()]
FACES = [_tri3(), _tri6(), _quad4()]
