"""Text constraints of add_constraint(txt, penalty) (formulation_ancestor.h:197, constraint.h:46-93),
parsed on the host (no GPU): the affine form coefficient x unknown + constant that the reference
obtains by differentiating the expression and substituting the unknowns by zero."""
import math

import pytest

from lmtpp_amd import _capi as A


@pytest.mark.parametrize("txt,t,coeff,b", [
    ("node[0].dep[0]", 0.0, {0: 1.0}, 0.0),
    ("node[12].dep[2] + 0.001*time", 3.0, {38: 1.0}, 0.003),
    ("2*node[5].dep[1] - 0.004", 0.0, {16: 2.0}, -0.004),
    ("-(node[7].dep - 1e-3)/4", 0.0, {21: -0.25}, 0.00025),
    ("node[3].dep[1]*sin(pi/2)**2 + cos(time)", 0.5, {10: 1.0}, math.cos(0.5)),
    ("node [ 3 ] . dep [ 1 ] - 2^3", 0.0, {10: 1.0}, -8.0),          # spaces ignored, ^ as power
    ("node[1].dep[0] + node[2].dep[1] - node[1].dep[0]", 0.0, {7: 1.0}, 0.0),   # cancels to one unknown
    ("sqrt(4)*exp(0)*node[0].dep[2] + abs(-1)*log(1)", 0.0, {2: 2.0}, 0.0),
])
def test_affine_form(txt, t, coeff, b):
    got, gb = A.constraint_parse(txt, t)
    assert got.keys() == coeff.keys()
    for k in coeff:
        assert got[k] == pytest.approx(coeff[k], rel=1e-15, abs=0)
    assert gb == pytest.approx(b, rel=1e-15, abs=1e-18)


@pytest.mark.parametrize("txt", ["node[0].dep[0]*node[1].dep[0]", "1/node[2].dep[1]", "sin(node[0].dep[0])",
                                 "node[0].dep[3]", "node[0].temperature", "foo + 1", "node[0].dep[0] +", "(1 + 2",
                                 "node[0].dep[0]**2", ""])
def test_rejects_non_linear_and_malformed(txt):
    with pytest.raises(A.FemGpuError):
        A.constraint_parse(txt)
