"""HIP code emission: common-subexpression elimination and C printing of the generated element
expressions, the counterpart of formal_lf's Write_code (include/codegen/write_code.h:77-129)."""
import sympy as sp

from . import elasticity_iso as F
from .elements import FACES, VOLUME, X

HEADER = """// elasticity_iso.inc -- GENERATED by `python -m lmtpp_amd.formal_hip`; do not edit.
//
// HIP device functions of the elasticity_iso formulation (formulations/formulation_elasticity_iso.py:45-60)
// and of the Tetra, Tetra_10, Hexa volume and Triangle, Triangle_6, Quad face elements
// (formulations/element_*.py), derived symbolically by lmtpp_amd/formal_hip:
//   hooke   : distinct entries of H = inverse of the isotropic compliance
//   block   : one Gauss point's contribution w detJ d2 res / d v_a d u_b to the 3x3 block (a, b)
//   load    : one Gauss point's contribution w detJ (-d res / d v_a) of the body force
//   strain, stress : after-solve recovery eps = grad_sym_col(u), sigma = H eps
//            (apply_on_elements_after_solve, formulations/formulation_elasticity_iso.py:64-85)
//   vol<K>  : Gauss table and shape functions with reference gradients of volume element K
//   face<F> : the same for face element F (tangent derivatives in d[0], d[1])
#pragma once

namespace femgpu {
namespace gen {
"""

FOOTER = """
}  // namespace gen
}  // namespace femgpu
"""


_POW = {"Pow": [(lambda b, e: e == 2, lambda b, e: "(%s*%s)" % (b, b)),
                (lambda b, e: e == 3, lambda b, e: "(%s*%s*%s)" % (b, b, b)),
                (lambda b, e: True, lambda b, e: "pow(%s, %s)" % (b, e))]}


def _c(expr):
    return sp.ccode(expr, standard="c99", user_functions=_POW)


def _cse_block(exprs, targets, op, indent="    ", prefix="t"):
    syms = sp.numbered_symbols(prefix)
    reps, red = sp.cse(exprs, symbols=syms, optimizations="basic")
    out = []
    for s, e in reps:
        out.append("%sconst double %s = %s;" % (indent, s, _c(e)))
    for t, e in zip(targets, red):
        out.append("%s%s %s %s;" % (indent, t, op, _c(e)))
    return out


def _const(txt):
    """Decimal strings of the reference tables kept as written (same doubles); fractions and the
    closed-form Quad abscissae written as the same double arithmetic the oracle uses."""
    if "sqrt" in txt or "/" in txt:
        return txt.replace("1/3", "(1.0 / 3.0)").replace("1/6", "(1.0 / 6.0)").replace("(1-", "(1.0 - ") \
                  .replace("(1+", "(1.0 + ").replace(")/2", ") / 2.0")
    return txt


def emit_hooke():
    _, (h0, h1, h2) = F.hooke()
    lines = ["// H = inverse of the compliance s1 = 1/E, s2 = -nu/E, s3 = (1+nu)/E (tensor shear)",
             "// (formal_lf/material_help.py:116-128): H[0] = H11, H[1] = H12, H[2] = H44.",
             "__host__ __device__ inline void hooke(double E, double nu, double H[3]) {"]
    lines += _cse_block([h0, h1, h2], ["H[0]", "H[1]", "H[2]"], "=", prefix="h")
    lines.append("}")
    return lines


def emit_block():
    ents = F.block_entries()
    ga, gb = F.GA, F.GB
    sub = {ga[i]: sp.Symbol("ga[%d]" % i) for i in range(3)}
    sub.update({gb[i]: sp.Symbol("gb[%d]" % i) for i in range(3)})
    sub.update({F.H0: sp.Symbol("H[0]"), F.H1: sp.Symbol("H[1]"), F.H2: sp.Symbol("H[2]")})
    ents = [e.xreplace(sub) for e in ents]
    wd = sp.Symbol("wd")
    lines = ["// k[3i+j] += w detJ d2(trace_sym_col(H eps(u), eps(v))) / d v_{a,i} d u_{b,j}",
             "__device__ inline void block(const double *ga, const double *gb, const double *H, double wd, double *k) {"]
    lines += _cse_block([wd * e for e in ents], ["k[%d]" % i for i in range(9)], "+=")
    lines.append("}")
    return lines


def emit_load():
    ents = F.load_entries()
    sub = {F.NA: sp.Symbol("Na"), F.BF[0]: sp.Symbol("bf0"), F.BF[1]: sp.Symbol("bf1"), F.BF[2]: sp.Symbol("bf2")}
    wd = sp.Symbol("wd")
    lines = ["// f[i] += w detJ (-d res / d v_{a,i}) of density * dot(acc - f_vol, v), static (acc = 0)",
             "__device__ inline void load(double Na, double bf0, double bf1, double bf2, double wd, double *f) {"]
    lines += _cse_block([wd * e.xreplace(sub) for e in ents], ["f[%d]" % i for i in range(3)], "+=")
    lines.append("}")
    return lines


def emit_strain_stress():
    sub = {F.GU[i]: sp.Symbol("G[%d]" % i) for i in range(9)}
    lines = ["// eps = grad_sym_col(u) from G[3i+j] = du_i/dx_j: Voigt [xx, yy, zz, xy, xz, yz], tensor shear",
             "__host__ __device__ inline void strain(const double *G, double *eps) {"]
    lines += _cse_block([e.xreplace(sub) for e in F.strain_entries()], ["eps[%d]" % i for i in range(6)], "=",
                        prefix="g")
    lines += ["}", "", "// sigma = H eps (H[0] = H11, H[1] = H12, H[2] = H44)",
              "__host__ __device__ inline void stress(const double *H, const double *eps, double *sig) {"]
    sub = {F.EPS[i]: sp.Symbol("eps[%d]" % i) for i in range(6)}
    sub.update({F.H0: sp.Symbol("H[0]"), F.H1: sp.Symbol("H[1]"), F.H2: sp.Symbol("H[2]")})
    lines += _cse_block([e.xreplace(sub) for e in F.stress_entries()], ["sig[%d]" % i for i in range(6)], "=",
                        prefix="r")
    lines.append("}")
    return lines


def _emit_element(el, tmpl):
    dim = el.dim
    lines = ["// %s: %d nodes, %d Gauss points" % (el.name, el.nb_nodes, len(el.gauss)),
             "template <> struct %s<%d> {" % (tmpl, el.kind),
             "    static constexpr int NPE = %d, NGP = %d;" % (el.nb_nodes, len(el.gauss)),
             "    __device__ static inline void gauss(int q, double *xi, double &w) {",
             "        switch (q) {"]
    for q, (w, pts) in enumerate(el.gauss):
        body = "w = %s; " % _const(w) + " ".join("xi[%d] = %s;" % (d, _const(p)) for d, p in enumerate(pts))
        lines.append("            %s: %s break;" % ("default" if q == len(el.gauss) - 1 else "case %d" % q, body))
    lines += ["        }", "    }",
              "    // value of shape function a at xi; reference gradient in d[0..%d]" % (dim - 1),
              "    __device__ static inline double shape(int a, const double *xi, double *d) {"]
    var = X[:dim]
    sub = {var[i]: sp.Symbol("xi[%d]" % i) for i in range(dim)}
    lines.append("        double N;")
    lines.append("        switch (a) {")
    for a, N in enumerate(el.shapes):
        exprs = [N] + [sp.diff(N, v) for v in var]
        exprs = [e.xreplace(sub) for e in exprs]
        lab = "default" if a == el.nb_nodes - 1 else "case %d" % a
        lines.append("            %s: {" % lab)
        lines += _cse_block(exprs, ["N"] + ["d[%d]" % i for i in range(dim)], "=", indent="                ",
                            prefix="s")
        lines.append("                break;")
        lines.append("            }")
    lines += ["        }", "        return N;", "    }", "};"]
    return lines


def generate():
    out = HEADER.splitlines()
    out += [""] + emit_hooke() + [""] + emit_block() + [""] + emit_load() + [""] + emit_strain_stress() + [""]
    out += ["template <int K> struct vol;", "template <int F> struct face;", ""]
    for el in VOLUME:
        out += _emit_element(el, "vol") + [""]
    for el in FACES:
        out += _emit_element(el, "face") + [""]
    out += FOOTER.splitlines()
    return "\n".join(out) + "\n"
