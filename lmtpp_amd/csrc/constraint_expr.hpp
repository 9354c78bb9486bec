// Textual constraints of the reference's add_constraint(txt, penalty) (formulation_ancestor.h:197,
// constraint.h:46-93): an expression whose zero is the constraint, linear in the unknowns
// node[i].dep[d], with numbers, `time`, + - * / ** (or ^), parentheses and the functions sin, cos,
// tan, exp, log, sqrt, abs of unknown-free arguments.  The reference differentiates the expression
// symbolically for the coefficients and substitutes the unknowns by 0 for the right-hand side;
// here every node evaluates directly to its affine form (constant + sum of coefficient x unknown),
// which is exact for the linear expressions the reference accepts and refuses the non-linear
// ones.  Host only.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>

namespace femgpu {

struct ExprNode;

class ConstraintExpr {
  public:
    // Parses txt (spaces ignored, as the reference strips them); returns an empty string or the
    // error message.
    std::string parse(const std::string &txt);
    // Affine form at time t: expr = b + sum_k coeff[k] u_k, u_k the DOF 3 node + comp; returns an
    // empty string or the error message (a product of two unknown-dependent factors, ...).
    std::string eval(double t, std::map<uint64_t, double> &coeff, double &b) const;
    bool uses_time() const { return time_; }
    const std::string &text() const { return txt_; }

  private:
    std::shared_ptr<ExprNode> root_;
    std::string txt_;
    bool time_ = false;
};

}  // namespace femgpu
