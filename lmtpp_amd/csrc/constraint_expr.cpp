// Parser and affine evaluation of the textual constraints (see constraint_expr.hpp).
#include "constraint_expr.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace femgpu {

struct ExprNode {
    enum Kind { NUM, TIME, UNK, NEG, ADD, SUB, MUL, DIV, POW, FUN } kind = NUM;
    double v = 0;          // NUM
    uint64_t dof = 0;      // UNK: 3 node + component
    std::string fun;       // FUN
    std::shared_ptr<ExprNode> a, b;
};

namespace {

using Node = std::shared_ptr<ExprNode>;

Node make(ExprNode::Kind k, Node a = nullptr, Node b = nullptr) {
    auto n = std::make_shared<ExprNode>();
    n->kind = k;
    n->a = a;
    n->b = b;
    return n;
}

// Recursive descent over: expr = term (('+' | '-') term)*; term = unary (('*' | '/') unary)*;
// unary = ('-' | '+') unary | power; power = atom (('**' | '^') unary)?;
// atom = number | 'time' | 'pi' | fun '(' expr ')' | 'node[' int '].dep' ('[' int ']')? | '(' expr ')'
class Parser {
  public:
    explicit Parser(const std::string &s) : s_(s) {}
    Node parse(std::string &err, bool &uses_time) {
        Node n = expr();
        if (err_.empty() && i_ != s_.size()) err_ = "unexpected '" + s_.substr(i_, 1) + "' at " + std::to_string(i_);
        err = err_;
        uses_time = time_;
        return err_.empty() ? n : nullptr;
    }

  private:
    const std::string &s_;
    size_t i_ = 0;
    std::string err_;
    bool time_ = false;

    bool eat(const char *t) {
        const size_t n = std::char_traits<char>::length(t);
        if (s_.compare(i_, n, t) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    Node fail(const std::string &m) {
        if (err_.empty()) err_ = m;
        return make(ExprNode::NUM);
    }
    Node expr() {
        Node n = term();
        while (err_.empty()) {
            if (eat("+")) n = make(ExprNode::ADD, n, term());
            else if (eat("-")) n = make(ExprNode::SUB, n, term());
            else break;
        }
        return n;
    }
    Node term() {
        Node n = unary();
        while (err_.empty()) {
            if (s_.compare(i_, 2, "**") == 0) break;
            if (eat("*")) n = make(ExprNode::MUL, n, unary());
            else if (eat("/")) n = make(ExprNode::DIV, n, unary());
            else break;
        }
        return n;
    }
    Node unary() {
        if (eat("-")) return make(ExprNode::NEG, unary());
        if (eat("+")) return unary();
        return power();
    }
    Node power() {
        Node n = atom();
        if (err_.empty() && (eat("**") || eat("^"))) n = make(ExprNode::POW, n, unary());
        return n;
    }
    bool integer(uint64_t &v) {
        const size_t i0 = i_;
        v = 0;
        while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) v = 10 * v + (uint64_t)(s_[i_++] - '0');
        return i_ > i0 && v < (1ull << 40);
    }
    Node atom() {
        if (i_ >= s_.size()) return fail("expression ends too early");
        const char c = s_[i_];
        if (std::isdigit((unsigned char)c) || c == '.') {
            char *end = nullptr;
            const double v = std::strtod(s_.c_str() + i_, &end);
            if (end == s_.c_str() + i_) return fail("bad number at " + std::to_string(i_));
            i_ = (size_t)(end - s_.c_str());
            Node n = make(ExprNode::NUM);
            n->v = v;
            return n;
        }
        if (eat("(")) {
            Node n = expr();
            if (err_.empty() && !eat(")")) return fail("missing ')'");
            return n;
        }
        if (eat("node[")) {
            uint64_t node = 0, comp = 0;
            if (!integer(node) || !eat("].")) return fail("syntax for node data is node[i].dep[d]");
            if (!eat("dep")) return fail("the unknown of elasticity_iso is dep (node[i].dep[d])");
            if (eat("[")) {
                if (!integer(comp) || !eat("]") || comp > 2) return fail("dep has components 0, 1, 2");
            }
            Node n = make(ExprNode::UNK);
            n->dof = 3 * node + comp;
            return n;
        }
        std::string id;
        while (i_ < s_.size() && (std::isalpha((unsigned char)s_[i_]) || s_[i_] == '_')) id += s_[i_++];
        if (id == "time") {
            time_ = true;
            return make(ExprNode::TIME);
        }
        if (id == "pi") {
            Node n = make(ExprNode::NUM);
            n->v = M_PI;
            return n;
        }
        static const char *funs[] = {"sin", "cos", "tan", "exp", "log", "sqrt", "abs"};
        for (const char *f : funs)
            if (id == f) {
                if (!eat("(")) return fail(id + " needs an argument in parentheses");
                Node n = make(ExprNode::FUN, expr());
                n->fun = id;
                if (err_.empty() && !eat(")")) return fail("missing ')'");
                return n;
            }
        return fail(id.empty() ? "unexpected '" + s_.substr(i_, 1) + "'" : "unknown symbol '" + id + "'");
    }
};

struct Affine {
    double b = 0;
    std::map<uint64_t, double> c;
    bool constant() const { return c.empty(); }
};

std::string evaluate(const ExprNode &n, double t, Affine &out) {
    Affine x, y;
    std::string e;
    if (n.a && !(e = evaluate(*n.a, t, x)).empty()) return e;
    if (n.b && !(e = evaluate(*n.b, t, y)).empty()) return e;
    switch (n.kind) {
        case ExprNode::NUM: out.b = n.v; break;
        case ExprNode::TIME: out.b = t; break;
        case ExprNode::UNK: out.c[n.dof] = 1.0; break;
        case ExprNode::NEG:
            out.b = -x.b;
            for (auto &kv : x.c) out.c[kv.first] = -kv.second;
            break;
        case ExprNode::ADD:
        case ExprNode::SUB: {
            const double sg = n.kind == ExprNode::ADD ? 1.0 : -1.0;
            out = x;
            out.b = x.b + sg * y.b;
            for (auto &kv : y.c) out.c[kv.first] += sg * kv.second;
            break;
        }
        case ExprNode::MUL: {
            if (!x.constant() && !y.constant()) return "not linear in the unknowns (product of two unknown terms)";
            const Affine &k = x.constant() ? x : y, &a = x.constant() ? y : x;
            out.b = k.b * a.b;
            for (auto &kv : a.c) out.c[kv.first] = k.b * kv.second;
            break;
        }
        case ExprNode::DIV:
            if (!y.constant()) return "not linear in the unknowns (division by an unknown term)";
            out.b = x.b / y.b;
            for (auto &kv : x.c) out.c[kv.first] = kv.second / y.b;
            break;
        case ExprNode::POW:
            if (!x.constant() || !y.constant()) return "not linear in the unknowns (power of an unknown term)";
            out.b = std::pow(x.b, y.b);
            break;
        case ExprNode::FUN: {
            if (!x.constant()) return "not linear in the unknowns (" + n.fun + " of an unknown term)";
            const double v = x.b;
            out.b = n.fun == "sin" ? std::sin(v) : n.fun == "cos" ? std::cos(v) : n.fun == "tan" ? std::tan(v)
                  : n.fun == "exp" ? std::exp(v) : n.fun == "log" ? std::log(v) : n.fun == "sqrt" ? std::sqrt(v)
                  : std::fabs(v);
            break;
        }
    }
    return "";
}

}  // namespace

std::string ConstraintExpr::parse(const std::string &txt) {
    std::string s;
    for (char c : txt)
        if (c != ' ' && c != '\t') s += c;
    txt_ = s;
    if (s.empty()) return "empty constraint";
    std::string err;
    Parser p(s);
    root_ = p.parse(err, time_);
    return err;
}

std::string ConstraintExpr::eval(double t, std::map<uint64_t, double> &coeff, double &b) const {
    if (!root_) return "constraint not parsed";
    Affine a;
    const std::string e = evaluate(*root_, t, a);
    if (!e.empty()) return e;
    coeff.clear();
    for (auto &kv : a.c)
        if (kv.second != 0.0) coeff[kv.first] = kv.second;
    b = a.b;
    return "";
}

}  // namespace femgpu
