// Approximate minimum degree node order: the order symamd gives a node graph.
//
// symamd(m) (util/symamd.h:27-36, util/symamd.cpp:16-36) hands the node adjacency of the mesh
// (util/GetLinkForOrdering.h: every pair of nodes of an element, rows sorted) to the AMD 1.1
// package vendored in amd/ (amd_order.c, amd_1.c, amd_2.c, amd_postorder.c, amd_post_tree.c;
// P. Amestoy, T. Davis, I. Duff, "An approximate minimum degree ordering algorithm", SIAM J.
// Matrix Anal. Appl. 17(4):886-905, 1996, and "Algorithm 837: AMD", ACM TOMS 30(3), 2004) with
// its default controls: dense threshold 10 sqrt(n) (at least 16, at most n) and aggressive
// absorption.  The order depends on every tie-break of that method, so the steps below are its
// steps, in its order: a quotient graph of variables and elements; the pivot is the head of the
// lowest non-empty degree list; the new element gathers the pivot's elements' variables, then
// the pivot's own; |Le \ Lme| through a counter array; approximate external degrees with
// aggressive absorption and mass elimination; the pivot's element put first in each neighbour's
// list; supervariables found through hash buckets of the element and variable indices; degree
// lists refilled head first; an assembly-tree postorder with each node's largest child last.
// The storage is this file's own: one list per variable or element rather than one workspace
// that is garbage-collected (collection moves lists but never reorders them, so the order is the
// same).  Host only.
#include <algorithm>
#include <climits>
#include <cmath>

#include "host_mesh.hpp"

namespace femgpu {
namespace {

constexpr int NONE = -1;

class AmdOrdering {
  public:
    // adj[i]: the neighbours of i other than i, ascending
    explicit AmdOrdering(std::vector<std::vector<int>> adj) : n_((int)adj.size()), list_(std::move(adj)) {}

    // Node at each position of the order.
    std::vector<uint32_t> run() {
        const int n = n_;
        if (n == 0) return {};
        init();
        int nel = 0, mindeg = 0, lemax = 0;
        int dense = (int)(10.0 * std::sqrt((double)n));
        dense = std::min(n, std::max(16, dense));
        for (int i = 0; i < n; ++i) {
            const int d = deg_[i];
            if (d == 0) {          // no neighbour: an element of its own, ordered at once
                front_[i] = 1;
                w_[i] = 0;
                ++nel;
            } else if (d > dense) {  // dense: left out, ordered last
                nv_[i] = 0;
                elen_[i] = NONE;
                ++nel;
            } else {
                push_degree(i, d);
            }
        }
        std::vector<int> lme, kept_e, kept_v;
        while (nel < n) {
            // pivot: head of the lowest non-empty degree list
            int d = mindeg;
            while (dhead_[d] == NONE) ++d;
            mindeg = d;
            const int me = dhead_[d];
            remove_degree(me);
            const int elenme = elen_[me];
            int nvpiv = nv_[me];
            nel += nvpiv;
            nv_[me] = -nvpiv;
            int degme = 0;

            // new element: the variables of the pivot's elements (each then absorbed), then the
            // pivot's own variables; a variable enters once (its nv is negated while in Lme)
            lme.clear();
            auto take = [&](int i) {
                const int nvi = nv_[i];
                if (nvi > 0) {
                    degme += nvi;
                    nv_[i] = -nvi;
                    lme.push_back(i);
                    remove_degree(i);
                }
            };
            std::vector<int> &Lp = list_[me];
            for (int k = 0; k < elenme; ++k) {
                const int e = Lp[k];
                for (int i : list_[e]) take(i);
                parent_[e] = me;
                w_[e] = 0;
                std::vector<int>().swap(list_[e]);
            }
            for (size_t k = (size_t)elenme; k < Lp.size(); ++k) take(Lp[k]);
            deg_[me] = degme;
            front_[me] = nvpiv + degme;
            list_[me] = lme;
            elen_[me] = NONE;
            check_flag();

            // scan 1: w[e] - wflg = |Le \ Lme| for every element met in Lme's element lists
            for (int i : lme) {
                const int eln = elen_[i];
                if (eln <= 0) continue;
                const int nvi = -nv_[i];
                const int64_t wnvi = wflg_ - nvi;
                for (int k = 0; k < eln; ++k) {
                    const int e = list_[i][k];
                    int64_t we = w_[e];
                    if (we >= wflg_) we -= nvi;
                    else if (we != 0) we = deg_[e] + wnvi;
                    w_[e] = we;
                }
            }

            // scan 2: approximate degrees, aggressive absorption, mass elimination, hash buckets
            for (int i : lme) {
                std::vector<int> &Li = list_[i];
                const int eln = elen_[i];
                uint32_t hash = 0;
                int64_t dg = 0;
                kept_e.clear();
                kept_v.clear();
                for (int k = 0; k < eln; ++k) {
                    const int e = Li[k];
                    const int64_t we = w_[e];
                    if (we == 0) continue;
                    const int64_t dext = we - wflg_;
                    if (dext > 0) {
                        dg += dext;
                        kept_e.push_back(e);
                        hash += (uint32_t)e;
                    } else {  // Le inside Lme: e absorbed into me
                        parent_[e] = me;
                        w_[e] = 0;
                        std::vector<int>().swap(list_[e]);
                    }
                }
                for (size_t k = (size_t)eln; k < Li.size(); ++k) {
                    const int j = Li[k];
                    if (nv_[j] > 0) {
                        dg += nv_[j];
                        kept_v.push_back(j);
                        hash += (uint32_t)j;
                    }
                }
                if (kept_e.empty() && kept_v.empty()) {  // only me left: eliminated with it
                    parent_[i] = me;
                    const int nvi = -nv_[i];
                    degme -= nvi;
                    nvpiv += nvi;
                    nel += nvi;
                    nv_[i] = 0;
                    elen_[i] = NONE;
                    std::vector<int>().swap(Li);
                    continue;
                }
                deg_[i] = (int)std::min<int64_t>(deg_[i], dg);
                // me first; the first kept element moves to the end of the element part and
                // the first kept variable to the end of the list
                Li.clear();
                Li.push_back(me);
                Li.insert(Li.end(), kept_e.begin() + (kept_e.empty() ? 0 : 1), kept_e.end());
                if (!kept_e.empty()) Li.push_back(kept_e[0]);
                Li.insert(Li.end(), kept_v.begin() + (kept_v.empty() ? 0 : 1), kept_v.end());
                if (!kept_v.empty()) Li.push_back(kept_v[0]);
                elen_[i] = (int)kept_e.size() + 1;
                const uint32_t h = hash % (uint32_t)n;
                bnext_[i] = bhead_[h];
                bhead_[h] = i;
                hkey_[i] = h;
            }
            deg_[me] = degme;
            lemax = std::max(lemax, degme);
            wflg_ += lemax;
            check_flag();

            // supervariables: within each bucket of Lme's hash keys, j is absorbed into an earlier
            // i with the same list (after me) and the same element count
            for (int i0 : lme) {
                if (nv_[i0] >= 0) continue;
                const uint32_t h = hkey_[i0];
                int i = bhead_[h];
                bhead_[h] = NONE;
                while (i != NONE && bnext_[i] != NONE) {
                    const std::vector<int> &Li = list_[i];
                    const size_t ln = Li.size();
                    const int eln = elen_[i];
                    for (size_t p = 1; p < ln; ++p) w_[Li[p]] = wflg_;
                    int jlast = i, j = bnext_[i];
                    while (j != NONE) {
                        const std::vector<int> &Lj = list_[j];
                        bool same = Lj.size() == ln && elen_[j] == eln;
                        for (size_t p = 1; same && p < ln; ++p) same = w_[Lj[p]] == wflg_;
                        if (same) {
                            parent_[j] = i;
                            nv_[i] += nv_[j];
                            nv_[j] = 0;
                            elen_[j] = NONE;
                            std::vector<int>().swap(list_[j]);
                            j = bnext_[j];
                            bnext_[jlast] = j;
                        } else {
                            jlast = j;
                            j = bnext_[j];
                        }
                    }
                    ++wflg_;
                    i = bnext_[i];
                }
            }

            // back into the degree lists (head first, in Lme order); Lme keeps its principal
            // variables
            const int nleft = n - nel;
            std::vector<int> &Le = list_[me];
            size_t kept = 0;
            for (int i : lme) {
                const int nvi = -nv_[i];
                if (nvi <= 0) continue;
                nv_[i] = nvi;
                const int dnew = std::min(deg_[i] + degme - nvi, nleft - nvi);
                push_degree(i, dnew);
                mindeg = std::min(mindeg, dnew);
                Le[kept++] = i;
            }
            Le.resize(kept);
            nv_[me] = nvpiv;
            if (kept == 0) {  // nothing left of me: a root of the assembly tree
                parent_[me] = NONE;
                w_[me] = 0;
            }
        }
        return positions();
    }

  private:
    int n_;
    std::vector<std::vector<int>> list_;  // variable: its elements (elen_ of them), then variables;
                                          // element: its variables
    std::vector<int> nv_, elen_, deg_, parent_, front_, dhead_, dnext_, dprev_, bhead_, bnext_;
    std::vector<uint32_t> hkey_;
    std::vector<int64_t> w_;
    int64_t wflg_ = 2, wbig_ = 0;

    void init() {
        const int n = n_;
        nv_.assign(n, 1);
        elen_.assign(n, 0);
        deg_.resize(n);
        for (int i = 0; i < n; ++i) deg_[i] = (int)list_[i].size();
        parent_.assign(n, NONE);
        front_.assign(n, 0);
        dhead_.assign(n, NONE);
        dnext_.assign(n, NONE);
        dprev_.assign(n, NONE);
        bhead_.assign(n, NONE);
        bnext_.assign(n, NONE);
        hkey_.assign(n, 0);
        w_.assign(n, 1);
        wflg_ = 2;
        wbig_ = (int64_t)INT_MAX - n;
    }
    void push_degree(int i, int d) {
        const int h = dhead_[d];
        if (h != NONE) dprev_[h] = i;
        dnext_[i] = h;
        dprev_[i] = NONE;
        dhead_[d] = i;
        deg_[i] = d;
    }
    void remove_degree(int i) {
        const int p = dprev_[i], q = dnext_[i];
        if (q != NONE) dprev_[q] = p;
        if (p != NONE) dnext_[p] = q;
        else dhead_[deg_[i]] = q;
    }
    // the flag stays below INT_MAX - n, as in the package's int version
    void check_flag() {
        if (wflg_ < wbig_) return;
        for (int x = 0; x < n_; ++x)
            if (w_[x] != 0) w_[x] = 1;
        wflg_ = 2;
    }

    // Postorder of the assembly tree (children in index order, the one with the largest front
    // moved last), each element's absorbed variables just before it, dense variables last.
    std::vector<uint32_t> positions() {
        const int n = n_;
        for (int i = 0; i < n; ++i) {  // non-principal variables point at their element
            if (nv_[i] != 0 || parent_[i] == NONE) continue;
            int e = parent_[i];
            while (nv_[e] == 0) e = parent_[e];
            for (int j = i; nv_[j] == 0;) {
                const int next = parent_[j];
                parent_[j] = e;
                j = next;
            }
        }
        std::vector<int> child(n, NONE), sibling(n, NONE);
        for (int j = n - 1; j >= 0; --j)
            if (nv_[j] > 0 && parent_[j] != NONE) {
                sibling[j] = child[parent_[j]];
                child[parent_[j]] = j;
            }
        for (int i = 0; i < n; ++i) {
            if (nv_[i] <= 0 || child[i] == NONE) continue;
            int prev = NONE, big = NONE, bigprev = NONE, bigsize = NONE;
            for (int f = child[i]; f != NONE; f = sibling[f]) {
                if (front_[f] >= bigsize) {
                    bigsize = front_[f];
                    bigprev = prev;
                    big = f;
                }
                prev = f;
            }
            if (sibling[big] != NONE) {
                if (bigprev == NONE) child[i] = sibling[big];
                else sibling[bigprev] = sibling[big];
                sibling[big] = NONE;
                sibling[prev] = big;
            }
        }
        std::vector<int> post;  // elements in postorder
        post.reserve(n);
        std::vector<int> stack;
        for (int r = 0; r < n; ++r) {
            if (parent_[r] != NONE || nv_[r] <= 0) continue;
            stack.push_back(r);
            while (!stack.empty()) {
                const int x = stack.back();
                if (child[x] != NONE) {  // children first, in list order
                    const size_t base = stack.size();
                    for (int f = child[x]; f != NONE; f = sibling[f]) stack.push_back(f);
                    std::reverse(stack.begin() + base, stack.end());
                    child[x] = NONE;
                } else {
                    stack.pop_back();
                    post.push_back(x);
                }
            }
        }
        std::vector<int> pos(n, NONE);
        int next = 0;
        for (int e : post) {
            pos[e] = next;
            next += nv_[e];
        }
        for (int i = 0; i < n; ++i) {
            if (nv_[i] != 0) continue;
            const int e = parent_[i];
            if (e != NONE) pos[i] = pos[e]++;
            else pos[i] = next++;
        }
        std::vector<uint32_t> order(n);
        for (int i = 0; i < n; ++i) order[pos[i]] = (uint32_t)i;
        return order;
    }
};

}  // namespace

std::vector<uint32_t> amd_order(const Pattern &pat, uint32_t r0, uint32_t r1) {
    std::vector<std::vector<int>> adj(r1 - r0);
    for (uint32_t r = r0; r < r1; ++r) {
        std::vector<int> &a = adj[r - r0];
        for (uint32_t p = pat.row_ptr[r]; p < pat.row_ptr[r + 1]; ++p)
            if (pat.col[p] != r) a.push_back((int)(pat.col[p] - r0));
        std::sort(a.begin(), a.end());
    }
    std::vector<uint32_t> o = AmdOrdering(std::move(adj)).run();
    for (uint32_t &v : o) v += r0;
    return o;
}

}  // namespace femgpu
