// Host-side mesh services (see host_mesh.hpp).
#include "host_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>

namespace femgpu {

namespace {

// Reference-space gradients of the shape functions at xi (formulations/element_*.py).
void shape_grad(int kind, const double xi[3], double (*dN)[3]) {
    const double x = xi[0], y = xi[1], z = xi[2];
    if (kind == TETRA) {
        static const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        std::memcpy(dN, d, sizeof d);
    } else if (kind == TETRA10) {
        const double L[4] = {1 - x - y - z, x, y, z};
        static const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        static const int edge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) dN[a][d] = (4 * L[a] - 1) * dL[a][d];
        for (int k = 0; k < 6; ++k)
            for (int d = 0; d < 3; ++d)
                dN[4 + k][d] = 4 * (dL[edge[k][0]][d] * L[edge[k][1]] + L[edge[k][0]] * dL[edge[k][1]][d]);
    } else {
        for (int a = 0; a < 8; ++a) {
            const int bx = (a == 1 || a == 2 || a == 5 || a == 6), by = (a == 2 || a == 3 || a == 6 || a == 7), bz = a >= 4;
            const double X = bx ? x : 1 - x, Y = by ? y : 1 - y, Z = bz ? z : 1 - z;
            const double sx = bx ? 1 : -1, sy = by ? 1 : -1, sz = bz ? 1 : -1;
            dN[a][0] = sx * Y * Z; dN[a][1] = X * sy * Z; dN[a][2] = X * Y * sz;
        }
    }
}

// Sign fix applied by make_rect: det J at xi = (1/2,1/2,1/2) (make_elem.py:28-45), node
// permutation of formulations/element_<E>.py when negative.
void orient(int kind, const double *coords, uint32_t *v) {
    const int npe = nodes_per_elem(kind);
    double dN[10][3], J[3][3] = {{0}};
    const double half[3] = {0.5, 0.5, 0.5};
    shape_grad(kind, half, dN);
    for (int a = 0; a < npe; ++a)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) J[j][i] += dN[a][j] * coords[3 * (size_t)v[a] + i];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (det >= 0) return;
    static const int p4[4] = {0, 2, 1, 3}, p10[10] = {0, 1, 3, 2, 4, 8, 7, 6, 5, 9}, p8[8] = {0, 3, 2, 1, 4, 7, 6, 5};
    const int *p = kind == TETRA ? p4 : kind == TETRA10 ? p10 : p8;
    uint32_t t[10];
    for (int a = 0; a < npe; ++a) t[a] = v[p[a]];
    std::memcpy(v, t, sizeof(uint32_t) * npe);
}

template <class F>
void parallel_for(size_t n, F &&f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536) nt = 1;
    if (nt == 1) { f(size_t(0), n); return; }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back([&, t] { f(n * t / nt, n * (t + 1) / nt); });
    for (auto &x : th) x.join();
}

// node -> incident elements (CSR)
void node_elems(uint32_t n_nodes, uint32_t n_elem, int npe, const uint32_t *conn, std::vector<uint32_t> &ptr,
                std::vector<uint32_t> &lst) {
    ptr.assign((size_t)n_nodes + 1, 0);
    for (size_t s = 0; s < (size_t)n_elem * npe; ++s) ptr[conn[s] + 1]++;
    for (uint32_t n = 0; n < n_nodes; ++n) ptr[n + 1] += ptr[n];
    lst.resize((size_t)n_elem * npe);
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    for (uint32_t e = 0; e < n_elem; ++e)
        for (int a = 0; a < npe; ++a) lst[fill[conn[(size_t)e * npe + a]]++] = e;
}

}  // namespace

bool rect_counts(int kind, const uint32_t nb[3], uint32_t &n_nodes, uint32_t &n_elem) {
    if (nb[0] < 2 || nb[1] < 2 || nb[2] < 2) return false;
    const uint64_t cells = uint64_t(nb[0] - 1) * (nb[1] - 1) * (nb[2] - 1);
    uint64_t nn, ne;
    if (kind == TETRA) { nn = uint64_t(nb[0]) * nb[1] * nb[2]; ne = 6 * cells; }
    else if (kind == HEXA) { nn = uint64_t(nb[0]) * nb[1] * nb[2]; ne = cells; }
    else if (kind == TETRA10) { nn = uint64_t(2 * nb[0] - 1) * (2 * nb[1] - 1) * (2 * nb[2] - 1); ne = 6 * cells; }
    else return false;
    if (nn > 0xffffffffull || ne > 0xffffffffull) return false;
    n_nodes = (uint32_t)nn; n_elem = (uint32_t)ne;
    return true;
}

int make_rect(int kind, const double X0[3], const double X1[3], const uint32_t nb[3], double *coords, uint32_t *conn) {
    uint32_t nn, ne;
    if (!rect_counts(kind, nb, nn, ne)) return 1;
    if (kind == TETRA10) {
        // Tetra_10 (make_rect.h:443-544): six tetrahedra per cell, each node placed at
        // O + tr + incr*(chgt*rela[i]) and merged with an existing node closer than
        // 16*|X1-X0|*eps (BestialNodeAdder, make_rect.h:26-39), i.e. numbered by first
        // appearance.  Merged points are exactly the half-step lattice points, so the
        // lookup is a hash on the rounded half-step index.
        static const double rela[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {.5, 0, 0},
                                           {.5, .5, 0}, {0, .5, 0}, {0, 0, .5}, {.5, 0, .5}, {0, .5, .5}};
        static const double tr[6][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
        static const double basis[6][3][3] = {
            {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},   {{-1, 0, 0}, {0, -1, 0}, {0, -1, 1}}, {{-1, 0, 1}, {-1, 1, 0}, {0, 0, 1}},
            {{-1, 0, 0}, {-1, 1, 0}, {-1, 1, -1}}, {{-1, 0, 0}, {-1, 0, 1}, {0, -1, 1}}, {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}}};
        const uint32_t ne3[3] = {nb[0] - 1, nb[1] - 1, nb[2] - 1};
        double incr[3];
        for (int d = 0; d < 3; ++d) incr[d] = (X1[d] - X0[d]) / double(ne3[d]);
        const uint64_t Lx = 2ull * ne3[0] + 1, Ly = 2ull * ne3[1] + 1, Lz = 2ull * ne3[2] + 1;
        std::vector<uint32_t> id(Lx * Ly * Lz, 0xffffffffu);  // dense table over the half-step lattice
        uint32_t next = 0;
        size_t e = 0;
        for (uint32_t i = 0; i < ne3[2]; ++i)
            for (uint32_t j = 0; j < ne3[1]; ++j)
                for (uint32_t k = 0; k < ne3[0]; ++k) {
                    const double O[3] = {X0[0] + double(k) * incr[0], X0[1] + double(j) * incr[1], X0[2] + double(i) * incr[2]};
                    for (int t = 0; t < 6; ++t) {
                        uint32_t v[10];
                        for (int r = 0; r < 10; ++r) {
                            double p[3];
                            int64_t q[3];
                            for (int d = 0; d < 3; ++d) {
                                double s = 0;
                                for (int c = 0; c < 3; ++c) s += basis[t][c][d] * rela[r][c];
                                const double base = t == 0 ? O[d] : O[d] + tr[t][d] * incr[d];
                                p[d] = base + incr[d] * s;
                                q[d] = std::llround((p[d] - X0[d]) / (0.5 * incr[d]));
                            }
                            uint32_t &slot = id[(uint64_t(q[2]) * Ly + uint64_t(q[1])) * Lx + uint64_t(q[0])];
                            if (slot == 0xffffffffu) {
                                slot = next++;
                                for (int d = 0; d < 3; ++d) coords[3 * (size_t)slot + d] = p[d];
                            }
                            v[r] = slot;
                        }
                        orient(TETRA10, coords, v);
                        std::memcpy(conn + 10 * e++, v, sizeof v);
                    }
                }
        return next == nn ? 0 : 5;
    }
    // Tetra / Hexa node lattice: Rectilinear_iterator running sums (util/rectilinear_iterator.h:40-51),
    // x fastest; the last point of each axis is the accumulated sum, not X1.
    double inc[3];
    const uint32_t n3[3] = {nb[0], nb[1], nb[2]};
    std::vector<std::vector<double>> axis(3);
    for (int d = 0; d < 3; ++d) {
        inc[d] = (X1[d] - X0[d]) / double(n3[d] - 1);
        axis[d].resize(n3[d]);
        double s = X0[d];
        for (uint32_t i = 0; i < n3[d]; ++i) { axis[d][i] = s; s += inc[d]; }
    }
    for (uint32_t z = 0; z < n3[2]; ++z)
        for (uint32_t y = 0; y < n3[1]; ++y)
            for (uint32_t x = 0; x < n3[0]; ++x) {
                const size_t n = (size_t(z) * n3[1] + y) * n3[0] + x;
                coords[3 * n] = axis[0][x]; coords[3 * n + 1] = axis[1][y]; coords[3 * n + 2] = axis[2][z];
            }
    const uint32_t nx = n3[0], nxy = n3[0] * n3[1];
    size_t e = 0;
    for (uint32_t i = 0; i + 1 < n3[2]; ++i)
        for (uint32_t j = 0; j + 1 < n3[1]; ++j)
            for (uint32_t k = 0; k + 1 < n3[0]; ++k) {
                const uint32_t b = i * nxy + j * nx + k;
                if (kind == HEXA) {  // make_rect.h:665-678
                    const uint32_t v[8] = {b, b + 1, b + 1 + nx, b + nx, b + nxy, b + nxy + 1, b + nxy + 1 + nx, b + nxy + nx};
                    std::memcpy(conn + 8 * e++, v, sizeof v);
                } else {  // make_rect.h:400-428: six tetrahedra from cell corners (bits x=1, y=2, z=4)
                    static const uint8_t corners[6][4] = {{0, 1, 2, 4}, {1, 2, 3, 5}, {1, 2, 4, 5}, {2, 4, 5, 6}, {2, 3, 5, 6}, {3, 5, 6, 7}};
                    for (int t = 0; t < 6; ++t) {
                        uint32_t v[4];
                        for (int c = 0; c < 4; ++c) {
                            const int m = corners[t][c];
                            v[c] = b + ((m >> 2) & 1) * nxy + ((m >> 1) & 1) * nx + (m & 1);
                        }
                        orient(TETRA, coords, v);
                        std::memcpy(conn + 4 * e++, v, sizeof v);
                    }
                }
            }
    return 0;
}

std::vector<uint32_t> skin_faces(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn) {
    // child faces (formulations/element_Tetra.py:17-22, element_Tetra_10.py:17-22, element_Hexa.py:22-29)
    static const int ft[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
    static const int f10[4][6] = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 8, 7}, {0, 3, 2, 7, 9, 6}, {1, 2, 3, 5, 9, 8}};
    static const int fh[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 4, 7, 3}, {1, 2, 6, 5}};
    const int npe = nodes_per_elem(kind), nfe = kind == HEXA ? 6 : 4, nv = kind == HEXA ? 4 : 3;
    const int npf = nodes_per_face(face_kind_of(kind));
    auto local = [&](int f, int v) { return kind == TETRA ? ft[f][v] : kind == TETRA10 ? f10[f][v] : fh[f][v]; };
    struct Key { uint32_t v[4]; uint64_t owner; };
    std::vector<Key> keys((size_t)n_elem * nfe);
    for (uint32_t e = 0; e < n_elem; ++e)
        for (int f = 0; f < nfe; ++f) {
            Key &k = keys[(size_t)e * nfe + f];
            k.v[3] = 0;
            for (int v = 0; v < nv; ++v) k.v[v] = conn[(size_t)e * npe + local(f, v)];
            std::sort(k.v, k.v + nv);
            k.owner = (size_t)e * nfe + f;
        }
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        return std::lexicographical_compare(a.v, a.v + 4, b.v, b.v + 4) || (std::equal(a.v, a.v + 4, b.v) && a.owner < b.owner);
    });
    std::vector<uint8_t> single(keys.size(), 0);
    for (size_t s = 0; s < keys.size();) {
        size_t t = s + 1;
        while (t < keys.size() && std::equal(keys[t].v, keys[t].v + 4, keys[s].v)) ++t;
        if (t - s == 1) single[keys[s].owner] = 1;
        s = t;
    }
    std::vector<uint32_t> out;
    for (size_t o = 0; o < single.size(); ++o)
        if (single[o]) {
            const uint32_t e = uint32_t(o / nfe);
            const int f = int(o % nfe);
            for (int v = 0; v < npf; ++v) out.push_back(conn[(size_t)e * npe + local(f, v)]);
        }
    (void)n_nodes;
    return out;
}

Pattern build_pattern(uint32_t n_nodes, uint32_t n_elem, int npe, const uint32_t *conn) {
    // ResizeMat::GetNZ (formulation.h:200-230): the block (a,b) of every ordered node pair of
    // every element; a row without any entry keeps its diagonal; rows sorted.
    std::vector<uint32_t> ptr, lst;
    node_elems(n_nodes, n_elem, npe, conn, ptr, lst);
    Pattern P;
    P.n_rows = n_nodes;
    std::vector<uint32_t> cnt(n_nodes);
    std::vector<std::vector<uint32_t>> rows(n_nodes);
    parallel_for(n_nodes, [&](size_t b, size_t e) {
        std::vector<uint32_t> buf;
        for (size_t n = b; n < e; ++n) {
            buf.clear();
            for (uint32_t s = ptr[n]; s < ptr[n + 1]; ++s)
                for (int a = 0; a < npe; ++a) buf.push_back(conn[(size_t)lst[s] * npe + a]);
            if (buf.empty()) buf.push_back((uint32_t)n);
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            rows[n] = buf;
        }
    });
    P.row_ptr.resize((size_t)n_nodes + 1);
    P.row_ptr[0] = 0;
    for (uint32_t n = 0; n < n_nodes; ++n) P.row_ptr[n + 1] = P.row_ptr[n] + (uint32_t)rows[n].size();
    P.col.resize(P.row_ptr[n_nodes]);
    parallel_for(n_nodes, [&](size_t b, size_t e) {
        for (size_t n = b; n < e; ++n) {
            std::copy(rows[n].begin(), rows[n].end(), P.col.begin() + P.row_ptr[n]);
            std::vector<uint32_t>().swap(rows[n]);
        }
    });
    finish_pattern(P);
    return P;
}

uint64_t sell_size(const Pattern &P, const std::vector<uint32_t> &perm) {
    const uint32_t n = P.n_rows;
    uint64_t tot = 0;
    for (uint32_t t0 = 0; t0 < n; t0 += 64) {
        uint32_t w = 0;
        for (uint32_t t = t0; t < std::min(n, t0 + 64); ++t) w = std::max(w, P.len(perm.empty() ? t : perm[t]));
        tot += 64ull * w;
    }
    return tot;
}

void finish_pattern(Pattern &P) {
    const uint32_t n = P.n_rows, ns = (n + 63) / 64;
    P.sell_ptr.assign((size_t)ns + 1, 0);
    for (uint32_t s = 0; s < ns; ++s) {
        uint32_t w = 0;
        for (uint32_t t = 64 * s; t < std::min(n, 64 * s + 64); ++t) w = std::max(w, P.len(P.row_of(t)));
        P.sell_ptr[s + 1] = P.sell_ptr[s] + 64 * w;
    }
    P.n_sell = P.sell_ptr[ns];
    P.diag.assign(n, 0);
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t *b = P.col.data() + P.row_ptr[r], *e = P.col.data() + P.row_ptr[r + 1];
        P.diag[r] = (uint32_t)P.pos(r, uint32_t(std::lower_bound(b, e, r) - b));
    }
}

std::vector<uint32_t> length_order(const Pattern &P, const std::vector<uint32_t> &part, uint32_t window,
                                   double min_gain) {
    const uint32_t n = P.n_rows;
    if (part.size() < 2 || part.front() != 0 || part.back() != n || window == 0) return {};
    std::vector<uint32_t> perm(n);
    for (uint32_t t = 0; t < n; ++t) perm[t] = t;
    bool changed = false;
    auto slices = [&](uint32_t a, uint32_t b, bool sorted) {  // positions of rows a..b as their own slices
        uint64_t tot = 0;
        for (uint32_t t0 = a; t0 < b; t0 += 64) {
            uint32_t w = 0;
            for (uint32_t t = t0; t < std::min(b, t0 + 64); ++t) w = std::max(w, P.len(sorted ? perm[t] : t));
            tot += 64ull * w;
        }
        return tot;
    };
    for (size_t s = 0; s + 1 < part.size(); ++s) {
        const uint32_t a = part[s], b = part[s + 1];
        if (a > b) return {};
        for (uint32_t w0 = a; w0 < b; w0 += window)
            std::stable_sort(perm.begin() + w0, perm.begin() + std::min(b, w0 + window),
                             [&](uint32_t x, uint32_t y) { return P.len(x) > P.len(y); });
        bool same = true;
        for (uint32_t t = a; t < b && same; ++t) same = perm[t] == t;
        const bool keep = !same && (min_gain < 0 || (double)slices(a, b, true) < (1.0 - min_gain) * (double)slices(a, b, false));
        if (!keep)
            for (uint32_t t = a; t < b; ++t) perm[t] = t;
        changed |= keep;
    }
    if (!changed) return {};
    return perm;
}

bool sort_rows_by_length(Pattern &P, const std::vector<uint32_t> &part, uint32_t window, double min_gain) {
    std::vector<uint32_t> perm = length_order(P, part, window, min_gain);
    if (perm.empty()) return false;
    const uint32_t n = P.n_rows;
    P.iperm.assign(n, 0);
    for (uint32_t t = 0; t < n; ++t) P.iperm[perm[t]] = t;
    P.perm = std::move(perm);
    finish_pattern(P);
    return true;
}

std::vector<uint32_t> sell_columns(const Pattern &P) {
    std::vector<uint32_t> c(std::max<uint64_t>(P.n_sell, 1));
    for (uint32_t t = 0; t < P.n_rows; ++t) {
        const uint32_t r = P.row_of(t), s = t >> 6, w = (P.sell_ptr[s + 1] - P.sell_ptr[s]) / 64, len = P.len(r);
        for (uint32_t k = 0; k < w; ++k) c[P.pos(r, k)] = k < len ? P.col[P.row_ptr[r] + k] : r;
    }
    return c;
}

std::vector<uint32_t> build_slot_map(const Pattern &pat, uint32_t n_elem, int npe, const uint32_t *conn) {
    std::vector<uint32_t> slot((size_t)n_elem * npe * npe);
    parallel_for(n_elem, [&](size_t b, size_t e) {
        for (size_t el = b; el < e; ++el) {
            const uint32_t *v = conn + el * npe;
            for (int a = 0; a < npe; ++a) {
                const uint32_t *r0 = pat.col.data() + pat.row_ptr[v[a]], *r1 = pat.col.data() + pat.row_ptr[v[a] + 1];
                for (int c = 0; c < npe; ++c)
                    slot[(el * npe + a) * npe + c] = (uint32_t)pat.pos(v[a], uint32_t(std::lower_bound(r0, r1, v[c]) - r0));
            }
        }
    });
    return slot;
}

Coloring color_elements(uint32_t n_nodes, uint32_t n_elem, int npe, const uint32_t *conn) {
    std::vector<uint32_t> ptr, lst;
    node_elems(n_nodes, n_elem, npe, conn, ptr, lst);
    std::vector<uint32_t> color(n_elem, 0xffffffffu);
    std::vector<uint64_t> seen;  // bitmask per 64 colours
    uint32_t nc = 0;
    for (uint32_t e = 0; e < n_elem; ++e) {
        seen.assign(nc / 64 + 1, 0);
        for (int a = 0; a < npe; ++a) {
            const uint32_t n = conn[(size_t)e * npe + a];
            for (uint32_t s = ptr[n]; s < ptr[n + 1]; ++s) {
                const uint32_t c = color[lst[s]];
                if (c != 0xffffffffu) seen[c / 64] |= 1ull << (c % 64);
            }
        }
        uint32_t c = 0;
        while (c / 64 < seen.size() && (seen[c / 64] >> (c % 64) & 1)) ++c;
        color[e] = c;
        nc = std::max(nc, c + 1);
    }
    Coloring C;
    C.color_ptr.assign(nc + 1, 0);
    for (uint32_t e = 0; e < n_elem; ++e) C.color_ptr[color[e] + 1]++;
    for (uint32_t c = 0; c < nc; ++c) C.color_ptr[c + 1] += C.color_ptr[c];
    C.order.resize(n_elem);
    std::vector<uint32_t> fill(C.color_ptr.begin(), C.color_ptr.end() - 1);
    for (uint32_t e = 0; e < n_elem; ++e) C.order[fill[color[e]]++] = e;
    return C;
}

RowGroups build_row_groups(uint32_t n_rows, uint32_t R, uint32_t n_elem, int npe, const uint32_t *conn,
                           const Coloring &col, const uint32_t *iperm) {
    RowGroups G;
    G.R = R;
    G.n_groups = (n_rows + R - 1) / R;
    G.n_colors = col.n_colors();
    const uint32_t nc = G.n_colors;
    std::vector<uint32_t> colour(n_elem), rank(n_elem);
    for (uint32_t c = 0; c < nc; ++c)
        for (uint32_t i = col.color_ptr[c]; i < col.color_ptr[c + 1]; ++i) {
            colour[col.order[i]] = c;
            rank[col.order[i]] = i;  // position in (colour, element) order
        }
    // (group, element) pairs, deduplicated per element
    std::vector<uint32_t> cnt((size_t)G.n_groups * nc + 1, 0);
    std::vector<uint32_t> gl;
    gl.reserve(4 * (size_t)n_elem);
    std::vector<uint64_t> pairs;
    pairs.reserve(4 * (size_t)n_elem);
    for (uint32_t e = 0; e < n_elem; ++e) {
        uint32_t gs[10];
        int ng = 0;
        for (int a = 0; a < npe; ++a) {
            const uint32_t v = conn[(size_t)e * npe + a], g = (iperm ? iperm[v] : v) / R;
            bool seen = false;
            for (int k = 0; k < ng; ++k) seen |= gs[k] == g;
            if (!seen) gs[ng++] = g;
        }
        for (int k = 0; k < ng; ++k) pairs.push_back(((uint64_t)gs[k] << 32) | rank[e]);
    }
    std::sort(pairs.begin(), pairs.end());
    G.elems.resize(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) {
        const uint32_t g = uint32_t(pairs[i] >> 32), e = col.order[uint32_t(pairs[i])];
        G.elems[i] = e;
        cnt[(size_t)g * nc + colour[e] + 1]++;
    }
    for (size_t i = 0; i + 1 < cnt.size(); ++i) cnt[i + 1] += cnt[i];
    // ptr[G*(nc+1)+c]: start of colour c in group G (colour nc = end of the group)
    G.ptr.assign((size_t)G.n_groups * (nc + 1), 0);
    for (uint32_t g = 0; g < G.n_groups; ++g)
        for (uint32_t c = 0; c <= nc; ++c) G.ptr[(size_t)g * (nc + 1) + c] = cnt[(size_t)g * nc + c];
    return G;
}

std::vector<uint32_t> cm_order(const Pattern &pat, uint32_t r0, uint32_t r1) {
    std::vector<uint32_t> order;
    if (r1 <= r0) return order;
    order.reserve(r1 - r0);
    auto degree = [&](uint32_t v) { return pat.len(v) - 1; };   // the pattern row holds v itself
    std::vector<uint8_t> seen(r1 - r0, 0);
    std::vector<uint32_t> nb;
    for (;;) {
        uint32_t start = r1;
        for (uint32_t v = r0; v < r1; ++v)
            if (!seen[v - r0] && (start == r1 || degree(v) < degree(start))) start = v;
        if (start == r1) break;
        size_t head = order.size();
        order.push_back(start);
        seen[start - r0] = 1;
        while (head < order.size()) {
            const uint32_t v = order[head++];
            nb.clear();
            for (uint32_t p = pat.row_ptr[v]; p < pat.row_ptr[v + 1]; ++p) {
                const uint32_t c = pat.col[p];
                if (c != v && c >= r0 && c < r1 && !seen[c - r0]) {
                    seen[c - r0] = 1;
                    nb.push_back(c);
                }
            }
            std::stable_sort(nb.begin(), nb.end(), [&](uint32_t a, uint32_t b) { return degree(a) < degree(b); });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    return order;
}

std::vector<uint32_t> balance_rows(const std::vector<uint32_t> &row_ptr, uint32_t n_parts) {
    const uint32_t n = (uint32_t)row_ptr.size() - 1;
    std::vector<uint32_t> cut(n_parts + 1);
    const double tot = double(row_ptr[n]) + n;  // blocks + one unit per row for the row overhead
    uint32_t r = 0;
    for (uint32_t k = 0; k <= n_parts; ++k) {
        const double target = tot * k / n_parts;
        while (r < n && double(row_ptr[r]) + r < target) ++r;
        cut[k] = r;
    }
    cut[n_parts] = n;
    return cut;
}

bool build_sym_sell(const Pattern &pat, SymSell &S) {
    const uint32_t n = pat.n_rows, ns = (n + 63) / 64;
    const uint32_t *rp = pat.row_ptr.data(), *col = pat.col.data();
    S = SymSell();
    S.lenU.resize(n);
    S.lenL.resize(n);
    std::vector<uint32_t> dslot(n);  // CSR slot of the diagonal
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t *b = col + rp[r], *e = col + rp[r + 1];
        const uint32_t d = uint32_t(std::lower_bound(b, e, r) - b);
        if (b + d == e || b[d] != r) return false;
        dslot[r] = d;
        S.lenL[r] = d;
        S.lenU[r] = rp[r + 1] - rp[r] - d;
    }
    S.sptrU.assign((size_t)ns + 1, 0);
    S.sptrL.assign((size_t)ns + 1, 0);
    for (uint32_t s = 0; s < ns; ++s) {
        uint32_t wu = 0, wl = 0;
        for (uint32_t r = 64 * s; r < std::min(n, 64 * s + 64); ++r) {
            wu = std::max(wu, S.lenU[r]);
            wl = std::max(wl, S.lenL[r]);
        }
        if (uint64_t(S.sptrU[s]) + 64ull * wu >= (1ull << 32) || uint64_t(S.sptrL[s]) + 64ull * wl >= (1ull << 32)) return false;
        S.sptrU[s + 1] = S.sptrU[s] + 64 * wu;
        S.sptrL[s + 1] = S.sptrL[s] + 64 * wl;
    }
    S.nU = S.sptrU[ns];
    S.nL = S.sptrL[ns];
    S.colU.assign(std::max<uint64_t>(S.nU, 1), 0);
    S.srcU.assign(std::max<uint64_t>(S.nU, 1), 0);
    S.colL.assign(std::max<uint64_t>(S.nL, 1), 0);
    S.tposL.assign(std::max<uint64_t>(S.nL, 1), 0);
    std::vector<uint8_t> bad(n, 0);
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t ri = b; ri < e; ++ri) {
            const uint32_t r = (uint32_t)ri, s = r >> 6;
            const uint32_t wu = (S.sptrU[s + 1] - S.sptrU[s]) / 64, wl = (S.sptrL[s + 1] - S.sptrL[s]) / 64;
            for (uint32_t k = 0; k < wu; ++k) {  // padding: own row, zero values (src = the diagonal)
                const uint64_t q = S.posU(r, k);
                const bool real = k < S.lenU[r];
                S.colU[q] = real ? col[rp[r] + dslot[r] + k] : r;
                S.srcU[q] = real ? (uint32_t)pat.pos(r, dslot[r] + k) : 0xffffffffu;
            }
            for (uint32_t k = 0; k < wl; ++k) {
                const uint64_t q = S.posL(r, k);
                if (k >= S.lenL[r]) { S.colL[q] = r; S.tposL[q] = 0xffffffffu; continue; }
                const uint32_t c = col[rp[r] + k];
                // row c holds r among its upper columns
                const uint32_t *cb = col + rp[c] + dslot[c], *ce = col + rp[c + 1];
                const uint32_t *f = std::lower_bound(cb, ce, r);
                if (f == ce || *f != r) { bad[r] = 1; continue; }
                S.colL[q] = c;
                S.tposL[q] = (uint32_t)S.posU(c, uint32_t(f - cb));
            }
        }
    });
    for (uint32_t r = 0; r < n; ++r)
        if (bad[r]) return false;
    return true;
}

bool build_ic0_schedule(const Pattern &pat, int ordering, Ic0Schedule &S, const std::vector<uint32_t> *base) {
    const uint32_t n = pat.n_rows;
    const uint32_t *rp = pat.row_ptr.data(), *col = pat.col.data();
    S = Ic0Schedule();
    S.ordering = ordering;
    S.n_rows = n;
    // elimination order
    std::vector<uint32_t> elim0(n), order0(n);
    if (ordering == IC0_COLOURED) {
        std::vector<uint32_t> colour(n, 0xffffffffu);
        std::vector<uint64_t> seen;
        uint32_t nc = 0;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t r = base ? (*base)[k] : k;
            seen.assign(nc / 64 + 1, 0);
            for (uint32_t p = rp[r]; p < rp[r + 1]; ++p) {
                const uint32_t c = colour[col[p]];
                if (c != 0xffffffffu) seen[c / 64] |= 1ull << (c % 64);
            }
            uint32_t c = 0;
            while (c / 64 < seen.size() && (seen[c / 64] >> (c % 64) & 1)) ++c;
            colour[r] = c;
            nc = std::max(nc, c + 1);
        }
        std::vector<uint32_t> cnt(nc + 1, 0);
        for (uint32_t r = 0; r < n; ++r) cnt[colour[r] + 1]++;
        for (uint32_t c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t r = base ? (*base)[k] : k;
            order0[cnt[colour[r]]++] = r;
        }
    } else {
        for (uint32_t k = 0; k < n; ++k) order0[k] = base ? (*base)[k] : k;
    }
    for (uint32_t i = 0; i < n; ++i) elim0[order0[i]] = i;
    // levels in elimination order
    std::vector<uint32_t> level(n, 0);
    uint32_t nl = n ? 1 : 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t r = order0[i];
        uint32_t l = 0;
        for (uint32_t p = rp[r]; p < rp[r + 1]; ++p)
            if (elim0[col[p]] < i) l = std::max(l, level[col[p]] + 1);
        level[r] = l;
        nl = std::max(nl, l + 1);
    }
    // rows t by (level, elimination position)
    S.lv_ptr.assign(nl + 1, 0);
    for (uint32_t r = 0; r < n; ++r) S.lv_ptr[level[r] + 1]++;
    for (uint32_t l = 0; l < nl; ++l) S.lv_ptr[l + 1] += S.lv_ptr[l];
    S.order.resize(n);
    {
        std::vector<uint32_t> fill(S.lv_ptr.begin(), S.lv_ptr.end() - 1);
        for (uint32_t i = 0; i < n; ++i) S.order[fill[level[order0[i]]]++] = order0[i];
    }
    std::vector<uint32_t> tof(n);
    for (uint32_t t = 0; t < n; ++t) tof[S.order[t]] = t;
    // rows: neighbours sorted by t
    S.len.resize(n);
    S.nlow.resize(n);
    const uint32_t ns = (n + 63) / 64;
    S.sptr.assign((size_t)ns + 1, 0);
    for (uint32_t t = 0; t < n; ++t) {
        const uint32_t r = S.order[t];
        S.len[t] = rp[r + 1] - rp[r];
        uint32_t lo = 0;
        for (uint32_t p = rp[r]; p < rp[r + 1]; ++p) lo += tof[col[p]] < t;
        S.nlow[t] = lo;
    }
    for (uint32_t s = 0; s < ns; ++s) {
        uint32_t w = 0;
        for (uint32_t t = 64 * s; t < std::min(n, 64 * s + 64); ++t) w = std::max(w, S.len[t]);
        S.sptr[s + 1] = S.sptr[s] + 64 * w;
    }
    S.n_sell = S.sptr[ns];
    const size_t np = std::max<uint64_t>(S.n_sell, 1);
    S.colN.assign(np, 0);
    S.colT.assign(np, 0);
    S.kpos.assign(np, 0);
    S.tpos.assign(np, 0);
    parallel_for(n, [&](size_t b, size_t e) {
        std::vector<std::pair<uint32_t, uint32_t>> nb;  // (t of neighbour, CSR slot in row r)
        for (size_t t = b; t < e; ++t) {
            const uint32_t r = S.order[t], s = (uint32_t)t >> 6, w = (S.sptr[s + 1] - S.sptr[s]) / 64;
            nb.clear();
            for (uint32_t p = rp[r]; p < rp[r + 1]; ++p) nb.emplace_back(tof[col[p]], p - rp[r]);
            std::sort(nb.begin(), nb.end());
            for (uint32_t k = 0; k < w; ++k) {
                const uint64_t q = S.pos((uint32_t)t, k);
                if (k < nb.size()) {
                    S.colT[q] = nb[k].first;
                    S.colN[q] = S.order[nb[k].first];
                    S.kpos[q] = (uint32_t)pat.pos(r, nb[k].second);
                } else {  // padding: own row, never read
                    S.colT[q] = (uint32_t)t;
                    S.colN[q] = r;
                    S.kpos[q] = pat.diag[r];
                }
            }
        }
    });
    // transposed slot of every lower block: row c holds t among its upper neighbours (the
    // pattern must be symmetric)
    std::vector<uint8_t> bad(n, 0);
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t t = b; t < e; ++t)
            for (uint32_t k = 0; k < S.nlow[t]; ++k) {
                const uint64_t q = S.pos((uint32_t)t, k);
                const uint32_t c = S.colT[q];
                uint32_t lo = S.nlow[c] + 1, hi = S.len[c];
                while (lo < hi) {
                    const uint32_t m = (lo + hi) / 2;
                    if (S.colT[S.pos(c, m)] < t) lo = m + 1; else hi = m;
                }
                if (lo == S.len[c] || S.colT[S.pos(c, lo)] != t) { bad[t] = 1; continue; }
                S.tpos[q] = (uint32_t)S.pos(c, lo);
            }
    });
    for (uint32_t t = 0; t < n; ++t)
        if (bad[t]) return false;
    return true;
}

}  // namespace femgpu
