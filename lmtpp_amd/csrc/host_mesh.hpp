// Host-side mesh services of the MI355X path: make_rect lattices, skin faces,
// the 3x3-block pattern, element colouring and the element->block scatter map.
// Pure C++ (no HIP); the device side consumes the arrays built here.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace femgpu {

enum ElemKind { TETRA = 0, TETRA10 = 1, HEXA = 2 };
enum FaceKind { TRI3 = 0, TRI6 = 1, QUAD4 = 2 };

inline int nodes_per_elem(int k) { return k == TETRA ? 4 : k == TETRA10 ? 10 : k == HEXA ? 8 : 0; }
inline int face_kind_of(int k) { return k == TETRA ? TRI3 : k == TETRA10 ? TRI6 : QUAD4; }
inline int nodes_per_face(int f) { return f == TRI3 ? 3 : f == TRI6 ? 6 : f == QUAD4 ? 4 : 0; }

// make_rect (mesh/make_rect.h): counts and lattice.
bool rect_counts(int kind, const uint32_t nb[3], uint32_t &n_nodes, uint32_t &n_elem);
int make_rect(int kind, const double X0[3], const double X1[3], const uint32_t nb[3], double *coords, uint32_t *conn);

// Faces with exactly one parent element, in (element, local face) order.
std::vector<uint32_t> skin_faces(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn);

// Full symmetric 3x3-block pattern: node adjacency, rows sorted, every row holds its diagonal.
// Device storage of the blocks is sliced ELL (SELL-64): storage rows t are grouped in slices of
// 64 consecutive t, slice s is as wide as its longest row, and block k of the row stored at t
// sits at position sell_ptr[t/64] + 64 k + t%64, so the 64 rows of a slice read their k-th blocks
// from 64 consecutive positions.  Storage row t holds block row perm[t]; perm and its inverse
// iperm are empty for the identity (t = r) and set by sort_rows_by_length.
struct Pattern {
    uint32_t n_rows = 0;
    std::vector<uint32_t> row_ptr, col;   // CSR view (the API and the reference's row order)
    std::vector<uint32_t> sell_ptr;       // n_slices + 1 slice starts, in block positions
    std::vector<uint32_t> diag;           // diag[r] = SELL position of block (r,r)
    std::vector<uint32_t> perm, iperm;    // storage row -> block row, block row -> storage row
    uint64_t n_sell = 0;                  // positions including padding
    uint32_t t_of(uint32_t r) const { return iperm.empty() ? r : iperm[r]; }
    uint32_t row_of(uint32_t t) const { return perm.empty() ? t : perm[t]; }
    uint32_t len(uint32_t r) const { return row_ptr[r + 1] - row_ptr[r]; }
    uint64_t pos(uint32_t r, uint32_t k) const {
        const uint32_t t = t_of(r);
        return sell_ptr[t >> 6] + 64ull * k + (t & 63u);
    }
    uint64_t pos_of_slot(uint32_t r, uint32_t slot) const { return pos(r, slot - row_ptr[r]); }
};
Pattern build_pattern(uint32_t n_nodes, uint32_t n_elem, int npe, const uint32_t *conn);
// Fills sell_ptr, n_sell and diag from row_ptr / col and the storage order.
void finish_pattern(Pattern &P);
// Positions the slices would take with rows stored in the order perm (empty: identity).
uint64_t sell_size(const Pattern &P, const std::vector<uint32_t> &perm);
// SELL-C-sigma: the rows of every part [part[i], part[i+1]) (parts cover 0..n_rows in increasing
// order: the sub-problems) are cut into windows of `window` rows from the part's start, and the
// rows of a window are stored by decreasing length (ties by row), so a slice holds rows of
// similar length and pads less.  A part keeps the sorted order when it saves more than min_gain of
// the positions its rows take as slices of their own (min_gain < 0: always), so a mesh gets the
// same order alone or in a batch.  Returns whether the storage order changed.
bool sort_rows_by_length(Pattern &P, const std::vector<uint32_t> &part, uint32_t window, double min_gain);
// The order sort_rows_by_length would set (needs only n_rows and row_ptr); empty for the identity.
std::vector<uint32_t> length_order(const Pattern &P, const std::vector<uint32_t> &part, uint32_t window,
                                   double min_gain);
// Column of every SELL position (padding positions hold their own row, with zero values).
std::vector<uint32_t> sell_columns(const Pattern &P);

// SELL position of block (conn[e][a], conn[e][b]) for every element and local pair: n_elem * npe * npe.
std::vector<uint32_t> build_slot_map(const Pattern &pat, uint32_t n_elem, int npe, const uint32_t *conn);

// Greedy first-fit colouring: no two elements of one colour share a node.
// order = elements sorted by colour (stable), color_ptr = offsets (n_colors + 1).
struct Coloring {
    std::vector<uint32_t> order, color_ptr;
    uint32_t n_colors() const { return color_ptr.empty() ? 0 : (uint32_t)color_ptr.size() - 1; }
};
Coloring color_elements(uint32_t n_nodes, uint32_t n_elem, int npe, const uint32_t *conn);

// Row-group assembly work lists: storage rows are cut into groups of R consecutive rows; group G
// lists the elements incident to its block rows, sorted by (colour, element), colour c of group G
// spanning elems[ptr[G*(nc+1)+c] .. ptr[G*(nc+1)+c+1]).  iperm: storage row of each block row
// (nullptr: identity).
struct RowGroups {
    uint32_t R = 0, n_groups = 0, n_colors = 0;
    std::vector<uint32_t> ptr, elems;
};
RowGroups build_row_groups(uint32_t n_rows, uint32_t R, uint32_t n_elem, int npe, const uint32_t *conn,
                           const Coloring &col, const uint32_t *iperm = nullptr);

// Cuthill-McKee visiting order of the nodes of [r0, r1) (rows with no coupling outside the range)
// as symrcm (util/symrcm.cpp:38-101) builds it: breadth-first from the node of lowest degree (the
// first one), each node's unvisited neighbours appended in order of increasing degree, ties in
// node order; a component the search did not reach restarts from its lowest-degree node (the
// reference numbers only the first component).  Degrees count neighbours other than the node.
std::vector<uint32_t> cm_order(const Pattern &pat, uint32_t r0, uint32_t r1);

// Approximate minimum degree order of the nodes of [r0, r1) (rows with no coupling outside the
// range): the order symamd (util/symamd.h:27-36) computes with the AMD 1.1 package it vendors,
// default controls.  Returns the node at each position.
std::vector<uint32_t> amd_order(const Pattern &pat, uint32_t r0, uint32_t r1);

// Contiguous row ranges, balanced by stored blocks, for n_parts workgroups.
std::vector<uint32_t> balance_rows(const std::vector<uint32_t> &row_ptr, uint32_t n_parts);

// Symmetric SELL-64 storage of K for the block SpMV: each row keeps its upper blocks (c >= r,
// diagonal first) with values; its lower blocks (c < r) are read transposed from row c's upper
// storage through a position index, the way the reference's MatWithTinyBlocks::partial_mul uses
// each stored (lower) block for both rows (containers/MatWithTinyBlocks.h:303-378).  Halves the
// value bytes streamed per product; the transposed reads of recently processed rows hit L2 or
// the Infinity Cache.
struct SymSell {
    std::vector<uint32_t> sptrU, lenU, colU, srcU;   // upper part; srcU = position of the block in K
    std::vector<uint32_t> sptrL, lenL, colL, tposL;  // lower part: column, position in the upper part
    uint64_t nU = 0, nL = 0;
    uint64_t posU(uint32_t r, uint32_t k) const { return sptrU[r >> 6] + 64ull * k + (r & 63u); }
    uint64_t posL(uint32_t r, uint32_t k) const { return sptrL[r >> 6] + 64ull * k + (r & 63u); }
};
// Returns false when the pattern is not symmetric or a position overflows 32 bits.
bool build_sym_sell(const Pattern &pat, SymSell &S);

// Block IC(0) schedule (containers/MatWithTinyBlocks.h:426-523 factorises in natural row order).
// An elimination order is chosen (IC0_NATURAL: node order, the reference's factor;
// IC0_COLOURED: first-fit colouring of the node graph, nodes sorted by colour), each node gets
// the level 1 + max level of its neighbours eliminated before it, and rows are renumbered
// t = 0..n-1 by (level, elimination position).  Rows of one level share no block, so a level is
// factorised and swept in parallel; reordering rows within a level leaves the factor unchanged.
// Factor storage is SELL-64 over t: row t holds its neighbours sorted by t -- nlow[t] strictly
// lower blocks, the diagonal block, then the upper blocks (the upper slot (t, c) receives L_ct,
// read transposed by the backward sweep).
enum { IC0_NATURAL = 0, IC0_COLOURED = 1 };
struct Ic0Schedule {
    int ordering = -1;
    uint32_t n_rows = 0;
    std::vector<uint32_t> order;     // t -> node
    std::vector<uint32_t> lv_ptr;    // level l covers rows t in [lv_ptr[l], lv_ptr[l+1])
    std::vector<uint32_t> sptr;      // slice starts (positions), n_slices + 1
    std::vector<uint32_t> len, nlow; // per row: stored blocks, strictly lower blocks
    uint64_t n_sell = 0;
    std::vector<uint32_t> colN;      // per position: neighbour node
    std::vector<uint32_t> colT;      // per position: neighbour row t
    std::vector<uint32_t> kpos;      // per position: SELL position of the block in K
    std::vector<uint32_t> tpos;      // lower positions: position of the transposed (upper) slot
    uint64_t pos(uint32_t t, uint32_t k) const { return sptr[t >> 6] + 64ull * k + (t & 63u); }
    uint32_t n_levels() const { return lv_ptr.empty() ? 0 : (uint32_t)lv_ptr.size() - 1; }
};
// base: the nodes in the order the elimination starts from (natural order, and the visiting
// order of the greedy colouring); null for 0..n-1.  Returns false when the pattern is not
// symmetric.
bool build_ic0_schedule(const Pattern &pat, int ordering, Ic0Schedule &S, const std::vector<uint32_t> *base = nullptr);

}  // namespace femgpu
