"""Matrix-free Hexa and Tetra_10 products (femgpu_set_spmv_config FEMGPU_SPMV_MATRIX_FREE) through
the C ABI.

The product K x is evaluated element by element from the mesh instead of read from the stored
matrix.  Pinned here against the oracle's independently assembled K (oracle.assemble + penalty,
whose element kernels are pinned in test_oracle_elem.py):
* FP64: |K_mf x - K x| <= 1e-12 (|K| |x|) per DOF; the FP32 product to 2e-5 of the same scale
  (the stored FP32 copy is held to 1e-6 beside it);
* the penalty diagonal, a node of no element (its row is the penalty alone) and the internal node
  order are covered, on general (jittered, rotated) elements, on axis-aligned boxes (the lattice
  and a rectilinear grid: the box-only kernel) and on a mesh mixing both in every block; Tetra_10
  elements with straight and with curved edges (mid-edge nodes off the chords);
* with assume_constant_matrix the product keeps the material of the last matrix assembly;
* solutions with the product in the loop match the oracle's solve to 1e-10 (FP64) and 1e-8
  (MIXED), with block Jacobi and block IC(0), and repeat bitwise;
* problems it cannot serve (Tetra, batches with a node of no element) are refused.
"""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import clamp_and_traction, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu

MF = (A.SPMV_MATRIX_FREE, 2, A.LOADS_NT)
FULL = (A.SPMV_FULL, 2, A.LOADS_NT)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _mesh(div, seed, isolated=False, shuffle=False, geometry="jittered"):
    """geometry: "jittered" (interior nodes moved: general elements), "lattice" (make_rect's
    axis-aligned boxes), "rectilinear" (boxes of unequal sizes), "rotated" (the lattice turned
    about the x axis: boxes that are not axis-aligned), "mixed" (nodes with x < 0.5
    jittered: box and general elements side by side in every element block)."""
    c, e = L.make_rect(L.HEXA, div)
    if geometry == "jittered":
        c = jitter_interior(c, 1.0 / max(div), 0.2, seed)
    elif geometry == "rectilinear":
        c = c ** 1.3
    elif geometry == "rotated":
        a = 0.3  # about the x axis, so the clamped and loaded planes x = 0, 1 stay in place
        R = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(a), -np.sin(a)], [0.0, np.sin(a), np.cos(a)]])
        c = c @ R.T
    elif geometry == "mixed":
        j = jitter_interior(c, 1.0 / max(div), 0.2, seed)
        c = np.where(c[:, :1] < 0.5, j, c)
    if shuffle:
        perm = np.random.default_rng(seed).permutation(len(c))
        new_of = np.empty(len(c), np.int64)
        new_of[perm] = np.arange(len(c))
        c, e = c[perm].copy(), new_of[e].astype(np.uint32)
    if isolated:  # a node no element uses, clamped below like the x = 0 face
        c = np.vstack([c, [[0.5, 0.5, 2.0]]])
    return c, e, lognormal_E(len(e), 0.4, seed)


def _tet10(cells, seed, curved=False, shuffle=False):
    """Tetra_10 on cells^3 cubes (corner nodes jittered); curved: interior mid-edge nodes moved
    off the chords by up to 5% of a cube edge, so J varies over each element."""
    c, e = tetra10_jittered(cells, seed=seed)
    if curved:
        mid = np.unique(e[:, 4:])
        mid = mid[np.all((c[mid] > 1e-9) & (c[mid] < 1 - 1e-9), axis=1)]
        c = c.copy()
        c[mid] += np.random.default_rng(seed + 1).uniform(-0.05, 0.05, (len(mid), 3)) / cells
    if shuffle:
        perm = np.random.default_rng(seed).permutation(len(c))
        new_of = np.empty(len(c), np.int64)
        new_of[perm] = np.arange(len(c))
        c, e = c[perm].copy(), new_of[e].astype(np.uint32)
    return c, e, lognormal_E(len(e), 0.4, seed)


def _system(oracle, c, e, E, isolated, order=A.NODE_ORDER_NATURAL, nu=0.3, kind=L.HEXA):
    faces, tr, dofs = clamp_and_traction(kind, c, e, (0.0, 0.4, -1.0))
    assert len(dofs) > 0
    if isolated:
        dofs = np.concatenate([dofs, 3 * (len(c) - 1) + np.arange(3)]).astype(np.uint32)
    with A.node_order_ctx(order):
        f = L.Formulation(kind, c, e)
    f.set_material(E, nu, (0.1, -0.2, 0.3))
    f.set_traction(faces, tr)
    f.add_constraints(dofs, None, 1e5)
    f.assemble()
    rp, col, vals, F, rc = oracle.assemble(kind, c, e, E, nu, bf=(0.1, -0.2, 0.3), faces=faces, face_trac=tr)
    assert rc == 0
    oracle.penalty(rp, col, vals, F, dofs, None, 1e5)
    return f, (rp, col, vals, F)


def _rel_err(y, ref, scale):
    return float(np.max(np.abs(y - ref) / scale))


@pytest.mark.parametrize("case", ["jittered", "isolated_node", "node_order", "lattice", "rectilinear", "rotated",
                                  "mixed", "lattice_isolated_node"])
def test_product_matches_oracle_matrix(oracle, case):
    say("matrix-free product", case)
    iso = case.endswith("isolated_node")
    geometry = case.replace("_isolated_node", "") if case in ("lattice", "rectilinear", "rotated", "mixed",
                                                             "lattice_isolated_node") else "jittered"
    c, e, E = _mesh((7, 6, 5), 11, isolated=iso, shuffle=case == "node_order", geometry=geometry)
    order = A.NODE_ORDER_RCM if case == "node_order" else A.NODE_ORDER_NATURAL
    f, (rp, col, vals, F) = _system(oracle, c, e, E, iso, order)
    boxes = f.matrix_free_info()["box_elements"]
    want = {"lattice": len(e), "rectilinear": len(e), "lattice_isolated_node": len(e)}.get(case, 0)
    if case == "mixed":
        assert 0 < boxes < len(e), boxes
    else:
        assert boxes == want, (boxes, want)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(3 * len(c))
    ref = oracle.spmv(rp, col, vals, x)
    scale = oracle.spmv(rp, col, np.abs(vals), np.abs(x))
    f.set_spmv_config(fp64=FULL, fp32=FULL)
    e_full64 = _rel_err(f.apply(x, A.FP64), ref, scale)
    e_full32 = _rel_err(f.apply(x, A.MIXED), ref, scale)
    f.set_spmv_config(fp64=MF, fp32=MF)
    assert f.spmv_config()["fp64"][0] == A.SPMV_MATRIX_FREE
    e_mf64 = _rel_err(f.apply(x, A.FP64), ref, scale)
    e_mf32 = _rel_err(f.apply(x, A.MIXED), ref, scale)
    say("  errors full64 %.2e full32 %.2e mf64 %.2e mf32 %.2e" % (e_full64, e_full32, e_mf64, e_mf32))
    assert e_full64 <= 1e-13 and e_full32 <= 1e-6
    assert e_mf64 <= 1e-12, e_mf64
    assert e_mf32 <= 2e-5, e_mf32
    # repeated products are bitwise equal (fixed colour order, no atomics)
    assert np.array_equal(f.apply(x, A.FP64), f.apply(x, A.FP64))
    assert np.array_equal(f.apply(x, A.MIXED), f.apply(x, A.MIXED))
    f.close()


@pytest.mark.parametrize("case", ["straight", "curved", "node_order"])
def test_tetra10_product_matches_oracle_matrix(oracle, case):
    say("matrix-free Tetra_10 product", case)
    c, e, E = _tet10(5, 21, curved=case == "curved", shuffle=case == "node_order")
    order = A.NODE_ORDER_RCM if case == "node_order" else A.NODE_ORDER_NATURAL
    f, (rp, col, vals, F) = _system(oracle, c, e, E, False, order, kind=L.TETRA10)
    x = np.random.default_rng(7).standard_normal(3 * len(c))
    ref = oracle.spmv(rp, col, vals, x)
    scale = oracle.spmv(rp, col, np.abs(vals), np.abs(x))
    f.set_spmv_config(fp64=MF, fp32=MF)
    info = f.matrix_free_info()
    assert info["box_elements"] == 0
    e_mf64 = _rel_err(f.apply(x, A.FP64), ref, scale)
    e_mf32 = _rel_err(f.apply(x, A.MIXED), ref, scale)
    say("  errors mf64 %.2e mf32 %.2e" % (e_mf64, e_mf32), info)
    assert e_mf64 <= 1e-12, e_mf64
    assert e_mf32 <= 2e-5, e_mf32
    assert np.array_equal(f.apply(x, A.FP64), f.apply(x, A.FP64))
    assert np.array_equal(f.apply(x, A.MIXED), f.apply(x, A.MIXED))
    f.close()


@pytest.mark.parametrize("curved", [False, True])
def test_tetra10_solutions_match_oracle(oracle, curved):
    say("matrix-free Tetra_10 solve", curved)
    c, e, E = _tet10(5, 23, curved=curved)
    f, (rp, col, vals, F) = _system(oracle, c, e, E, False, kind=L.TETRA10)
    x, _, _, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14, block=True)
    assert rc == 0
    f.precond = L.BLOCK_JACOBI
    iters = {}
    for cfg in (FULL, MF):
        f.set_spmv_config(fp64=cfg, fp32=cfg)
        for prec, tol in ((L.FP64, 1e-10), (L.MIXED, 1e-8)):
            f.precision = prec
            f.set_result(np.zeros_like(F))
            assert f.solve_system(rel_tol=1e-12), (cfg, prec)
            u = f.get_result()
            assert np.linalg.norm(u - x) <= tol * np.linalg.norm(x), (cfg, prec)
            iters[(cfg[0], prec)] = f.last_stats.iters
            if cfg == MF:
                f.set_result(np.zeros_like(F))
                assert f.solve_system(rel_tol=1e-12)
                assert np.array_equal(f.get_result(), u)
    say("  iterations", iters)
    a, b = iters[(A.SPMV_FULL, L.FP64)], iters[(A.SPMV_MATRIX_FREE, L.FP64)]
    assert abs(a - b) <= 2, (a, b)
    f.close()


def test_product_keeps_the_assembled_material(oracle):
    say("matrix-free snapshot")
    c, e, E = _mesh((5, 5, 4), 12)
    f, (rp, col, vals, F) = _system(oracle, c, e, E, False)
    f.set_spmv_config(fp64=MF, fp32=MF)
    x = np.random.default_rng(6).standard_normal(3 * len(c))
    y0 = f.apply(x, A.FP64)
    f.set_material(3.0 * E, 0.25)            # not assembled: K, and the product, keep E and nu
    assert np.array_equal(f.apply(x, A.FP64), y0)
    f.assemble()
    _, (rp2, col2, vals2, _) = _system(oracle, c, e, 3.0 * E, False, nu=0.25)
    ref = oracle.spmv(rp2, col2, vals2, x)
    scale = oracle.spmv(rp2, col2, np.abs(vals2), np.abs(x))
    assert _rel_err(f.apply(x, A.FP64), ref, scale) <= 1e-12
    f.close()


@pytest.mark.parametrize("precond,geometry", [(L.BLOCK_JACOBI, "jittered"), (L.BLOCK_IC0, "jittered"),
                                              (L.BLOCK_JACOBI, "lattice"), (L.BLOCK_JACOBI, "mixed")])
def test_solutions_match_oracle(oracle, precond, geometry):
    say("matrix-free solve", precond, geometry)
    c, e, E = _mesh((8, 6, 6), 13, geometry=geometry)
    f, (rp, col, vals, F) = _system(oracle, c, e, E, False)
    x, _, _, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14, block=True)
    assert rc == 0
    f.precond = precond
    iters = {}
    for cfg in (FULL, MF):
        f.set_spmv_config(fp64=cfg, fp32=cfg)
        for prec, tol in ((L.FP64, 1e-10), (L.MIXED, 1e-8)):
            f.precision = prec
            f.set_result(np.zeros_like(F))
            assert f.solve_system(rel_tol=1e-12), (cfg, prec)
            u = f.get_result()
            assert np.linalg.norm(u - x) <= tol * np.linalg.norm(x), (cfg, prec)
            iters[(cfg[0], prec)] = f.last_stats.iters
            if cfg == MF:
                f.set_result(np.zeros_like(F))
                assert f.solve_system(rel_tol=1e-12)
                assert np.array_equal(f.get_result(), u)          # deterministic
    say("  iterations", iters)
    # FP64: the same operator to ~1e-15, so the same iteration count; MIXED: the FP32 element
    # product is more accurate than the rounded FP32 matrix, so it may need fewer, never many more
    a, b = iters[(A.SPMV_FULL, L.FP64)], iters[(A.SPMV_MATRIX_FREE, L.FP64)]
    assert abs(a - b) <= 2, (a, b)
    a, b = iters[(A.SPMV_FULL, L.MIXED)], iters[(A.SPMV_MATRIX_FREE, L.MIXED)]
    assert b <= a + max(5, 0.1 * a), (a, b)
    f.close()


def test_time_iteration_keeps_the_result(oracle):
    """femgpu_time_iteration times the two kernels of the CG iteration in a solve of its own (into
    the scratch vector): positive times, the problem's result and its next solve unchanged, and a
    refusal where the iteration is not the two-kernel one (the stored product)."""
    say("time_iteration")
    c, e, E = _mesh((6, 5, 5), 3)
    f, (rp, col, vals, F) = _system(oracle, c, e, E, False)
    f.set_spmv_config(fp64=MF, fp32=MF)
    f.precond, f.precision = L.BLOCK_JACOBI, L.MIXED
    f.set_result(np.zeros(3 * len(c)))
    assert f.solve_system(rel_tol=1e-10)
    u0, it0 = f.get_result(), f.last_stats.iters
    for prec in (L.MIXED, L.FP64):
        tb, tu = f.time_iteration(prec, 20)
        assert tb > 0 and tu > 0, (prec, tb, tu)
    assert np.array_equal(f.get_result(), u0)
    f.set_result(np.zeros(3 * len(c)))
    assert f.solve_system(rel_tol=1e-10)
    assert np.array_equal(f.get_result(), u0) and f.last_stats.iters == it0
    f.set_spmv_config(fp64=FULL, fp32=FULL)
    with pytest.raises(L.FemGpuError):
        f.time_iteration(L.MIXED, 20)
    f.close()


def test_refused_where_it_cannot_serve():
    say("matrix-free refused")
    c, e = L.make_rect(L.TETRA, (4, 4, 4))
    f = L.Formulation(L.TETRA, c, e)
    with pytest.raises(L.FemGpuError):
        f.set_spmv_config(fp64=MF, fp32=FULL)
    f.close()
    # a batch whose meshes have a node of no element (its pass is served for a single problem only)
    c, e = L.make_rect(L.HEXA, (3, 3, 3))
    meshes = [(np.vstack([c, [[0.5, 0.5, 2.0]]]), e), L.make_rect(L.HEXA, (4, 3, 3))]
    b = L.Batch(L.HEXA, meshes)
    with pytest.raises(L.FemGpuError):
        b.set_spmv_config(fp64=MF, fp32=MF)
    b.close()


@pytest.mark.parametrize("precision,assembly,kind", [(L.FP64, "staged", L.HEXA), (L.MIXED, "gather", L.HEXA),
                                                    (L.MIXED, "staged", L.TETRA10)])
def test_batch_matches_individual_solves(oracle, precision, assembly, kind):
    """Matrix-free products in a batch: each sub-problem's blocks are its own (a block never
    straddles two), so every sub-problem gives bitwise the solution and iteration count of the
    same problem solved alone with the matrix-free product; lattices (box elements) and jittered
    meshes of several sizes, moduli and Poisson ratios side by side, and the oracle's solution."""
    from problems import clamp_and_traction
    say("matrix-free batch", precision, assembly, kind)
    rng = np.random.default_rng(17)
    items = []
    for i in range(5):
        div = tuple(int(v) for v in rng.integers(3, 8, size=3))
        if kind == L.TETRA10:
            c, e, E = _tet10(2 + i % 3, 30 + i, curved=i % 2 == 1)
        else:
            c, e, E = _mesh(div, 30 + i, geometry="lattice" if i % 2 == 0 else "jittered")
        items.append((c, e, E, float(rng.uniform(0.2, 0.4)), rng.standard_normal(3)))
    b = L.Batch(kind, [(c, e) for c, e, *_ in items])
    singles = []
    for i, (c, e, E, nu, tr) in enumerate(items):
        faces, trs, dofs = clamp_and_traction(kind, c, e, tr)
        b.set_material(i, E, nu)
        b.set_traction(i, faces, trs)
        b.add_constraints(i, dofs, None, 1e5)
        f = L.Formulation(kind, c, e)
        f.set_material(E, nu)
        f.set_traction(faces, trs)
        f.add_constraints(dofs, None, 1e5)
        f.assemble()
        f.set_spmv_config(fp64=MF, fp32=MF)
        f.precision = precision
        f.precond = L.JACOBI
        f.set_result(np.zeros(3 * len(c)))
        assert f.solve_system(rel_tol=1e-11)
        singles.append((f.get_result(), f.last_stats.iters))
        rp, col, vals, F, rc = oracle.assemble(kind, c, e, E, nu, faces=faces, face_trac=trs)
        assert rc == 0
        oracle.penalty(rp, col, vals, F, dofs, None, 1e5)
        x, _, _, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14)
        assert rc == 0
        tol = 1e-9 if precision == L.FP64 else 1e-8
        assert np.linalg.norm(singles[-1][0] - x) <= tol * np.linalg.norm(x), i
        f.close()
    if assembly == "gather":   # the batch's K and F are bitwise the staged passes'
        b.set_assembly_rows(A.ASSEMBLY_GATHER)
    b.assemble()
    b.set_spmv_config(fp64=MF, fp32=MF)
    ok, ms = b.solve(rel_tol=1e-11, precision=precision, precond=L.JACOBI)
    assert ok and ms > 0
    for i, (u, it) in enumerate(singles):
        say("  sub-problem", i, "iterations", b.last_stats[i].iters, it)
        assert b.last_stats[i].iters == it
        assert np.array_equal(b.get_result(i), u)
    b.close()
