"""GPU parity of the HIP path (through the C ABI) against the oracle and the reference fixtures.

Tolerances: assembled blocks agree to 1e-13 of the matrix scale (the reference's generated
kernels and summation order are not reproducible bit for bit); patterns and DOF numbering are
exact; FP64 solutions agree with the oracle's dense solve to 1e-10 relative; MIXED solutions to
1e-8 relative.
"""
import json
import os
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from conftest import GOLDEN
from lmtpp_amd import _capi as A
from problems import E0, NU0, build, fd_gradient, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def oracle_system(O, kind, c, e, E, nu, bf, faces, tr, dofs, P=1e5):
    rp, col, vals, F, rc = O.assemble(kind, c, e, E, nu, bf=bf, faces=faces, face_trac=tr)
    M = O.penalty(rp, col, vals, F, dofs, None, P)
    return rp, col, vals, F, rc, M


def dense(rp, col, vals):
    n = 3 * (len(rp) - 1)
    A = np.zeros((n, n))
    for r in range(len(rp) - 1):
        for p in range(rp[r], rp[r + 1]):
            A[3 * r:3 * r + 3, 3 * col[p]:3 * col[p] + 3] = vals[p]
    return A


def small_case(kind, seed):
    if kind == L.TETRA10:
        c, e = tetra10_jittered(3, seed=seed)
    else:
        c, e = L.make_rect(kind, (5, 4, 4))
        c = jitter_interior(c, 0.25, 0.15, seed)
    E = lognormal_E(len(e), 0.5, seed)
    return c, e, E


@pytest.mark.parametrize("kind", [L.TETRA, L.TETRA10, L.HEXA])
def test_assembly_matches_oracle(oracle, kind):
    say("assembly", kind)
    c, e, E = small_case(kind, 11 + kind)
    bf = (0.3, -0.2, -9.81 * 7.8e-9 * 1e6)
    f, faces, tr, dofs = build(kind, c, e, E, bf=bf)
    f.assemble()
    rp, col, vals = f.get_mat()
    F = f.get_sollicitation()
    orp, ocol, ovals, oF, rc, M = oracle_system(oracle, kind, c, e, E, NU0, bf, faces, tr, dofs)
    assert rc == 0
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)          # pattern: exact
    scale = np.abs(ovals).max()
    assert np.abs(vals - ovals).max() <= 1e-13 * scale
    blk = np.abs(ovals).max(axis=(1, 2))
    big = blk > 1e-3 * scale
    assert (np.abs(vals - ovals).max(axis=(1, 2))[big] <= 1e-14 * 8 * blk[big]).all()
    assert np.abs(F - oF).max() <= 1e-13 * np.abs(oF).max()
    assert f.max_diag() == pytest.approx(M, rel=1e-14)
    # exact symmetry of the stored full matrix
    A = dense(rp, col, vals)
    assert np.array_equal(A, A.T)


def test_config0_solution_matches_reference_and_oracle(oracle):
    say("config0 solve")
    g = np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))
    c, e = L.make_rect(L.TETRA, (11, 11, 11))
    f, faces, tr, dofs = build(L.TETRA, c, e, E0)
    f.assemble()
    rp, col, vals = f.get_mat()
    F = f.get_sollicitation()
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(col, g["col"])
    assert np.abs(vals - g["vals"]).max() <= 1e-13 * np.abs(g["vals"]).max()
    assert np.abs(F - g["F"]).max() <= 1e-15
    xs = np.linalg.solve(dense(g["row_ptr"], g["col"], g["vals"]), g["F"])
    for prec in (L.JACOBI, L.BLOCK_JACOBI):
        f.precond = prec
        f.precision = L.FP64
        f.set_result(np.zeros_like(F))
        assert f.solve_system(rel_tol=1e-12), f.last_stats.status
        u = f.get_result()
        assert f.last_stats.rel_res <= 1e-12, f.last_stats.rel_res
        assert np.linalg.norm(u - xs) <= 1e-10 * np.linalg.norm(xs)
        # the reference's own block IC(0)-PCG result, to its stop test (max|r| <= 1e-10 max|F|)
        assert np.linalg.norm(u - g["x"]) <= 1e-9 * np.linalg.norm(g["x"])
    f.precision = L.MIXED
    f.precond = L.JACOBI
    f.set_result(np.zeros_like(F))
    assert f.solve_system(rel_tol=1e-12)
    u = f.get_result()
    assert f.last_stats.refreshes >= 1
    assert np.linalg.norm(u - xs) <= 1e-8 * np.linalg.norm(xs)


def test_abs_maxnorm_stop_matches_reference_semantics(oracle):
    say("abs maxnorm")
    c, e = L.make_rect(L.TETRA, (6, 5, 5))
    f, *_ = build(L.TETRA, c, e, E0)
    f.assemble()
    F = f.get_sollicitation()
    crit = 1e-9 * np.abs(F).max()
    assert f.solve_system(crit=crit)
    assert f.last_stats.max_abs_res <= crit
    rp, col, vals = f.get_mat()
    r = F - oracle.spmv(rp, col, vals, f.get_result())
    assert np.abs(r).max() <= 1.001 * crit


@pytest.mark.parametrize("kind", [L.TETRA10, L.HEXA])
def test_solve_matches_oracle_pcg(oracle, kind):
    say("solve parity", kind)
    c, e, E = small_case(kind, 3)
    f, faces, tr, dofs = build(kind, c, e, E, bf=(0.0, 0.0, -9.81 * 7.8e-9))
    assert f.solve(rel_tol=1e-12)
    u = f.get_result()
    orp, ocol, ovals, oF, _, _ = oracle_system(oracle, kind, c, e, E, NU0, (0.0, 0.0, -9.81 * 7.8e-9), faces, tr, dofs)
    xs = np.linalg.solve(dense(orp, ocol, ovals), oF)
    assert np.linalg.norm(u - xs) <= 1e-10 * np.linalg.norm(xs)
    ux = f.update_variables()
    assert ux.shape == (len(c), 3)


def test_block_system_known_answers():
    say("bsr kat")
    with open(os.path.join(GOLDEN, "ref_cg_kat.json")) as fh:
        k = json.load(fh)
    A = np.zeros((3, 3))
    for i, row in enumerate(k["sym3"]["lower"]):
        for j, v in enumerate(row):
            A[i, j] = A[j, i] = v
    s = L.BlockSystem([0, 1], [0], A.reshape(1, 3, 3), np.array(k["sym3"]["b"], float))
    s.precond = L.BLOCK_JACOBI
    assert s.solve_system(rel_tol=1e-15)
    assert np.allclose(s.get_result(), k["sym3"]["x"], rtol=1e-12, atol=0)
    # 10x10 dense system padded with an identity block row to 12 = 4 block rows
    A10 = np.array(k["dense10"]["rows"], float)
    A12 = np.eye(12)
    A12[:10, :10] = A10
    b12 = np.zeros(12)
    b12[:10] = k["dense10"]["b"]
    rp = np.arange(0, 17, 4)
    col = np.tile(np.arange(4), 4)
    vals = np.array([A12[3 * r:3 * r + 3, 3 * c:3 * c + 3] for r in range(4) for c in range(4)])
    s = L.BlockSystem(rp, col, vals, b12)
    s.precond = L.BLOCK_JACOBI
    # cond(A10) = 1.9e4: FP64 attains about ||A|| ||x|| eps / ||b|| = 4e-13 of relative residual
    assert s.solve_system(rel_tol=2e-12), s.last_stats.status
    x = s.get_result()[:10]
    xr = np.array(k["dense10"]["x"])
    assert (np.abs(x - xr) / (1e-20 + np.abs(x) + np.abs(xr))).max() < 1e-10


def test_adjoint_gradient_matches_oracle_and_fd(oracle):
    say("adjoint")
    kind = L.HEXA
    c, e = L.make_rect(kind, (6, 6, 6))
    E = lognormal_E(len(e), 0.3, 5)
    f, faces, tr, dofs = build(kind, c, e, E)
    assert f.solve(rel_tol=1e-11)
    u = f.get_result()
    rng = np.random.default_rng(1)
    top = np.where(np.abs(c[:, 2] - 1.0) < 1e-9)[0]
    obs = rng.choice(top, size=max(3, len(top) // 10), replace=False)
    odofs = (3 * obs[:, None] + np.arange(3)).ravel()
    u_meas = u[odofs] * (1 + 0.05 * rng.standard_normal(len(odofs)))
    dJdu = np.zeros_like(u)
    dJdu[odofs] = u[odofs] - u_meas
    g = f.adjoint(dJdu, rel_tol=1e-11)
    # oracle adjoint: dense solve with the oracle's K (penalty scale held fixed)
    orp, ocol, ovals, oF, _, _ = oracle_system(oracle, kind, c, e, E, NU0, (0, 0, 0), faces, tr, dofs)
    Ad = dense(orp, ocol, ovals)
    lam = np.linalg.solve(Ad, dJdu)
    ug = np.linalg.solve(Ad, oF)
    go = oracle.elem_grad_E(kind, e, c, E, NU0, ug, lam)
    gmax = np.abs(go).max()
    big = np.abs(go) >= 1e-3 * gmax
    assert (np.abs(g - go)[big] / np.abs(go)[big]).max() <= 1e-8
    assert np.abs(g - go).max() <= 1e-9 * gmax
    # element-gradient kernel alone, same u and lambda
    assert np.abs(f.elem_grad(ug, lam) - go).max() <= 1e-12 * gmax
    # central finite differences through the GPU solve (step 1e-4 E_e), per-element relative error
    res = u[odofs] - u_meas
    n_rel = 0
    for el in rng.choice(len(e), size=8, replace=False):
        fd = fd_gradient(f, oracle, E, u, res, odofs, el)
        if fd is None:
            continue
        if abs(g[el]) >= 1e-3 * gmax:
            assert abs(fd - g[el]) <= 1e-5 * abs(g[el]), (el, fd, g[el])
            n_rel += 1
        else:
            assert abs(fd - g[el]) <= 1e-8 * gmax, (el, fd, g[el])
    assert n_rel >= 4


def test_error_paths():
    say("errors")
    c, e = L.make_rect(L.TETRA, (3, 3, 3))
    f = L.Formulation(L.TETRA, c, e)
    with pytest.raises(L.FemGpuError) as ei:
        f.assemble()
    assert ei.value.status == L._capi.NOT_READY
    with pytest.raises(L.FemGpuError) as ei:
        f.set_material(np.full(len(e), np.nan), 0.3)
    assert ei.value.status == L._capi.NAN
    bad = e.copy()
    bad[0, [1, 2]] = bad[0, [2, 1]]   # inverted element
    g = L.Formulation(L.TETRA, c, bad)
    g.set_material(1.0, 0.3)
    with pytest.raises(L.FemGpuError) as ei:
        g.assemble()
    assert ei.value.status == L._capi.NEG_JACOBIAN
    # no Dirichlet condition: singular K, iteration cap reported, nothing hangs
    h = L.Formulation(L.TETRA, c, e)
    h.set_material(1.0, 0.3, (0.0, 0.0, 1.0))
    h.max_iter = 50
    h.assemble()
    assert not h.solve_system(rel_tol=1e-30)
    assert h.last_status in (L._capi.MAX_ITER, L._capi.NOT_SPD)
    with pytest.raises(L.FemGpuError):
        h.add_constraint(10 ** 6, 0.0, 1e5)


def test_bitwise_reproducible():
    say("determinism")
    c, e, E = small_case(L.HEXA, 9)
    outs = []
    for _ in range(2):
        f, *_ = build(L.HEXA, c, e, E)
        assert f.solve(rel_tol=1e-12)
        outs.append((f.get_mat()[2], f.get_result()))
        f.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_cpp_facade_example(tmp_path):
    """The C++ FormulationGpu facade (include/femgpu_formulation.hpp) driven by the example
    program, with the reference's stop test max|r| <= 1e-10 max|F| on config 0."""
    import subprocess
    say("cpp facade")
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "lmtpp_amd", "bin", "elastic_cube")
    out = str(tmp_path / "u.bin")
    r = subprocess.run([exe, "tetra", "11", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    u = np.fromfile(out)
    g = np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))
    assert np.linalg.norm(u - g["x"]) <= 1e-8 * np.linalg.norm(g["x"])
    # with the reference's own preconditioner: its iteration count and its solution
    r = subprocess.run([exe, "tetra", "11", out, "ic0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert " %d iterations" % int(g["iters"]) in r.stdout, r.stdout
    u = np.fromfile(out)
    assert np.linalg.norm(u - g["x"]) <= 1e-10 * np.linalg.norm(g["x"])
    # ProblemGpu time steps: traction x k, constant K and factor, warm starts
    r = subprocess.run([exe, "tetra", "11", out, "ramp"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps = [ln for ln in r.stdout.splitlines() if ln.startswith("step ")]
    assert len(steps) == 3, r.stdout
    tip = [float(ln.split("tip uz ")[1]) for ln in steps]
    times = [float(ln.split("time ")[1].split(":")[0]) for ln in steps]
    assert times == [-1.0, 0.0, 1.0]
    for k in (2, 3):
        assert abs(tip[k - 1] - k * tip[0]) <= 1e-9 * abs(k * tip[0])


def _batch_inputs(seed=4, n=3):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        cells = int(rng.integers(3, 7))
        c, e = L.make_rect(L.HEXA, (cells + 1,) * 3)
        E = 210000.0 * np.exp(0.5 * rng.standard_normal(len(e)))
        nu = float(rng.uniform(0.2, 0.4))
        tr = rng.standard_normal(3)
        out.append((c, e, E, nu, tr))
    return out


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_batch_matches_individual_solves(precision):
    """Sub-problems of a batch give bitwise the same K, F and solution as separate problems
    (same element colours, same reduction order)."""
    from problems import clamp_and_traction
    say("batch", precision)
    items = _batch_inputs()
    b = L.Batch(L.HEXA, [(c, e) for c, e, *_ in items])
    singles = []
    for i, (c, e, E, nu, tr) in enumerate(items):
        faces, trs, dofs = clamp_and_traction(L.HEXA, c, e, tr)
        b.set_material(i, E, nu)
        b.set_traction(i, faces, trs)
        b.add_constraints(i, dofs, None, 1e5)
        f = L.Formulation(L.HEXA, c, e)
        f.set_material(E, nu)
        f.set_traction(faces, trs)
        f.add_constraints(dofs, None, 1e5)
        f.precision = precision
        assert f.solve(rel_tol=1e-11)
        singles.append((f.get_result(), f.last_stats.iters))
    b.assemble()
    ok, ms = b.solve(rel_tol=1e-11, precision=precision)
    assert ok and ms > 0
    for i in range(len(items)):
        u = b.get_result(i)
        assert b.last_stats[i].iters == singles[i][1]
        assert np.array_equal(u, singles[i][0])


def test_batch_autotune_keeps_results_and_picks_a_variant():
    """femgpu_batch_autotune: the batch's results are left as they were, the chosen product is
    one of the timed variants and is the configuration in force, a solve after it matches the
    single problems' FP64 solutions, and the chunk timing of that solve is consistent."""
    from problems import clamp_and_traction
    say("batch autotune")
    items = _batch_inputs(seed=12)
    b = L.Batch(L.HEXA, [(c, e) for c, e, *_ in items])
    refs = []
    for i, (c, e, E, nu, tr) in enumerate(items):
        faces, trs, dofs = clamp_and_traction(L.HEXA, c, e, tr)
        b.set_material(i, E, nu)
        b.set_traction(i, faces, trs)
        b.add_constraints(i, dofs, None, 1e5)
        f = L.Formulation(L.HEXA, c, e)
        f.set_material(E, nu)
        f.set_traction(faces, trs)
        f.add_constraints(dofs, None, 1e5)
        assert f.solve(rel_tol=1e-12)
        refs.append(f.get_result())
        f.close()
    b.assemble()
    ok, _ = b.solve(rel_tol=1e-6, precision=L.FP64)
    assert ok
    before = [b.get_result(i) for i in range(len(items))]
    prod = b.autotune(rel_tol=1e-10, precision=L.MIXED, precond=L.BLOCK_JACOBI)
    assert prod in (A.BATCH_STORED_PLAIN, A.BATCH_STORED_NT, A.BATCH_STORED_NT_COL16, A.BATCH_MATRIX_FREE)
    for i in range(len(items)):
        assert np.array_equal(b.get_result(i), before[i])
    b.clear_result()
    ok, ms = b.solve(rel_tol=1e-10, precision=L.MIXED, precond=L.BLOCK_JACOBI)
    assert ok
    for i in range(len(items)):
        assert np.linalg.norm(b.get_result(i) - refs[i]) <= 1e-8 * np.linalg.norm(refs[i])
    n, chunk_ms, gaps, gap_ms = b.chunk_stats()
    assert n >= 1 and 0 < chunk_ms <= ms * 1.01 and 0 <= gaps <= n - 1 and gap_ms >= 0
    b.close()


def test_graph_replay_is_bitwise_equal():
    say("graph replay")
    c, e, E = small_case(L.HEXA, 21)
    res = []
    for graph in (False, True, True):
        f, *_ = build(L.HEXA, c, e, E)
        f.set_graph_replay(graph)
        f.precision = L.MIXED
        assert f.solve(rel_tol=1e-11)
        res.append(f.get_result())
        f.close()
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[1], res[2])


def test_autotune_keeps_results():
    say("autotune")
    c, e, E = small_case(L.HEXA, 22)
    f, *_ = build(L.HEXA, c, e, E)
    assert f.solve(rel_tol=1e-11)
    u0 = f.get_result()
    cfg = f.autotune()
    # stored formats 0/1, or the matrix-free product (a Hexa problem with one sub-problem)
    assert all(fm in (0, 1, 2) and u in (1, 2, 3) and ld in (0, 1, 2) for fm, u, ld in cfg.values())
    f.set_result(np.zeros_like(u0))
    assert f.solve(rel_tol=1e-11)
    assert np.linalg.norm(f.get_result() - u0) <= 1e-9 * np.linalg.norm(u0)


@pytest.mark.parametrize("kind", [L.TETRA, L.TETRA10, L.HEXA])
def test_symmetric_spmv_format_matches_oracle(oracle, kind):
    """The symmetric SELL storage (upper blocks stored, lower ones read transposed) gives the
    oracle's solution in FP64 and MIXED, and the same residual as the full storage."""
    say("sym spmv", kind)
    c, e, E = small_case(kind, 30 + kind)
    f, faces, tr, dofs = build(kind, c, e, E)
    f.assemble()
    rp, col, vals = f.get_mat()
    F = f.get_sollicitation()
    x_ref, _, _, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-13)
    assert rc == 0
    its = {}
    for fmt, nt in ((L._capi.SPMV_FULL, 0), (L._capi.SPMV_FULL, 1), (L._capi.SPMV_FULL, 2), (L._capi.SPMV_SYM, 0)):
        for u in (1, 3):
            f.set_spmv_config(fp64=(fmt, u, nt), fp32=(fmt, u, nt))
            for prec, tol in ((L.FP64, 1e-10), (L.MIXED, 1e-8)):
                f.precision = prec
                f.set_result(np.zeros_like(F))
                assert f.solve_system(rel_tol=1e-11), (fmt, u, prec, f.last_stats.status)
                assert f.last_stats.rel_res <= 1e-11
                assert np.linalg.norm(f.get_result() - x_ref) <= tol * np.linalg.norm(x_ref)
                its[(fmt, nt, u, prec)] = f.last_stats.iters
    # same Krylov process up to rounding: iteration counts within a few percent
    for (fmt, nt, u, prec), n in its.items():
        assert abs(n - its[(0, 0, 1, prec)]) <= max(3, 0.05 * n)
    assert f.spmv_config() == {"fp64": (1, 3, 0), "fp32": (1, 3, 0)}


@pytest.mark.parametrize("kind", [L.TETRA, L.TETRA10, L.HEXA])
def test_assembly_variants_agree(oracle, kind):
    """Colour-pass, staged, gather and row-group element loops (every admissible group size) give
    the oracle's K and F and agree with one another to rounding (same per-entry summation order);
    the staged passes and the gather give bitwise the colour passes' K and F."""
    say("assembly variants", kind)
    c, e, E = small_case(kind, 40 + kind)
    bf = (0.1, 0.2, -0.3)
    f, faces, tr, dofs = build(kind, c, e, E, bf=bf)
    orp, ocol, ovals, oF, rc, M = oracle_system(oracle, kind, c, e, E, NU0, bf, faces, tr, dofs)
    _, rmax = f.assembly_rows()
    assert rmax >= 8
    out = {}
    for R in (0, L._capi.ASSEMBLY_STAGED, L._capi.ASSEMBLY_GATHER, rmax, rmax // 2, 8):
        f.set_assembly_rows(R)
        f.assemble()
        rp, col, vals = f.get_mat()
        F = f.get_sollicitation()
        scale = np.abs(ovals).max()
        assert np.abs(vals - ovals).max() <= 1e-13 * scale, R
        assert np.abs(F - oF).max() <= 1e-13 * np.abs(oF).max(), R
        out[R] = (vals, F)
    for V in (L._capi.ASSEMBLY_STAGED, L._capi.ASSEMBLY_GATHER):   # same adds, same order
        assert np.array_equal(out[V][0], out[0][0]) and np.array_equal(out[V][1], out[0][1]), V
    for R, (v, F) in out.items():
        assert np.abs(v - out[0][0]).max() <= 1e-15 * np.abs(out[0][0]).max(), R
        assert np.abs(F - out[0][1]).max() <= 1e-15 * np.abs(out[0][1]).max(), R
    assert f.assembly_rows()[0] == 8
    with pytest.raises(L.FemGpuError):
        f.set_assembly_rows(3)


@pytest.mark.parametrize("geometry", ["lattice", "rectilinear"])
def test_box_gather_assembly_matches_oracle(oracle, geometry):
    """FEMGPU_ASSEMBLY_GATHER_BOX (every element an axis-aligned box): K and F against the oracle's
    assembly to 1e-13 of their scales and against the element-block gather to 1e-15 (E_e times the
    unit-modulus matrix of the element's edge-length class, rounding aside), with a body force, a
    traction and per-element moduli; refused on a mesh with a non-box element."""
    say("box gather", geometry)
    c, e = L.make_rect(L.HEXA, (7, 6, 5))
    if geometry == "rectilinear":          # monotone maps of each axis keep the elements boxes
        c = np.stack([c[:, 0] ** 2, 0.5 * (c[:, 1] + c[:, 1] ** 3), np.sqrt(c[:, 2])], axis=1)
    E = lognormal_E(len(e), 0.5, 11)
    bf = (0.3, -0.2, 0.9)
    f, faces, tr, dofs = build(L.HEXA, c, e, E, bf=bf)
    rp, col, ovals, oF, rc = oracle.assemble(L.HEXA, c, e, E, NU0, bf=bf, faces=faces, face_trac=tr)
    assert rc == 0
    oracle.penalty(rp, col, ovals, oF, dofs, None, 1e5)
    got = {}
    for R in (L._capi.ASSEMBLY_GATHER, L._capi.ASSEMBLY_GATHER_BOX):
        f.set_assembly_rows(R)
        f.assemble()
        _, _, vals = f.get_mat()
        F = f.get_sollicitation()
        scale = f.max_diag()
        d = np.abs(vals - ovals) / np.maximum(scale, np.abs(ovals))
        assert d.max() <= 1e-13, (R, d.max())
        assert np.abs(F - oF).max() <= 1e-13 * np.abs(oF).max(), R
        got[R] = (vals, F)
    (v0, F0), (v1, F1) = got[L._capi.ASSEMBLY_GATHER], got[L._capi.ASSEMBLY_GATHER_BOX]
    assert (np.abs(v1 - v0) / np.maximum(f.max_diag(), np.abs(v0))).max() <= 1e-15
    assert np.abs(F1 - F0).max() <= 1e-15 * np.abs(F0).max()
    # exact symmetry: a position and its transpose sum the same class blocks
    n = len(c)
    for r in range(0, n, 7):
        for p in range(rp[r], rp[r + 1]):
            q = rp[col[p]] + np.searchsorted(col[rp[col[p]]:rp[col[p] + 1]], r)
            assert np.array_equal(v1[p], v1[q].T)
    f.close()
    cj = jitter_interior(L.make_rect(L.HEXA, (5, 4, 4))[0], 0.25, 0.1, 2)
    g, *_ = build(L.HEXA, cj, L.make_rect(L.HEXA, (5, 4, 4))[1], 210000.0)
    with pytest.raises(L.FemGpuError):
        g.set_assembly_rows(L._capi.ASSEMBLY_GATHER_BOX)
    g.close()
