"""Benchmark of the MI355X assemble + PCG path on Hexa8 100^3 elasticity (3,090,903 DOF).

One step = assembly of K and F on the GPU from the device-resident mesh and material, then the
mixed-precision (FP32 inner / FP64 refreshed) block-Jacobi PCG solve from a zero initial guess to
the relative residual that was calibrated, before timing, to give <= 0.5e-8 relative L2 error in
FP64 against the FP64 oracle solution.  That solution comes from the CPU leg: the oracle (the
plain-C restatement of the reference's path) assembles the same system on the host cores and
solves it by block-Jacobi PCG to a relative residual of 1e-12, and its K and F are compared with
the GPU's block by block.  ``value`` is DOF / step seconds (MDOF/s) of the whole step; the
FP64-only solve with the same preconditioner, the assembly rate, the product rooflines, the other
preconditioners (scalar Jacobi, block IC(0) in the reference's natural order and in multicolour
order), the after-solve recovery, configurations 0, 2, 3, 4 and the CPU baseline (the CPU leg's
time for the same workload and tolerance) are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 100]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak (packed FP32 issue)
FP64_VECTOR_PEAK_TFLOPS = 78.6    # MI355X published FP64 vector peak
METRIC = "solve s & MDOF/s to 1e-8 rel. err, FP64 vs mixed PCG; assembly Melem/s"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def spmv_bytes(nnzb, n_nodes, fp32, fmt=0, col16=0):
    from lmtpp_amd.traffic import spmv_bytes as sb
    return sb(nnzb, n_nodes, fp32, fmt, col16)


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return (d.get(name) or {}).get("traffic")


def host_cpu():
    """(logical CPUs of this process, CPU model name) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return n, model


def cpu_leg(kind, c, e, E, faces, tr, dofs, f, batch_items=()):
    """The CPU leg: the oracle (oracle/fem_oracle.c, the plain-C restatement of the reference's
    path) on the host cores, in the same run as the GPU.

    * All host cores (at most 16, the GPU box's CPU share): the full config-1 workload --
      coloured element loop, penalty, block-Jacobi PCG to a relative residual of 1e-12 with the
      time of every iteration recorded, so the time to any tolerance is read off afterwards.
      Its solution is the FP64 oracle solution the GPU results are compared with; its K and F are
      compared with the GPU's assembled K and F, block by block.
    * One core: the serial restatement (the reference's element-order scatter) on a Hexa8 48^3
      sample, assembly and 300 PCG iterations, scaled to the full workload.
    * The batch members batch_items = [(index, c, e, E, nu, faces, trac, dofs)] solved by the
      oracle, for the batch's error against the oracle.
    Returns a dict; ["u"] is the oracle solution."""
    from oracle import oracle as O
    import lmtpp_amd as L
    from problems import clamp_and_traction, lognormal_E
    O.build()
    n_cpu, model = host_cpu()
    threads = O.set_threads(min(16, n_cpu))
    out = {"threads": threads, "host_cpus": n_cpu, "cpu_model": model,
           "build": "gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"}
    pat = O.pattern(kind, len(c), e)
    t0 = time.perf_counter()
    rp, col, vals, F, rc = O.assemble(kind, c, e, E, 0.3, faces=faces, face_trac=tr, pat=pat)
    M = O.penalty(rp, col, vals, F, dofs, None, 1e5)
    out["assembly_s"] = time.perf_counter() - t0
    assert rc == 0
    x, it, rr, rc, hist = O.pcg_jacobi_hist(rp, col, vals, F, rel_tol=1e-12, block=True, max_iter=20000)
    out["pcg_iters"], out["pcg_s"], out["pcg_status"] = it, float(hist[-1, 1]) if len(hist) else 0.0, rc
    out["hist"] = hist
    out["true_rel_res"] = float(np.linalg.norm(F - O.spmv(rp, col, vals, x)) / np.linalg.norm(F))
    out["u"] = x
    log("CPU leg (%d threads, %s): oracle assembly %.2f s, block-Jacobi PCG %d iterations to 1e-12 in %.1f s, "
        "true rel. res %.2e" % (threads, model, out["assembly_s"], it, out["pcg_s"], out["true_rel_res"]))
    # the GPU's assembled K and F against the oracle's, every block
    grp, gcol, gvals = f.get_mat()
    same = np.array_equal(grp, rp) and np.array_equal(gcol, col)
    # each value's difference over max(max|diag K| before the constraints, |value|): the matrix
    # scale, or the value's own size on the penalty-carrying diagonal entries
    dk = 0.0
    step = 1 << 22
    for i in range(0, len(col), step):
        ov = vals[i:i + step]
        dk = max(dk, float((np.abs(gvals[i:i + step] - ov) / np.maximum(M, np.abs(ov))).max()))
    out["K_vs_oracle"] = {"pattern_equal": bool(same), "blocks": int(len(col)),
                          "max_rel_diff": dk,
                          "F_max_abs_diff_over_max_F": float(np.abs(f.get_sollicitation() - F).max() / np.abs(F).max())}
    del grp, gcol, gvals, vals
    log("GPU K vs oracle K: pattern %s, max |dK| / max(max|diag K|, |K|) = %.2e" % (
        "equal" if same else "DIFFERS", out["K_vs_oracle"]["max_rel_diff"]))
    # batch members
    out["batch"] = []
    for i, bc, be, bE, bnu, bfaces, btr, bdofs in batch_items:
        brp, bcol, bvals, bF, _ = O.assemble(L.HEXA, bc, be, bE, bnu, faces=bfaces, face_trac=btr)
        O.penalty(brp, bcol, bvals, bF, bdofs, None, 1e5)
        bx, _, _, brc = O.pcg_jacobi(brp, bcol, bvals, bF, rel_tol=1e-12, block=True)
        out["batch"].append((i, bx, brc))
    # one core, serial restatement, bounded sample
    O.set_threads(1)
    sn = 48
    sc, se = L.make_rect(L.HEXA, (sn + 1,) * 3)
    sE = lognormal_E(len(se), 0.5, seed=0)
    sfaces, str_, sdofs = clamp_and_traction(L.HEXA, sc, se)
    spat = O.pattern(O.HEXA, len(sc), se)
    t0 = time.perf_counter()
    srp, scol, svals, sF, _ = O.assemble(O.HEXA, sc, se, sE, 0.3, faces=sfaces, face_trac=str_, pat=spat)
    O.penalty(srp, scol, svals, sF, sdofs, None, 1e5)
    t_asm1 = time.perf_counter() - t0
    its = 300
    t0 = time.perf_counter()
    O.pcg_jacobi(srp, scol, svals, sF, rel_tol=0.0, max_iter=its, block=True)
    out["one_core"] = {"sample": "Hexa8 %d^3 (%d elements, %d DOF)" % (sn, len(se), 3 * len(sc)),
                       "assembly_s": t_asm1, "s_per_iter": (time.perf_counter() - t0) / its,
                       "elements": len(se), "dof": 3 * len(sc)}
    return out


def cpu_baseline(cpu, tol, ndof, n_elem):
    """cpu_baseline from the CPU leg: the all-core oracle's assembly plus its PCG time to the
    relative residual tol the GPU headline solves to, on the full workload."""
    hist = cpu["hist"]
    k = int(np.argmax(hist[:, 0] <= tol)) if np.any(hist[:, 0] <= tol) else len(hist) - 1
    t_solve = float(hist[k, 1])
    t_total = cpu["assembly_s"] + t_solve
    one = cpu["one_core"]
    t_one = one["assembly_s"] * n_elem / one["elements"] + one["s_per_iter"] * (ndof / one["dof"]) * (k + 1)
    return {"value": round(ndof / t_total / 1e6, 4), "unit": "MDOF/s", "cores": cpu["threads"], "kind": "port",
            "sample": ("the whole workload (%d elements, %d DOF) on %d host threads (%s, %d CPUs visible): "
                       "oracle/fem_oracle.c (%s), coloured element loop %.2f s + block-Jacobi PCG to rel. res. %.0e "
                       "in %d iterations, %.2f s (from the PCG's per-iteration clock)" % (
                           n_elem, ndof, cpu["threads"], cpu["cpu_model"], cpu["host_cpus"], cpu["build"],
                           cpu["assembly_s"], tol, k + 1, t_solve)),
            "total_s": round(t_total, 3),
            "one_core": {"value": round(ndof / t_one / 1e6, 5), "unit": "MDOF/s", "cores": 1,
                         "sample": ("serial restatement (element-order scatter, as the reference's default "
                                    "nb_threads_assemble_matrix = 1) on %s: assembly %.2f s, %.4f s per PCG iteration "
                                    "(300 timed), scaled to the full workload and %d iterations = %.1f s" % (
                                        one["sample"], one["assembly_s"], one["s_per_iter"], k + 1, t_one)),
                         "est_total_s": round(t_one, 2)}}


def calibrated_solves(f, ndof, log_tag, rel_ref=1e-11):
    """FP64 reference solve (to rel_ref, or to where FP64 stagnates above it), then the largest
    relative-residual tolerance whose FP64 solution is within 0.5e-8 of it, then the MIXED
    tolerance reaching 1e-8 (tightened by 3x if needed)."""
    import lmtpp_amd as L
    from lmtpp_amd import _capi as A
    f.precision = L.FP64
    f.set_result(np.zeros(ndof))
    ok = f.solve_system(rel_tol=rel_ref)
    assert ok or f.last_stats.status == A.STAGNATED, f.last_stats.status
    log("%s: FP64 reference %d iters, status %d, rel res %.2e" % (log_tag, f.last_stats.iters, f.last_stats.status,
                                                                  f.last_stats.rel_res))
    u_ref = f.get_result()
    nref = np.linalg.norm(u_ref)

    def solve(prec, tol):
        f.precision = prec
        f.set_result(np.zeros(ndof))
        ok = f.solve_system(rel_tol=tol)
        return ok, f.last_stats, np.linalg.norm(f.get_result() - u_ref) / nref

    tol = None
    for cand in [1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11, 3e-12]:
        ok, st, err = solve(L.FP64, cand)
        if ok and err <= 0.5e-8:
            tol = cand
            break
    assert tol is not None, "no tolerance reached 1e-8 error"
    ok64, st64, err64 = solve(L.FP64, tol)
    mtol = tol
    for _ in range(5):
        okm, stm, errm = solve(L.MIXED, mtol)
        if okm and errm <= 1e-8:
            break
        mtol /= 3
    log("%s: tol %.0e FP64 %d it %.1f ms err %.1e | MIXED tol %.0e %d it %d refreshes %.1f ms err %.1e" % (
        log_tag, tol, st64.iters, st64.solve_ms, err64, mtol, stm.iters, stm.refreshes, stm.solve_ms, errm))
    return u_ref, tol, mtol, (st64, err64), (stm, errm), solve


def config0():
    import lmtpp_amd as L
    from problems import build
    c, e = L.make_rect(L.TETRA, (11, 11, 11))
    f, *_ = build(L.TETRA, c, e, 210000.0)
    f.assemble()
    F = f.get_sollicitation()
    f.precision = L.FP64
    f.set_result(np.zeros_like(F))
    crit = 1e-10 * np.abs(F).max()
    ok = f.solve_system(crit=crit)   # the reference's stop test
    st = f.last_stats
    # the reference's own preconditioner: block IC(0) in natural order (its factor, on the GPU)
    f.precond = L.BLOCK_IC0
    f.set_result(np.zeros_like(F))
    ok_ic = f.solve_system(crit=crit)
    st_ic = f.last_stats
    setup_ic = f.precond_setup_ms()
    return {"workload": "Tetra4 11^3 points (6000 elements, 3993 DOF), E=210000, nu=0.3, clamp x=0, traction x=1; "
                        "FP64 PCG with the reference stop test max|r| <= 1e-10 max|F|",
            "ok": bool(ok), "iters": st.iters, "solve_ms": round(st.solve_ms, 3), "assembly_ms": round(f.assemble_ms(), 4),
            "max_abs_res": st.max_abs_res,
            "block_ic0_natural": {"ok": bool(ok_ic), "iters": st_ic.iters, "solve_ms": round(st_ic.solve_ms, 3),
                                  "factor_ms": round(setup_ic, 3), "max_abs_res": st_ic.max_abs_res},
            "reference_ic0_pcg_iters": int(np.load(os.path.join(ROOT, "tests", "golden",
                                                               "ref_tinyblock_pcg_tetra11.npz"))["iters"])}


def config2():
    import lmtpp_amd as L
    from problems import build, tetra10_jittered
    c, e = tetra10_jittered(40, seed=0)
    f, *_ = build(L.TETRA10, c, e, 210000.0, bf=(0.0, 0.0, -9.81 * 7.8e-9))
    sz = f.sizes()
    ndof = 3 * sz["n_nodes"]
    f.assemble()
    cfg = f.autotune()   # product (stored SpMV variant or matrix-free) and element-loop variant
    asm = []
    for _ in range(3):
        f.assemble()
        asm.append(f.assemble_ms())
    # the product alone: the autotuner's pick beside the stored matrix's fastest variant
    prod_ms = {"picked": {"fp64": round(f.time_spmv(L.FP64, 30), 4), "fp32": round(f.time_spmv(L.MIXED, 30), 4)}}
    f.set_spmv_config(fp64=(0, 2, 2), fp32=(0, 3, 2))
    prod_ms["stored"] = {"fp64": round(f.time_spmv(L.FP64, 30), 4), "fp32": round(f.time_spmv(L.MIXED, 30), 4)}
    f.set_spmv_config(fp64=cfg["fp64"], fp32=cfg["fp32"])
    f.precond = L.JACOBI
    u_ref, tol, mtol, (st64, err64), (stm, errm), solve = calibrated_solves(f, ndof, "cfg2")
    # block Jacobi at the same tolerances (fewer FP64 iterations, more FP32 ones on this mesh)
    f.precond = L.BLOCK_JACOBI
    _, bst64, berr64 = solve(L.FP64, tol)
    _, bstm, berrm = solve(L.MIXED, mtol)
    a = float(np.median(asm))
    return {"workload": "Tetra_10 40^3 cells x 6 (384,000 elements, %d DOF), interior corner nodes jittered +-0.2h, "
                        "E=210000, nu=0.3, body force (0,0,-9.81*7.8e-9) + traction; Jacobi PCG (block Jacobi beside)" % ndof,
            "elements": sz["n_elem"], "dof": ndof, "blocks": sz["nnzb"], "assembly_ms": round(a, 3),
            "product": {"fp64": "matrix-free" if cfg["fp64"][0] == 2 else "stored",
                        "fp32": "matrix-free" if cfg["fp32"][0] == 2 else "stored"},
            "product_ms": prod_ms,
            "assembly_melem_s": round(sz["n_elem"] / (a * 1e-3) / 1e6, 2),
            "fp64_solve_ms": round(st64.solve_ms, 2), "fp64_iters": st64.iters, "rel_err_fp64": err64,
            "mixed_solve_ms": round(stm.solve_ms, 2), "mixed_iters": stm.iters, "rel_err_mixed": errm,
            "mixed_vs_fp64_speedup": round(st64.solve_ms / stm.solve_ms, 3),
            "solve_mdof_s_mixed": round(ndof / (stm.solve_ms * 1e-3) / 1e6, 3),
            "block_jacobi": {"fp64_solve_ms": round(bst64.solve_ms, 2), "fp64_iters": bst64.iters, "rel_err_fp64": berr64,
                             "mixed_solve_ms": round(bstm.solve_ms, 2), "mixed_iters": bstm.iters, "rel_err_mixed": berrm}}


def config3(steps=50):
    """Adjoint identification: Hexa8 64^3, 262,144 per-element moduli; Adam on log E with
    gradients from femgpu_adjoint (one forward + one adjoint solve per step)."""
    import lmtpp_amd as L
    from problems import build, config3 as config3_problem
    c, e, E_true, odofs, _ = config3_problem(n=64)
    f, *_ = build(L.HEXA, c, e, E_true)
    f.assemble()
    f.autotune()                             # product variant for this mesh (matrix-free on Hexa)
    assert f.solve(rel_tol=1e-11)
    u_true = f.get_result()
    u_meas = u_true[odofs]
    logE = np.full(len(e), np.log(210000.0))
    m1 = np.zeros_like(logE)
    m2 = np.zeros_like(logE)
    lr, b1, b2 = 0.02, 0.9, 0.999
    scale = 1.0 / np.dot(u_meas, u_meas)
    hist, t_steps = [], []
    u = np.zeros(3 * len(c))
    for it in range(steps):
        t0 = time.perf_counter()
        E = np.exp(logE)
        f.set_material(E, 0.3)
        f.set_result(u)                      # previous solution as initial guess (shift())
        assert f.solve(rel_tol=1e-10)
        u = f.get_result()
        r = u[odofs] - u_meas
        J = 0.5 * np.dot(r, r)
        dJdu = np.zeros_like(u)
        dJdu[odofs] = r
        gE = f.adjoint(dJdu, rel_tol=1e-10)
        g = gE * E * scale                   # d(J*scale)/d(log E)
        m1 = b1 * m1 + (1 - b1) * g
        m2 = b2 * m2 + (1 - b2) * g * g
        logE -= lr * (m1 / (1 - b1 ** (it + 1))) / (np.sqrt(m2 / (1 - b2 ** (it + 1))) + 1e-12)
        t_steps.append(time.perf_counter() - t0)
        hist.append(J)
    return {"workload": "Hexa8 64^3 (262,144 elements, 823,875 DOF); E_true = 210000 exp(0.3 z~), z~ = 3 passes of the "
                        "27-point box filter on N(0,1) (zero-flux edges, rescaled to unit std); 5%% of top-face nodes "
                        "observed (3 comps); J = 1/2 ||u_obs - u_meas||^2; %d Adam steps on log E (lr 0.02), forward "
                        "(warm-started) + adjoint FP64 solves to 1e-10 per step" % steps,
            "params": len(e), "steps": steps, "J_first": hist[0], "J_last": hist[-1], "J_ratio": hist[-1] / hist[0],
            "s_per_step": round(float(np.median(t_steps)), 4), "total_s": round(float(np.sum(t_steps)), 2),
            "rel_err_E_first": None, "rel_err_E_last": float(np.linalg.norm(np.exp(logE) - E_true) / np.linalg.norm(E_true))}


def config4_inputs(n_models=256):
    """The 256 cubes of the batch: (meshes [(coords, conn)], items [(E, nu, traction)])."""
    import lmtpp_amd as L
    rng = np.random.default_rng(4)
    meshes, items = [], []
    for i in range(n_models):
        side = int(rng.integers(8, 33))
        c, e = L.make_rect(L.HEXA, (side + 1,) * 3)
        E = 210000.0 * np.exp(0.5 * rng.standard_normal(len(e)))
        nu = float(rng.uniform(0.2, 0.4))
        tr = rng.standard_normal(3)
        meshes.append((c, e))
        items.append((E, nu, tr))
    return meshes, items


def config4_oracle_members(meshes, items, k=8, seed=5):
    """k seeded members of the batch with their clamp and traction, for the oracle's solves."""
    import lmtpp_amd as L
    from problems import clamp_and_traction
    pick = np.random.default_rng(seed).choice(len(meshes), size=k, replace=False)
    out = []
    for i in sorted(int(v) for v in pick):
        (c, e), (E, nu, tr) = meshes[i], items[i]
        faces, trs, dofs = clamp_and_traction(L.HEXA, c, e, tr)
        out.append((i, c, e, E, nu, faces, trs, dofs))
    return out


def config4(meshes, items, oracle_refs=()):
    """256 independent Hexa8 cubes of uneven size, assembled and solved (MIXED) in one batch."""
    import lmtpp_amd as L
    from problems import clamp_and_traction
    n_models = len(meshes)
    t0 = time.perf_counter()
    b = L.Batch(L.HEXA, meshes)
    setup = time.perf_counter() - t0
    ndof = 0
    for i, ((c, e), (E, nu, tr)) in enumerate(zip(meshes, items)):
        faces, trs, dofs = clamp_and_traction(L.HEXA, c, e, tr)
        b.set_material(i, E, nu)
        b.set_traction(i, faces, trs)
        b.add_constraints(i, dofs, None, 1e5)
        ndof += 3 * len(c)
    b.assemble()
    # FP64 references (tight), then the timed MIXED batch solve
    b.clear_result()
    ok, _ = b.solve(rel_tol=1e-11, precision=L.FP64, max_iter=40000)
    refs = [b.get_result(i) for i in range(n_models)]
    # product variant and element loop chosen inside the library by timing whole batched solves
    prod = b.autotune(rel_tol=1e-10, precision=L.MIXED, precond=L.BLOCK_JACOBI, max_iter=40000)
    t_asm = []
    for _ in range(2):
        b.assemble()
        t_asm.append(b.assemble_ms())
    times, polls = [], []
    for _ in range(2):
        b.assemble()
        b.clear_result()
        t0 = time.perf_counter()
        ok, ms = b.solve(rel_tol=1e-10, precision=L.MIXED, precond=L.BLOCK_JACOBI, max_iter=40000)
        times.append((time.perf_counter() - t0, ms))
        polls.append(b.chunk_stats())
    errs = [np.linalg.norm(b.get_result(i) - refs[i]) / np.linalg.norm(refs[i]) for i in range(n_models)]
    oerr = {int(i): float(np.linalg.norm(b.get_result(i) - x) / np.linalg.norm(x)) for i, x, rc in oracle_refs if rc == 0}
    ms = float(np.median([t[1] for t in times]))
    iters = [st.iters for st in b.last_stats]
    return {"workload": "256 Hexa8 cubes, side ~ U{8..32} cells (seed 4), lognormal E (sigma 0.5), nu ~ U[0.2,0.4], "
                        "traction ~ N(0,1) on x=1, clamp x=0; one batch: per-problem MIXED block-Jacobi PCG to rel. res. 1e-10, "
                        "iteration chunks replayed from a captured hipGraph; product and element loop from femgpu_batch_autotune",
            "models": n_models, "total_dof": ndof, "batch_setup_s": round(setup, 2),
            "assembly_ms": round(float(np.median(t_asm)), 3), "solve_ms": round(ms, 2),
            "aggregate_mdof_s": round(ndof / (ms * 1e-3) / 1e6, 3), "max_rel_err_vs_fp64": float(max(errs)),
            "rel_err_vs_oracle": oerr, "max_rel_err_vs_oracle": max(oerr.values()) if oerr else None,
            "median_rel_err": float(np.median(errs)), "iters_min": int(min(iters)), "iters_max": int(max(iters)),
            "all_ok": bool(ok), "product": {2: "stored, nontemporal_col16", 1: "stored, nontemporal", 0: "stored, plain",
                                            -1: "matrix-free"}[prod],
            "product_choice": "femgpu_batch_autotune",
            "host_poll": {"chunks": polls[-1][0], "device_ms_in_chunks": round(polls[-1][1], 2),
                          "gaps": polls[-1][2], "gap_ms": round(polls[-1][3], 3),
                          "gap_us_per_chunk": round(1e3 * polls[-1][3] / max(1, polls[-1][2]), 1),
                          # that gap at every chunk transition (refreshes add device work there)
                          "est_fraction_of_solve": round(polls[-1][3] / max(1, polls[-1][2]) *
                                                         max(0, polls[-1][0] - 1) / ms, 4)},
            "assembly": {0xffffffff: "staged", 0xfffffffe: "gather", 0xfffffffd: "gather_box"}.get(
                b.assembly_rows()[0], "rows %d" % b.assembly_rows()[0])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="cells per side (Hexa8 n^3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-configs", action="store_true", help="skip configurations 0, 2, 3, 4")
    ap.add_argument("--configs", default="0,2,3,4", help="which of configurations 0, 2, 3, 4 to run (comma list)")
    ap.add_argument("--no-precond", action="store_true", help="skip the preconditioner comparison")
    ap.add_argument("--no-stepping", action="store_true", help="skip the time-stepping measurement")
    ap.add_argument("--no-natural-ic0", action="store_true",
                    help="skip the natural-order block IC(0) solve (hundreds of level launches per iteration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; a multi-GPU run would give each rank whole "
              "sub-problems (batches) or a node-range shard with halo exchange, not run here "
              "(--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import lmtpp_amd as L
    from problems import build, lognormal_E

    n = args.n
    kind = L.HEXA
    t0 = time.perf_counter()
    c, e = L.make_rect(kind, (n + 1,) * 3)
    E = lognormal_E(len(e), 0.5, seed=0)
    f, faces, tr, dofs = build(kind, c, e, E)
    sz = f.sizes()
    ndof = 3 * sz["n_nodes"]
    log("setup %.1f s: %d elem, %d nodes, %d DOF, %d blocks, %d colours" % (
        time.perf_counter() - t0, sz["n_elem"], sz["n_nodes"], ndof, sz["nnzb"], sz["n_colors"]))

    # -- references (untimed) ------------------------------------------------------------------
    f.assemble()
    spmv_cfg = f.autotune()
    fmt_name = {0: "full", 1: "sym", 2: "matrix-free"}
    log("autotuned SpMV: FP64 %s/%d/%d, FP32 %s/%d/%d (format / SELL slots per loop trip / loads)" % (
        fmt_name[spmv_cfg["fp64"][0]], spmv_cfg["fp64"][1], spmv_cfg["fp64"][2], fmt_name[spmv_cfg["fp32"][0]],
        spmv_cfg["fp32"][1], spmv_cfg["fp32"][2]))
    f.precision, f.precond = L.FP64, L.BLOCK_JACOBI
    f.max_iter = 40000
    f.set_result(np.zeros(ndof))
    gpu_ref_tol = 3e-11
    assert f.solve_system(rel_tol=gpu_ref_tol), f.last_stats.status
    assert f.last_stats.rel_res <= gpu_ref_tol
    u_gpu64 = f.get_result()
    gpu_ref = {"rel_tol": gpu_ref_tol, "iters": int(f.last_stats.iters), "rel_res": f.last_stats.rel_res,
               "true_residual_checks": int(f.last_stats.refreshes), "solve_ms": round(f.last_stats.solve_ms, 2)}
    log("FP64 GPU reference solve: %d iters, true rel res %.2e (tol %.0e), %.1f ms" % (
        f.last_stats.iters, f.last_stats.rel_res, gpu_ref_tol, f.last_stats.solve_ms))
    # the CPU leg: the oracle on the host cores, whose solution is the FP64 oracle solution
    wanted = set() if args.no_configs else set(args.configs.split(","))
    c4 = config4_inputs() if "4" in wanted else None
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_leg(kind, c, e, E, faces, tr, dofs, f,
                          config4_oracle_members(*c4) if c4 else ())
        except Exception as ex:  # the baseline is reported, never the measured value
            log("CPU leg failed: %r" % (ex,))
    if cpu is not None:
        u_ref, ref_name = cpu["u"], "FP64 oracle (oracle/fem_oracle.c block-Jacobi PCG to 1e-12, host CPU)"
        gpu_ref["rel_err_vs_oracle"] = float(np.linalg.norm(u_gpu64 - u_ref) / np.linalg.norm(u_ref))
        log("FP64 GPU reference vs oracle: rel. err %.2e" % gpu_ref["rel_err_vs_oracle"])
    else:
        u_ref, ref_name = u_gpu64, "FP64 GPU solve to a true relative residual of %.0e" % gpu_ref_tol
    nref = np.linalg.norm(u_ref)

    def solve(prec, tol, precond=L.BLOCK_JACOBI):
        f.precision, f.precond = prec, precond
        f.set_result(np.zeros(ndof))
        ok = f.solve_system(rel_tol=tol)
        st = f.last_stats
        err = np.linalg.norm(f.get_result() - u_ref) / nref
        return ok, st, err

    tol = None
    for cand in [1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11]:
        ok, st, err = solve(L.FP64, cand)
        log("FP64 tol %.0e: %d iters, err %.2e" % (cand, st.iters, err))
        if ok and err <= 0.5e-8:
            tol = cand
            break
    assert tol is not None, "no tolerance reached 1e-8 error"
    ok, st64, err64 = solve(L.FP64, tol)
    mtol = tol
    for _ in range(4):
        ok, stm, errm = solve(L.MIXED, mtol)
        log("MIXED tol %.0e: %d iters, %d refreshes, err %.2e, %.1f ms" % (mtol, stm.iters, stm.refreshes, errm,
                                                                           stm.solve_ms))
        if ok and errm <= 1e-8:
            break
        mtol /= 3
    assert errm <= 1e-8

    # -- timed steps ---------------------------------------------------------------------------
    f.precision, f.precond = L.MIXED, L.BLOCK_JACOBI

    def step():
        f.assemble()
        f.set_result(np.zeros(ndof))
        ok = f.solve_system(rel_tol=mtol)
        assert ok
        return f.last_stats

    for _ in range(args.warmup):
        step()
    asm_ms, sol_ms, errs = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        asm_ms.append(f.assemble_ms())
        sol_ms.append(st.solve_ms)
    wall = time.perf_counter() - t0
    err_timed = np.linalg.norm(f.get_result() - u_ref) / nref
    ms_step = wall / args.steps * 1e3
    value = ndof / (ms_step / 1e3) / 1e6

    # FP64-only solve at the same error target, for the FP64 vs mixed comparison
    fp64_ms = []
    for _ in range(max(1, min(3, args.steps))):
        ok, st, err = solve(L.FP64, tol)
        fp64_ms.append(st.solve_ms)

    # -- the other preconditioners at the same tolerances ------------------------------------------
    precond = {}
    for name, pc, precs in () if args.no_precond else (("jacobi", L.JACOBI, (L.FP64, L.MIXED)), ("block_jacobi", L.BLOCK_JACOBI, (L.FP64, L.MIXED)),
                            ("block_ic0_coloured", L.BLOCK_IC0_COLOURED, (L.FP64, L.MIXED)),
                            ("block_ic0_natural", L.BLOCK_IC0, () if args.no_natural_ic0 else (L.FP64,))):
        row = {}
        for prec in precs:
            ok, st, err = solve(prec, tol if prec == L.FP64 else mtol, pc)
            key = "fp64" if prec == L.FP64 else "mixed"
            row[key] = {"ok": bool(ok), "iters": int(st.iters), "solve_ms": round(st.solve_ms, 2),
                        "setup_ms": round(f.precond_setup_ms(), 3), "rel_err": float(err)}
            log("%s %s: %d iters %.1f ms (setup %.1f ms) err %.1e" % (name, key, st.iters, st.solve_ms,
                                                                      f.precond_setup_ms(), err))
        precond[name] = row

    # -- after-solve recovery (strain, stress, trace per element) ------------------------------------
    f.precision, f.precond = L.FP64, L.BLOCK_JACOBI
    f.set_result(u_ref)
    as_ms = []
    for _ in range(5):
        f.call_after_solve()
        as_ms.append(f.after_solve_ms())
    as_bytes = sz["n_nodes"] * 48 + sz["n_elem"] * (4 * 8 + 8 + 13 * 8)
    a_ms = float(np.median(as_ms))

    # -- time steps (ProblemAncestor::iteration): traction ramp, constant K, warm starts ----------
    ts_rows = []
    if not args.no_stepping:
        f.precision, f.precond = L.MIXED, L.BLOCK_JACOBI
        f.assume_constant_matrix = True
        f.set_initial_time_step(1.0)
        f.set_result(np.zeros(ndof))
        prob = L.Problem([f], rel_tol=mtol)
        for k in range(1, 6):
            f.set_traction(faces, k * tr)
            t1 = time.perf_counter()
            prob.iteration()
            wall_k = (time.perf_counter() - t1) * 1e3
            st = f.last_stats
            ts_rows.append({"step": k, "time": prob.time, "iters": int(st.iters), "solve_ms": round(st.solve_ms, 2),
                            "assembly_ms": round(f.assemble_ms(), 3), "wall_ms": round(wall_k, 2),
                            "rel_err": float(np.linalg.norm(f.get_result() - k * u_ref) / (k * nref))})
            log("time step %d: %d iters, solve %.1f ms, assembly %.2f ms, wall %.1f ms, err %.1e" % (
                k, st.iters, st.solve_ms, f.assemble_ms(), wall_k, ts_rows[-1]["rel_err"]))
        prob.close()
        f.assume_constant_matrix = False

    # -- roofline of the dominant kernel, HIP events on the problem's stream --------------------
    # The stored-format product (the full storage with the tuned loads when the autotuner picked
    # the matrix-free one) is HBM-bound; the matrix-free element-block kernel is bound by vector
    # arithmetic.
    from lmtpp_amd import _capi as A
    from lmtpp_amd.traffic import mf_block_bytes, mf_flops
    reps = 50
    mf = A.SPMV_MATRIX_FREE
    stored = {k: ((A.SPMV_FULL if v[0] == mf else v[0]), v[1], v[2]) for k, v in spmv_cfg.items()}
    f.set_spmv_config(fp64=stored["fp64"], fp32=stored["fp32"])
    t32 = f.time_spmv(L.MIXED, reps)
    t64 = f.time_spmv(L.FP64, reps)
    f.set_spmv_config(fp64=spmv_cfg["fp64"], fp32=spmv_cfg["fp32"])
    sfmt32, sfmt64 = stored["fp32"][0], stored["fp64"][0]
    b32 = spmv_bytes(sz["nnzb"], sz["n_nodes"], True, sfmt32, stored["fp32"][2] == 2)
    b64 = spmv_bytes(sz["nnzb"], sz["n_nodes"], False, sfmt64, stored["fp64"][2] == 2)
    load_key = {0: "_plain", 1: "", 2: "_col16"}
    ach32 = b32 / (t32 * 1e-3) / 1e9
    ach64 = b64 / (t64 * 1e-3) / 1e9
    roof_stored32 = {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic("spmv_fp32" + ("_sym" if sfmt32 else load_key[stored["fp32"][2]])),
                     "kernel": "k_spmv<float,0,%d,%d,%d> (FP32 SELL-64 block SpMV, %s storage, loads %d)" % (
                         stored["fp32"][1], sfmt32, stored["fp32"][2], fmt_name[sfmt32], stored["fp32"][2]),
                     "avg_launch_ms": round(t32, 4), "algorithmic_bytes": b32}
    roof_stored64 = {"achieved": round(ach64, 1), "frac": round(ach64 / HBM_PEAK_GBS, 4),
                     "avg_launch_ms": round(t64, 4), "algorithmic_bytes": b64,
                     "traffic": load_traffic("spmv_fp64" + ("_sym" if sfmt64 else load_key[stored["fp64"][2]])),
                     "kernel": "k_spmv<double,0,%d,%d,%d>" % (stored["fp64"][1], sfmt64, stored["fp64"][2])}

    def mf_roof(fp32):
        # The element-block kernel of the matrix-free product.  Its arithmetic intensity (about
        # 10 FLOP/B in FP32 on box elements) sits at the ridge of the unpacked vector rate against
        # HBM (78.6 TFLOP/s / 8 TB/s), so both roofs are given; "achieved" is the HBM one
        # (algorithmic bytes / launch time), the vector one beside it.
        tb, tp = f.time_matrix_free(L.MIXED if fp32 else L.FP64, reps)
        info = f.matrix_free_info()
        flops = mf_flops(sz["n_elem"], info)
        nbytes = mf_block_bytes(sz["n_elem"], sz["n_nodes"], info, fp32)
        peak = FP32_VECTOR_PEAK_TFLOPS if fp32 else FP64_VECTOR_PEAK_TFLOPS
        ach = flops / (tb * 1e-3) / 1e12
        gbs = nbytes / (tb * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(gbs / HBM_PEAK_GBS, 4),
             "traffic": load_traffic("mf_block_" + ("fp32" if fp32 else "fp64")),
             "kernel": "k_mf_block<%s,false,false,%s,2,false,true> (%s matrix-free Hexa product, element blocks)" % (
                 "float" if fp32 else "double", "true" if info["box_elements"] == sz["n_elem"] else "false",
                 "FP32" if fp32 else "FP64"),
             "avg_launch_ms": round(tb, 4), "algorithmic_bytes": nbytes, "algorithmic_flops": flops,
             "flop_per_byte": round(flops / nbytes, 2),
             "valu": {"achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s", "frac": round(ach / peak, 4)},
             "product_ms": round(tp, 4), "layout": info}
        if fp32:  # packed FP32 issue is disabled in this build: the reachable vector rate is the FP64 one
            r["valu"]["frac_of_unpacked_fp32"] = round(ach / FP64_VECTOR_PEAK_TFLOPS, 4)
        return r

    def iter_kernel_roof(fp32):
        # The dominant kernel of a step: the CG iteration's element blocks (the matrix-free product
        # with the direction update and the previous vector update's end folded in), its device
        # time from femgpu_time_iteration (50 iterations after 4 warm-up ones: the kernels' shares
        # from HIP events around each launch, the iteration's time from events around its chunks
        # on the problem's stream); its algorithmic bytes from mf_block_iter_bytes.
        from lmtpp_amd.traffic import mf_block_iter_bytes
        tb, tu = f.time_iteration(L.MIXED if fp32 else L.FP64, 50)
        info = f.matrix_free_info()
        flops = mf_flops(sz["n_elem"], info)
        nbytes = mf_block_iter_bytes(sz["n_elem"], sz["n_nodes"], info, fp32)
        peak = FP32_VECTOR_PEAK_TFLOPS if fp32 else FP64_VECTOR_PEAK_TFLOPS
        ach = flops / (tb * 1e-3) / 1e12
        gbs = nbytes / (tb * 1e-3) / 1e9
        boxes = "true" if info["box_elements"] == sz["n_elem"] else "false"
        r = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(gbs / HBM_PEAK_GBS, 4),
             "traffic": load_traffic("mf_block_iter_" + ("fp32" if fp32 else "fp64")),
             "kernel": "k_mf_block<%s,true,true,%s,2,true,true> (%s CG iteration: matrix-free Hexa product with the "
                       "direction update and the previous vector update's end folded in)" % (
                           "float" if fp32 else "double", boxes, "FP32" if fp32 else "FP64"),
             "avg_launch_ms": round(tb, 4), "algorithmic_bytes": nbytes, "algorithmic_flops": flops,
             "flop_per_byte": round(flops / nbytes, 2),
             "valu": {"achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s", "frac": round(ach / peak, 4)},
             "vector_update_ms": round(tu, 4), "layout": info}
        if fp32:
            r["valu"]["frac_of_unpacked_fp32"] = round(ach / FP64_VECTOR_PEAK_TFLOPS, 4)
        return r

    prod32 = mf_roof(True) if spmv_cfg["fp32"][0] == mf else None
    prod64 = mf_roof(False) if spmv_cfg["fp64"][0] == mf else None
    roof32 = iter_kernel_roof(True) if prod32 else roof_stored32
    roof64 = iter_kernel_roof(False) if prod64 else roof_stored64

    def iter_roof(prec, tol_p):
        # A whole CG iteration (block Jacobi, matrix-free product: the element-block kernel with
        # the direction update and the previous update's end folded in, then the vector update):
        # its algorithmic bytes over its device time, the iteration chunks' time (HIP events on
        # the problem's stream, host polls excluded) over the iteration count.
        from lmtpp_amd.traffic import mf_block_iter_bytes, update_iter_bytes
        ok, st, err = solve(prec, tol_p)
        nch, cms, ngaps, gms = f.chunk_stats()
        fp32 = prec == L.MIXED
        info = f.matrix_free_info()
        kb = mf_block_iter_bytes(sz["n_elem"], sz["n_nodes"], info, fp32)
        ub = update_iter_bytes(sz["n_nodes"], info, fp32, 1)
        us = cms * 1e3 / max(1, st.iters)
        gbs = (kb + ub) / (us * 1e-6) / 1e9
        key = "fp32" if fp32 else "fp64"
        t1, t2 = load_traffic("mf_block_iter_" + key), load_traffic("update_iter_" + key)
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "us_per_iteration": round(us, 2),
                "algorithmic_bytes": {"element_blocks": kb, "vector_update": ub},
                "traffic": (t1 + t2) if t1 and t2 else None,
                "kernels": "k_mf_block<%s,true,true,true,2,true,true> + k_update_mf<1,%s,true>" % (
                    ("float", "float") if fp32 else ("double", "double")),
                "iters": int(st.iters), "chunks": int(nch), "chunk_ms": round(cms, 2), "host_gap_ms": round(gms, 3)}

    roof_iter = None
    if spmv_cfg["fp32"][0] == mf and spmv_cfg["fp64"][0] == mf:
        f.precond = L.BLOCK_JACOBI
        roof_iter = {"mixed": iter_roof(L.MIXED, mtol), "fp64": iter_roof(L.FP64, tol)}
        log("iteration: MIXED %.1f us (%.2f of HBM), FP64 %.1f us (%.2f of HBM)" % (
            roof_iter["mixed"]["us_per_iteration"], roof_iter["mixed"]["frac"], roof_iter["fp64"]["us_per_iteration"],
            roof_iter["fp64"]["frac"]))

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "MDOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-inner/f64-refined",
        "data": "synthetic (make_rect Hexa lattice, lognormal E from seed 0)",
        "config": {"workload": "Hexa8 %d^3 elasticity, E_e=210000*exp(0.5 z_e), nu=0.3, clamp x=0 (penalty 1e5), "
                               "traction (0,0,-1) on x=1; step = GPU assembly + mixed block-Jacobi PCG to 1e-8 rel. err"
                               % n,
                   "elements": sz["n_elem"], "nodes": sz["n_nodes"], "dof": ndof, "blocks": sz["nnzb"],
                   "parallelism": "single GPU"},
        "mixed_solve_ms": round(float(np.median(sol_ms)), 3),
        "fp64_solve_ms": round(float(np.median(fp64_ms)), 3),
        "mixed_vs_fp64_speedup": round(float(np.median(fp64_ms) / np.median(sol_ms)), 3),
        "solve_mdof_s_mixed": round(ndof / (np.median(sol_ms) * 1e-3) / 1e6, 3),
        "solve_mdof_s_fp64": round(ndof / (np.median(fp64_ms) * 1e-3) / 1e6, 3),
        "assembly_ms": round(float(np.median(asm_ms)), 3),
        "assembly_melem_s": round(sz["n_elem"] / (np.median(asm_ms) * 1e-3) / 1e6, 3),
        "rel_err_mixed": float(err_timed), "rel_err_fp64": float(err64), "rel_err_against": ref_name,
        "fp64_gpu_reference": gpu_ref,
        "rel_res_tol": mtol, "iters_mixed": int(stm.iters), "refreshes_mixed": int(stm.refreshes),
        "iters_fp64": int(st64.iters),
        "spmv_formats": {"fp64": fmt_name[spmv_cfg["fp64"][0]], "fp32": fmt_name[spmv_cfg["fp32"][0]]},
        "roofline": roof32,
        "roofline_fp64": roof64,
        "roofline_product": prod32,
        "roofline_fp64_product": prod64,
        "roofline_iteration": roof_iter,
        "roofline_stored_spmv": roof_stored32,
        "roofline_stored_spmv_fp64": roof_stored64,
        "preconditioners": precond,
        "time_stepping": {"workload": "5 steps of ProblemAncestor::iteration, traction x step, assume_constant_matrix "
                                      "(K, its FP32 copy and the block-Jacobi factors built at step 1; steps 2-5 "
                                      "assemble F only), MIXED PCG warm-started from the previous step",
                          "steps": ts_rows} if ts_rows else None,
        "after_solve": {"kernel": "k_after_solve<2>", "ms": round(a_ms, 4),
                        "melem_s": round(sz["n_elem"] / (a_ms * 1e-3) / 1e6, 1), "algorithmic_bytes": as_bytes,
                        "achieved_gbs": round(as_bytes / (a_ms * 1e-3) / 1e9, 1)},
    }
    from lmtpp_amd import _capi as A
    gbs, _ = A.stream_copy_gbs(1 << 27, 20)
    rgbs, _ = A.stream_read_gbs(1 << 28, 20)
    out["stream_copy_gbs"] = round(gbs, 1)
    out["stream_read_gbs"] = round(rgbs, 1)
    out["roofline_stored_spmv"]["frac_of_measured_copy"] = round(ach32 / gbs, 4)
    out["roofline_stored_spmv"]["frac_of_measured_read"] = round(ach32 / rgbs, 4)
    f.close()
    if wanted:
        out["configs"] = {}
        oracle_refs = cpu["batch"] if cpu is not None else ()
        for name, fn in (("0", config0), ("2", config2), ("3", config3),
                         ("4", lambda: config4(c4[0], c4[1], oracle_refs))):
            if name not in wanted:
                continue
            try:
                t0 = time.perf_counter()
                out["configs"][name] = fn()
                log("config %s done in %.1f s" % (name, time.perf_counter() - t0))
            except Exception as ex:  # a secondary configuration never hides the headline line
                log("config %s failed: %r" % (name, ex))
                out["configs"][name] = {"error": repr(ex)}
    if cpu is not None:
        out["cpu_baseline"] = cpu_baseline(cpu, mtol, ndof, sz["n_elem"])
        out["oracle_check"] = {"true_rel_res": cpu["true_rel_res"], "pcg_iters": cpu["pcg_iters"],
                               "K_vs_oracle": cpu["K_vs_oracle"]}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
