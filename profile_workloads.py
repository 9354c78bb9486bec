"""Workloads run under rocprofv3 for the summaries in profiles/ (one counter set per run).

    python -u profile_workloads.py <workload> [--n 100]

workloads, each on the Hexa8 n^3 lattice with a lognormal (sigma 0.5) modulus, the x=0 face
clamped and a traction on x=1 (the elastic-cube workload of bench.py):

  traffic    the plain copy kernel (k_copy of lmtpp_amd.traffic.COPY_DOUBLES doubles), then 5 FP64
             and 5 FP32 block SpMV launches per storage variant (full storage with non-temporal
             loads, with 16-bit column offsets, with plain loads; symmetric storage), then the
             matrix-free product (element-block kernel and shared-node pass), then 64 CG
             iterations (block Jacobi, matrix-free) in MIXED and in FP64 (the iteration's
             element-block kernel and vector update);
  assembly   three assemblies with each element-loop variant: staged colour passes, gather, box
             gather;
  matfree    the matrix-free product: its element-block kernel and the whole product, 3 launches
             each, FP32 then FP64;
  iteration  the solves bench.py times: the box matrix-free product with block Jacobi, MIXED then
             FP64 (--precision mixed or fp64: that one alone), to a relative residual of 1e-8 from
             a zero initial guess.

FETCH_SIZE / WRITE_SIZE runs of `traffic` are read back with `python -m lmtpp_amd.traffic parse`.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))

import numpy as np  # noqa: E402

import lmtpp_amd as L  # noqa: E402
from lmtpp_amd import _capi as A  # noqa: E402
from lmtpp_amd.traffic import COPY_DOUBLES  # noqa: E402
from problems import build, lognormal_E  # noqa: E402


def _cube(n):
    c, e = L.make_rect(L.HEXA, (n + 1,) * 3)
    f, *_ = build(L.HEXA, c, e, lognormal_E(len(e), 0.5, 0))
    return f, c


def traffic(n):
    gbs, ms = A.stream_copy_gbs(COPY_DOUBLES, 5)
    print("copy: %.1f GB/s (%.3f ms per launch)" % (gbs, ms), flush=True)
    f, c = _cube(n)
    f.assemble()
    f.set_result(np.zeros(3 * len(c)))
    sz = f.sizes()
    for fmt, nt in ((A.SPMV_FULL, A.LOADS_NT), (A.SPMV_FULL, A.LOADS_NT_COL16), (A.SPMV_FULL, A.LOADS_PLAIN),
                    (A.SPMV_SYM, A.LOADS_PLAIN)):
        f.set_spmv_config(fp64=(fmt, 2, nt), fp32=(fmt, 2, nt))
        t64 = f.time_spmv(L.FP64, 5)
        t32 = f.time_spmv(L.MIXED, 5)
        print(json.dumps({"format": fmt, "loads": nt, "nnzb": sz["nnzb"], "n_nodes": sz["n_nodes"], "spmv_fp64_ms": t64,
                          "spmv_fp32_ms": t32}), flush=True)
    f.set_spmv_config(fp64=(A.SPMV_MATRIX_FREE, 2, A.LOADS_NT), fp32=(A.SPMV_MATRIX_FREE, 2, A.LOADS_NT))
    b64, p64 = f.time_matrix_free(L.FP64, 5)
    b32, p32 = f.time_matrix_free(L.MIXED, 5)
    print(json.dumps({"format": A.SPMV_MATRIX_FREE, "info": f.matrix_free_info(), "block_fp64_ms": b64,
                      "product_fp64_ms": p64, "block_fp32_ms": b32, "product_fp32_ms": p32}), flush=True)
    # the CG iteration's two kernels: 64 block-Jacobi iterations in each precision
    f.precond = L.BLOCK_JACOBI
    for prec in (L.MIXED, L.FP64):
        f.precision = prec
        f.max_iter = 64
        f.set_result(np.zeros(3 * len(c)))
        f.solve_system(rel_tol=1e-30)
        print("iterations (precision %d): %d, status %d" % (prec, f.last_stats.iters, f.last_stats.status), flush=True)
    f.close()


def assembly(n):
    f, _ = _cube(n)
    for name, R in (("staged", A.ASSEMBLY_STAGED), ("gather", A.ASSEMBLY_GATHER),
                    ("gather_box", A.ASSEMBLY_GATHER_BOX)):
        f.set_assembly_rows(R)
        ts = []
        for _ in range(3):
            f.assemble()
            ts.append(f.assemble_ms())
        print("%s: assembly %s ms" % (name, " ".join("%.3f" % t for t in ts)), flush=True)
    f.close()


def matfree(n):
    f, _ = _cube(n)
    f.assemble()
    f.set_spmv_config(fp64=(A.SPMV_MATRIX_FREE, 2, A.LOADS_NT), fp32=(A.SPMV_MATRIX_FREE, 2, A.LOADS_NT))
    for prec, name in ((L.MIXED, "fp32"), (L.FP64, "fp64")):
        b, p = f.time_matrix_free(prec, 3)
        print("%s: element blocks %.1f us, product %.1f us" % (name, b * 1e3, p * 1e3), flush=True)
    print(json.dumps(f.matrix_free_info()), flush=True)
    f.close()


def iteration(n, only=None):
    f, c = _cube(n)
    f.assemble()
    mf = (A.SPMV_MATRIX_FREE, 2, A.LOADS_NT)
    f.set_spmv_config(fp64=mf, fp32=mf)
    f.precond = L.BLOCK_JACOBI
    for prec, name in ((L.MIXED, "mixed"), (L.FP64, "fp64")):
        if only and name != only:
            continue
        f.precision = prec
        f.set_result(np.zeros(3 * len(c)))
        f.solve_system(rel_tol=1e-8)
        st = f.last_stats
        print("%s: %d iterations, %d refreshes, %.2f ms, status %d" % (name, st.iters, st.refreshes, st.solve_ms,
                                                                        st.status), flush=True)
    f.close()


WORKLOADS = {"traffic": traffic, "assembly": assembly, "matfree": matfree, "iteration": iteration}

if __name__ == "__main__":
    if len(sys.argv) < 2 or sys.argv[1] not in WORKLOADS:
        print(__doc__)
        sys.exit(2)
    n = int(sys.argv[sys.argv.index("--n") + 1]) if "--n" in sys.argv else 100
    if sys.argv[1] == "iteration" and "--precision" in sys.argv:
        iteration(n, sys.argv[sys.argv.index("--precision") + 1])
    else:
        WORKLOADS[sys.argv[1]](n)
