"""HBM traffic of the block SpMV, from rocprofv3 FETCH_SIZE / WRITE_SIZE counters.

Workload (run it under rocprofv3 --pmc, one counter per run):

    python -m lmtpp_amd.traffic run [--n 100]

launches the plain copy kernel (k_copy, 2^26 doubles: exactly 512 MiB read and 512 MiB written),
then 5 FP64 and 5 FP32 block SpMV launches on Hexa8 n^3.  Then

    python -m lmtpp_amd.traffic parse <fetch_dir> <write_dir> > profiles/spmv_traffic.json

reads the per-dispatch counters (KiB), scales each counter by the copy kernel's known bytes
(factor = known KiB / counted KiB), and reports bytes per SpMV launch next to the algorithmic
bytes.
"""
import csv
import glob
import json
import os
import sys

COPY_DOUBLES = 1 << 26


def run(n):
    import numpy as np

    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, here)
    sys.path.insert(0, os.path.join(here, "tests"))
    import lmtpp_amd as L
    from lmtpp_amd import _capi as A
    from problems import build, lognormal_E

    gbs, ms = A.stream_copy_gbs(COPY_DOUBLES, 5)
    print("copy: %.1f GB/s (%.3f ms per launch)" % (gbs, ms), flush=True)
    c, e = L.make_rect(L.HEXA, (n + 1,) * 3)
    f, *_ = build(L.HEXA, c, e, lognormal_E(len(e), 0.5, 0))
    f.assemble()
    f.set_result(np.zeros(3 * len(c)))
    t64 = f.time_spmv(L.FP64, 5)
    t32 = f.time_spmv(L.MIXED, 5)
    sz = f.sizes()
    print(json.dumps({"nnzb": sz["nnzb"], "n_nodes": sz["n_nodes"], "spmv_fp64_ms": t64, "spmv_fp32_ms": t32}), flush=True)


def _counters(d):
    rows = []
    for p in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(p) as f:
            rows += list(csv.DictReader(f))
    per = {}
    for r in rows:
        name = r.get("Kernel_Name", "")
        key = "copy" if "k_copy" in name else ("spmv_fp64" if "k_spmv<double" in name else
                                                 "spmv_fp32" if "k_spmv<float" in name else None)
        if key is None:
            continue
        per.setdefault(key, []).append(float(r["Counter_Value"]))
    return per


def parse(fetch_dir, write_dir, nnzb, n_nodes):
    fe, wr = _counters(fetch_dir), _counters(write_dir)
    known_kib = 8.0 * COPY_DOUBLES / 1024.0
    f_scale = known_kib / (sum(fe["copy"]) / len(fe["copy"]))
    w_scale = known_kib / (sum(wr["copy"]) / len(wr["copy"]))
    ndof = 3 * n_nodes
    out = {"copy_scale": {"FETCH_SIZE": f_scale, "WRITE_SIZE": w_scale},
           "note": "bytes per launch = mean counter (KiB) x 1024 x scale; scale = known copy KiB / counted KiB"}
    for key, alg in (("spmv_fp64", nnzb * 76 + n_nodes * 4 + ndof * 16), ("spmv_fp32", nnzb * 40 + n_nodes * 4 + ndof * 8)):
        if key not in fe or key not in wr:
            continue
        fb = sum(fe[key]) / len(fe[key]) * 1024.0 * f_scale
        wb = sum(wr[key]) / len(wr[key]) * 1024.0 * w_scale
        out[key] = {"fetch_bytes": fb, "write_bytes": wb, "traffic": fb + wb, "algorithmic_bytes": alg,
                    "traffic_over_algorithmic": (fb + wb) / alg, "dispatches": len(fe[key])}
    return out


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "run":
        n = int(sys.argv[sys.argv.index("--n") + 1]) if "--n" in sys.argv else 100
        run(n)
    elif len(sys.argv) >= 4 and sys.argv[1] == "parse":
        nnzb = int(sys.argv[4]) if len(sys.argv) > 4 else 27270901
        nn = int(sys.argv[5]) if len(sys.argv) > 5 else 1030301
        print(json.dumps(parse(sys.argv[2], sys.argv[3], nnzb, nn), indent=1))
    else:
        print(__doc__)
        sys.exit(2)
