"""Byte and arithmetic model of the product kernels, and the reader of rocprofv3 FETCH_SIZE /
WRITE_SIZE counters that compares them with measured HBM traffic.

The algorithmic bytes and flops of a launch (spmv_bytes, mf_block_bytes, mf_halo_bytes, mf_flops)
(and mf_block_iter_bytes, update_iter_bytes for the two kernels of a CG iteration) are what the
roofline fractions are computed on.  The counter workloads that run under rocprofv3
live outside the package (profile_workloads.py at the repository root); their counters are read
back with

    python -m lmtpp_amd.traffic parse <fetch_dir> <write_dir> [nnzb n_nodes] > profiles/spmv_traffic.json

which scales each per-dispatch counter (KiB) by the plain copy kernel's known bytes (k_copy of
COPY_DOUBLES doubles: exactly 512 MiB read and 512 MiB written; factor = known KiB / counted KiB)
and reports bytes per launch next to the algorithmic bytes.
"""
import csv
import glob
import json
import os
import re
import sys

COPY_DOUBLES = 1 << 26


# Arithmetic of one element in the matrix-free product as written (k_mf_block; an FMA counts 2).
# General element (hexa_apply): lattice differences 72, 18 bilinear interpolations x 20 = 360, per
# Gauss point adjugate and det 32 + w/det 1 + displacement gradient 45 + trace 2 + stress 19 +
# projection onto the adjugate 54 = 153 (x 8 = 1224), 9 transposed projections x 24 = 216, nodal
# forces 48, Hooke law 8.
MF_FLOPS_PER_ELEM = 72 + 360 + 1224 + 216 + 48 + 8
# Axis-aligned box (hexa_apply_box, Hooke coefficients from the element's record): displacement
# differences 36, 9 bilinear interpolations x 20 = 180, diagonal adjugate, w/det and the axis
# scales 8, per axis (box_axis) coefficients 9 + four marginal sums 12 + 12 values 20 = 41
# (x 3 = 123), 9 transposed projections x 24 = 216, nodal forces 48.
MF_BOX_FLOPS_PER_ELEM = 36 + 180 + 8 + 123 + 216 + 48


def mf_flops(n_elem, info):
    """Arithmetic of one element-block launch: box elements on the reduced path, the rest on the
    general one."""
    nb = info.get("box_elements", 0)
    return nb * MF_BOX_FLOPS_PER_ELEM + (n_elem - nb) * MF_FLOPS_PER_ELEM


def mf_block_bytes(n_elem, n_nodes, info, fp32):
    """Algorithmic bytes of one element-block launch: per element its 8 block-local indices and
    sub-colour (17 B) and either its record (a box: edge lengths and Hooke coefficients, 24 B in
    FP32, 48 B in FP64) or its 8 nodes, element id and modulus (44 B) with the coordinates of the
    general elements' nodes once (24 B per node); per block its local-node range; per node x once;
    per block-local node entry its global node and slot (8 B; the lists are stored at a fixed
    stride per block, and the padding is not counted); the stores of the nodes a block owns and of
    the shared nodes' partials."""
    s = 12 if fp32 else 24
    nb = info.get("box_elements", 0)
    owned = n_nodes - info["shared_nodes"] - info["isolated_nodes"]
    coords = n_nodes * 24 if nb < n_elem else 0
    return (nb * (17 + 2 * s) + (n_elem - nb) * 61 + coords + info["blocks"] * 4 + n_nodes * s +
            info["local_entries"] * 8 + (owned + info["partial_slots"]) * s)


def mf_halo_bytes(info, fp32):
    """Algorithmic bytes of the shared-node pass: node and range (8 B) per shared node, its
    partials read once, its y written once."""
    s = 12 if fp32 else 24
    return info["shared_nodes"] * (8 + s) + info["partial_slots"] * s


def spmv_bytes(nnzb, n_nodes, fp32, fmt, col16=0):
    """Algorithmic bytes of one block SpMV: every stored value and index read once, x read once,
    y written once.  Full format: 4-byte columns, or 2-byte column offsets (col16).  Symmetric
    format: upper blocks (incl. diagonal) with values and columns, lower blocks as column +
    position (their values are the upper blocks of other rows)."""
    v = 36 if fp32 else 72
    ndof = 3 * n_nodes
    vec = ndof * (8 if fp32 else 16)
    if fmt == 0:
        return nnzb * (v + (2 if col16 else 4)) + n_nodes * 4 + vec
    nu, nl = (nnzb + n_nodes) // 2, (nnzb - n_nodes) // 2
    return nu * (v + 4) + nl * 8 + n_nodes * 8 + vec


def mf_block_iter_bytes(n_elem, n_nodes, info, fp32):
    """Algorithmic bytes of the CG iteration's element-block launch (k_mf_block with the direction
    update and the end of the previous vector update folded in): mf_block_bytes with the input
    formed from two vectors (z and the stored direction: one more vector read at every node), the
    new direction stored once per node, and the p.q partial per block.  The r.z partials of the
    previous update every block combines come from L2 and are not counted."""
    s = 12 if fp32 else 24
    return (mf_block_bytes(n_elem, n_nodes, info, fp32) + n_nodes * s +
            (n_nodes - info["isolated_nodes"]) * s + info["blocks"] * 8)


def update_iter_bytes(n_nodes, info, fp32, precond=1):
    """Algorithmic bytes of the CG iteration's vector update (k_update_mf): per node x and r read
    and written, the direction read, the preconditioner's values read (3 for Jacobi, 6 for block
    Jacobi; bfloat16 in FP32, FP64 otherwise), z written, its partial range (4 B); q read at the
    nodes of one block and the partials of the shared nodes read instead."""
    s = 12 if fp32 else 24
    m = (3 if precond == 0 else 6) * (2 if fp32 else 8)
    return (n_nodes * (6 * s + m + 4) - info["shared_nodes"] * s + info["partial_slots"] * s)


def _counters(d):
    rows = []
    for p in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(p) as f:
            rows += list(csv.DictReader(f))
    per = {}
    pat = re.compile(r"k_spmv<(double|float), 0, (\d), (\d), (\d)>")
    for r in rows:
        name = r.get("Kernel_Name", "")
        m = pat.search(name)
        mf = re.search(r"k_mf_(block)<(double|float), false, false, (?:true|false)(?:, \d+)?(?:, false)?(?:, (?:true|false))?>",
                       name) or re.search(r"k_mf_(halo)<(double|float)(?:, false)?>", name)
        it = re.search(r"k_mf_block<(double|float), true, true, (?:true|false), \d+, true(?:, (?:true|false))?>", name)
        up = re.search(r"k_update_mf<\d, (double|float), true>", name)
        if "k_copy" in name:
            key = "copy"
        elif it:
            key = "mf_block_iter_%s" % ("fp64" if it.group(1) == "double" else "fp32")
        elif up:
            key = "update_iter_%s" % ("fp64" if up.group(1) == "double" else "fp32")
        elif mf:
            key = "mf_%s_%s" % (mf.group(1), "fp64" if mf.group(2) == "double" else "fp32")
        elif m:
            kind = "_sym" if m.group(3) == "1" else {"0": "_plain", "1": "", "2": "_col16"}[m.group(4)]
            key = "spmv_%s%s" % ("fp64" if m.group(1) == "double" else "fp32", kind)
        else:
            continue
        per.setdefault(key, []).append(float(r["Counter_Value"]))
    return per


def parse(fetch_dir, write_dir, nnzb, n_nodes):
    fe, wr = _counters(fetch_dir), _counters(write_dir)
    known_kib = 8.0 * COPY_DOUBLES / 1024.0
    f_scale = known_kib / (sum(fe["copy"]) / len(fe["copy"]))
    w_scale = known_kib / (sum(wr["copy"]) / len(wr["copy"]))
    out = {"copy_scale": {"FETCH_SIZE": f_scale, "WRITE_SIZE": w_scale},
           "note": "bytes per launch = mean counter (KiB) x 1024 x scale; scale = known copy KiB / counted KiB"}
    for key, fp32, fmt, c16 in (("spmv_fp64", False, 0, 0), ("spmv_fp32", True, 0, 0), ("spmv_fp64_col16", False, 0, 1),
                                ("spmv_fp32_col16", True, 0, 1), ("spmv_fp64_plain", False, 0, 0),
                                ("spmv_fp32_plain", True, 0, 0), ("spmv_fp64_sym", False, 1, 0), ("spmv_fp32_sym", True, 1, 0)):
        alg = spmv_bytes(nnzb, n_nodes, fp32, fmt, c16)
        if key not in fe or key not in wr:
            continue
        fb = sum(fe[key]) / len(fe[key]) * 1024.0 * f_scale
        wb = sum(wr[key]) / len(wr[key]) * 1024.0 * w_scale
        out[key] = {"fetch_bytes": fb, "write_bytes": wb, "traffic": fb + wb, "algorithmic_bytes": alg,
                    "traffic_over_algorithmic": (fb + wb) / alg, "dispatches": len(fe[key])}
    # matrix-free kernels: measured bytes (their algorithmic bytes depend on the block layout and
    # are computed where the layout is known: mf_block_bytes, mf_halo_bytes)
    for key in ("mf_block_fp64", "mf_block_fp32", "mf_halo_fp64", "mf_halo_fp32", "mf_block_iter_fp64",
                "mf_block_iter_fp32", "update_iter_fp64", "update_iter_fp32"):
        if key not in fe or key not in wr:
            continue
        fb = sum(fe[key]) / len(fe[key]) * 1024.0 * f_scale
        wb = sum(wr[key]) / len(wr[key]) * 1024.0 * w_scale
        out[key] = {"fetch_bytes": fb, "write_bytes": wb, "traffic": fb + wb, "dispatches": len(fe[key])}
    return out


if __name__ == "__main__":
    if len(sys.argv) >= 4 and sys.argv[1] == "parse":
        nnzb = int(sys.argv[4]) if len(sys.argv) > 4 else 27270901
        nn = int(sys.argv[5]) if len(sys.argv) > 5 else 1030301
        print(json.dumps(parse(sys.argv[2], sys.argv[3], nnzb, nn), indent=1))
    else:
        print(__doc__)
        sys.exit(2)
