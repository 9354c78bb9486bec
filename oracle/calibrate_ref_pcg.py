"""Times the reference's own compiled block IC(0)-PCG on this host, beside its published timing.

The reference publishes 0.430416 s for chol_incomp + solve_using_incomplete_chol_factorize
(containers/MatWithTinyBlocks.h:426-523, 1101-1151) on a 27,783-DOF Hexa problem
(make_rect(m, Hexa(), 0, 1, 21): 20^3 cells, both x faces held by penalty constraints, a start
vector 10% of the solution's norm away from it, crit 1e-11; include/mesh/exemple_probleme_elastique.txt
:78-95,137-187,216-231) on an Intel Core2 6700 at 2.66 GHz.  This script runs the same two calls,
compiled from the reference's sources (oracle/_ref/ref_tinyblock_pcg, `make -C oracle ref`), on:

  hexa20   a system of that size and structure assembled by the oracle: Hexa8 20^3 cells,
           E 210000, nu 0.3, the x=0 and x=1 faces held by penalty 1e5 (u_x = 0.1 x, u_y = u_z = 0),
           start = solution + 0.1 ||solution|| in every entry; stop at max|r| <= 1e-10 max|b| (the
           published crit is absolute and its problem's scale is not stated);
  config0  the 3,993-DOF Tetra cube (10^3 cells x 6, clamp at x=0, traction (0,0,-1) at x=1,
           penalty 1e5), zero start, max|r| <= 1e-10 max|b|.

It reports the iterations, the seconds of the factorisation and of the PCG (single thread, as the
reference runs), and the relative residual reached.  Run where the reference sources and the
oracle are built:

    python oracle/calibrate_ref_pcg.py
"""
import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle import oracle as O  # noqa: E402

PUBLISHED_S = 0.430416


def _cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def hexa20():
    c, e = O.make_rect(O.HEXA, (21, 21, 21))
    rp, col, vals, F, _ = O.assemble(O.HEXA, c, e, 210000.0, 0.3)
    ends = np.where((c[:, 0] < 1e-6) | (c[:, 0] > 1 - 1e-6))[0]
    dofs = (3 * ends[:, None] + np.arange(3)).ravel()
    val = np.where(dofs % 3 == 0, 0.1 * c[dofs // 3, 0], 0.0)
    O.penalty(rp, col, vals, F, dofs, val, 1e5)
    x, _, _, rc = O.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14, block=True)
    start = x + 0.1 * np.linalg.norm(x)
    return rp, col, vals, F, start


def config0():
    c, e = O.make_rect(O.TETRA, (11, 11, 11))
    faces = O.skin_faces(O.TETRA, len(c), e)
    fx = faces[np.all(c[faces][:, :, 0] > 1 - 1e-6, axis=1)]
    rp, col, vals, F, _ = O.assemble(O.TETRA, c, e, 210000.0, 0.3, faces=fx, face_trac=(0, 0, -1))
    clamp = np.where(c[:, 0] < 1e-6)[0]
    O.penalty(rp, col, vals, F, (3 * clamp[:, None] + np.arange(3)).ravel(), None, 1e5)
    return rp, col, vals, F, np.zeros_like(F)


def run(name, rp, col, vals, F, x0, tmp):
    crit = 1e-10 * np.abs(F).max()
    inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    with open(inp, "wb") as f:
        f.write(np.array([len(rp) - 1, len(col)], dtype=np.uint32).tobytes())
        f.write(np.asarray(rp, np.uint32).tobytes())
        f.write(np.asarray(col, np.uint32).tobytes())
        f.write(np.asarray(vals, np.float64).tobytes())
        f.write(np.asarray(F, np.float64).tobytes())
        f.write(np.asarray(x0, np.float64).tobytes())
        f.write(np.array([crit], np.float64).tobytes())
    line = subprocess.check_output([os.path.join(HERE, "_ref", "ref_tinyblock_pcg"), inp, out]).decode().split()
    iters, t_fac, t_pcg = int(line[1]), float(line[3]), float(line[5])
    with open(out, "rb") as f:
        f.read(4)
        x = np.frombuffer(f.read(8 * len(F)), dtype=np.float64)
    r = F - O.spmv(rp, col, vals, x)
    return {"system": name, "dof": len(F), "iterations": iters, "factor_s": t_fac, "pcg_s": t_pcg,
            "total_s": t_fac + t_pcg, "rel_residual": float(np.linalg.norm(r) / np.linalg.norm(F)),
            "max_abs_residual_over_max_b": float(np.abs(r).max() / np.abs(F).max())}


def main():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so", "_ref/ref_tinyblock_pcg"])
    tmp = tempfile.mkdtemp()
    rows = [run("hexa20", *hexa20(), tmp), run("config0", *config0(), tmp)]
    print(json.dumps({"cpu": _cpu(), "threads": 1, "published_s": PUBLISHED_S,
                      "published_cpu": "Intel Core2 6700 @ 2.66 GHz", "runs": rows}, indent=1))


if __name__ == "__main__":
    main()
