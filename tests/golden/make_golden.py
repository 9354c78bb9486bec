"""Writes the golden fixtures of tests/golden/ from the reference builds in oracle/_ref.

oracle/_ref is compiled from the LMT++ sources (oracle/Makefile, target `ref`); this script is
run where those sources exist.  Fixtures are data only (inputs and the reference's outputs):
  ref_mesh_<kind>_<nx>_<ny>_<nz>.npz  coords, conn   -- make_rect (mesh/make_rect.h)
  ref_tinyblock_pcg_tetra11.npz        K, F, crit, x, iters, factor
        -- MatWithTinyBlocks chol_incomp + solve_using_incomplete_chol_factorize
           (containers/MatWithTinyBlocks.h:426-523,1101-1151) on the config-0 system
  ref_cg_kat.json                       known-answer systems of
        tests/include/containers/conjugate_gradient_test.cpp:15-71
  ref_after_solve_hexa_4_3_5.npz       coords, conn, u, E, nu -> eps, sigma, tr_epsilon
        -- the generated Hexa after_solve of std_formulations/
           definition_probleme_elasticity_dep_static.h:721-789 on a jittered make_rect mesh
  ref_write_vtu_hexa_3_4_2.npz         coords, conn, u, E, nu, eps, sigma, tr -> vtu (file bytes)
        -- write_mesh_vtk(filename, m) (mesh/write_mesh_vtk.h:434-656,827-831) on that mesh
  ref_write_avs.npz                    per case <name>_{coords,conn,u,E,nu,eps,sigma,tr,display,avs}:
        write_avs(mesh, name, display_fields, Ascii()) (mesh/write_avs.h:389-530) on that Hexa
        mesh (displays all / dep,sigma,poisson_ratio,tr_epsilon / f_vol,epsilon / none) and on
        field-less Tetra and Tetra_10 meshes (all)
  ref_symamd.npz                       <mesh>: symamd(m) (util/symamd.h:27-36 with the vendored AMD 1.1)
        node positions of make_rect meshes; graph<k>_{row_ptr,col,pos}: symamd(ind)
        (util/symamd.cpp:16-36) on random symmetric graphs, some with dense rows and empty ones
  ref_symrcm.npz                       tetra_7_6_5, hexa_7_6_5: symrcm(m) (util/symrcm.cpp:38-101)
        node orders of make_rect meshes

    python tests/golden/make_golden.py [after_solve|vtu|avs|symrcm|symamd]   (no argument: every fixture)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

MESHES = [("tetra", O.TETRA, (4, 4, 4)), ("tetra", O.TETRA, (4, 3, 5)), ("hexa", O.HEXA, (3, 4, 5)),
          ("tetra10", O.TETRA10, (3, 3, 3)), ("tetra10", O.TETRA10, (4, 3, 2))]


def config0_system():
    c, e = O.make_rect(O.TETRA, (11, 11, 11))
    faces = O.skin_faces(O.TETRA, len(c), e)
    fx = faces[np.all(c[faces][:, :, 0] > 1 - 1e-6, axis=1)]
    rp, col, vals, F, _ = O.assemble(O.TETRA, c, e, 210000.0, 0.3, faces=fx, face_trac=(0, 0, -1))
    clamp = np.where(c[:, 0] < 1e-6)[0]
    dofs = (3 * clamp[:, None] + np.arange(3)).ravel()
    O.penalty(rp, col, vals, F, dofs, None, 1e5)
    return rp, col, vals, F


def after_solve_fixture(tmp):
    nb = (4, 3, 5)
    c, e = O.ref_mesh("hexa", nb, tmp)
    rng = np.random.default_rng(2024)
    inside = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
    c = c.copy()
    c[inside] += rng.uniform(-0.06, 0.06, size=(int(inside.sum()), 3))
    u = 1e-3 * rng.standard_normal(3 * len(c))
    E, nu = 273000.0, 0.27
    eps, sig, tr = O.ref_after_solve_hexa(nb, c, u, E, nu, tmp)
    np.savez_compressed(os.path.join(HERE, "ref_after_solve_hexa_4_3_5.npz"), coords=c, conn=e, u=u, E=E, nu=nu,
                        eps=eps, sigma=sig, tr=tr)


def vtu_fixture(tmp):
    nb = (3, 4, 2)
    c, e = O.ref_mesh("hexa", nb, tmp)
    rng = np.random.default_rng(77)
    inside = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
    c = c.copy()
    c[inside] += rng.uniform(-0.05, 0.05, size=(int(inside.sum()), 3))
    u = 1e-3 * rng.standard_normal(3 * len(c))
    E, nu = 190000.0, 0.31
    data, eps, sig, tr = O.ref_write_vtu_hexa(nb, c, u, E, nu, tmp)
    np.savez_compressed(os.path.join(HERE, "ref_write_vtu_hexa_3_4_2.npz"), coords=c, conn=e, u=u, E=E, nu=nu,
                        eps=eps, sigma=sig, tr=tr, vtu=np.frombuffer(data, dtype=np.uint8))


def avs_fixture(tmp):
    out = {}
    rng = np.random.default_rng(91)
    cases = [("hexa_all", "hexa", (4, 3, 3), ["all"]),
             ("hexa_some", "hexa", (4, 3, 3), ["dep", "sigma", "poisson_ratio", "tr_epsilon"]),
             ("hexa_vol_eps", "hexa", (4, 3, 3), ["f_vol", "epsilon"]),
             ("hexa_none", "hexa", (4, 3, 3), ["none"]),
             ("tetra_all", "tetra", (3, 3, 3), ["all"]),
             ("tetra10_all", "tetra10", (3, 3, 3), ["all"])]
    for name, kind, nb, display in cases:
        c, e = O.ref_mesh(kind, nb, tmp)
        inside = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
        c = c.copy()
        c[inside] += rng.uniform(-0.05, 0.05, size=(int(inside.sum()), 3))
        u = 1e-3 * rng.standard_normal(3 * len(c))
        E, nu = 201000.0, 0.29
        data, eps, sig, tr = O.ref_write_avs(kind, nb, c, u, E, nu, display, tmp)
        out.update({name + "_coords": c, name + "_conn": e, name + "_u": u, name + "_E": E, name + "_nu": nu,
                    name + "_display": np.array(display), name + "_avs": np.frombuffer(data, dtype=np.uint8)})
        if eps is not None:
            out.update({name + "_eps": eps, name + "_sigma": sig, name + "_tr": tr})
    np.savez_compressed(os.path.join(HERE, "ref_write_avs.npz"), **out)


def symamd_fixture(tmp):
    out = {}
    for name, nb in (("tetra", (7, 6, 5)), ("hexa", (7, 6, 5)), ("tetra10", (4, 3, 3)), ("hexa", (13, 12, 11))):
        out["%s_%d_%d_%d" % ((name,) + nb)] = O.ref_symamd(name, nb, tmp)
    rng = np.random.default_rng(5)
    for k, n in enumerate((40, 150, 300, 500)):
        M = rng.random((n, n)) < 6.0 / n
        if k >= 2:   # rows denser than max(16, 10 sqrt(n)), and rows with no neighbour
            for h in rng.integers(0, n, 3):
                M[h, :] = rng.random(n) < 0.9
            for z in rng.integers(0, n, 4):
                M[z, :] = False
                M[:, z] = False
        M = M | M.T
        np.fill_diagonal(M, True)
        rp = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.uint32)
        col = np.concatenate([np.nonzero(row)[0] for row in M]).astype(np.uint32)
        out.update({"graph%d_row_ptr" % k: rp, "graph%d_col" % k: col,
                    "graph%d_pos" % k: O.ref_symamd_graph(rp, col, tmp)})
    np.savez_compressed(os.path.join(HERE, "ref_symamd.npz"), **out)


def symrcm_fixture(tmp):
    out = {}
    for name, nb in (("tetra", (7, 6, 5)), ("hexa", (7, 6, 5))):
        out["%s_%d_%d_%d" % ((name,) + nb)] = O.ref_symrcm(name, nb, tmp)
    np.savez_compressed(os.path.join(HERE, "ref_symrcm.npz"), **out)


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "liboracle.so"])
    tmp = tempfile.mkdtemp()
    if sys.argv[1:] == ["symrcm"]:
        symrcm_fixture(tmp)
        return
    if sys.argv[1:] == ["after_solve"]:
        after_solve_fixture(tmp)
        return
    if sys.argv[1:] == ["vtu"]:
        vtu_fixture(tmp)
        return
    if sys.argv[1:] == ["symamd"]:
        symamd_fixture(tmp)
        return
    if sys.argv[1:] == ["avs"]:
        avs_fixture(tmp)
        return
    after_solve_fixture(tmp)
    vtu_fixture(tmp)
    avs_fixture(tmp)
    symrcm_fixture(tmp)
    symamd_fixture(tmp)
    for name, kind, nb in MESHES:
        c, e = O.ref_mesh(name, nb, tmp)
        np.savez_compressed(os.path.join(HERE, "ref_mesh_%s_%d_%d_%d.npz" % ((name,) + nb)), coords=c, conn=e)
    rp, col, vals, F = config0_system()
    crit = 1e-10 * np.abs(F).max()
    x, iters, L, present = O.ref_tinyblock_pcg(rp, col, vals, F, np.zeros_like(F), crit, tmp)
    np.savez_compressed(os.path.join(HERE, "ref_tinyblock_pcg_tetra11.npz"), row_ptr=rp, col=col, vals=vals, F=F,
                        crit=crit, x=x, iters=iters, factor=L, present=present)
    kat = {
        "source": "tests/include/containers/conjugate_gradient_test.cpp:15-71",
        "sym3": {"lower": [[1], [2, 5], [-3, -7, 13]], "b": [3, 4, 5], "x": [11, 2, 4], "rel_tol": 1e-12},
        "dense10": {
            "rows": [[348, 295, 221, 165, 221, 311, 311, 279, 216, 200],
                     [295, 333, 113, 152, 174, 271, 244, 250, 169, 179],
                     [221, 113, 277, 125, 166, 215, 272, 159, 207, 134],
                     [165, 152, 125, 152, 155, 135, 206, 107, 143, 143],
                     [221, 174, 166, 155, 270, 250, 280, 170, 219, 151],
                     [311, 271, 215, 135, 250, 412, 389, 287, 247, 167],
                     [311, 244, 272, 206, 280, 389, 474, 259, 259, 219],
                     [279, 250, 159, 107, 170, 287, 259, 274, 140, 159],
                     [216, 169, 207, 143, 219, 247, 259, 140, 297, 141],
                     [200, 179, 134, 143, 151, 167, 219, 159, 141, 165]],
            "b": [0, 8, 1, 4, 6, 6, 0, 9, 0, 6],
            "x": [-1.4176810081375044408, -3.0067729851774937789, 3.5389214608316368121, 7.5278767154067853482,
                  -.95330870992906846260, 16.004952790217101221, -11.900092892559983095, -7.5660154416662888416,
                  -6.0846051557343736576, 8.5770626957989029109],
            "rel_tol": 1e-12},
    }
    with open(os.path.join(HERE, "ref_cg_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print("golden fixtures written; reference PCG iterations:", iters)


if __name__ == "__main__":
    main()
