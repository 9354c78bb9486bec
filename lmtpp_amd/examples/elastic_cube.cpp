// Elastic cube through the C++ facade: the counterpart of the reference's example
// (include/mesh/exemple_probleme_elastique.txt:59-213) on the GPU path.
//   elastic_cube <tetra|tetra10|hexa> <points per side> <out.bin> [jacobi|ic0|ramp]
// Clamps x = 0 by penalty (1e5, as the reference example), applies the traction (0,0,-1) on
// x = 1, E = 210000, nu = 0.3, solves with the reference stop test max|r| <= 1e-10 max|F| (Jacobi
// PCG, or the reference's block IC(0) PCG with "ic0"), writes the displacements (3 doubles per
// node) to out.bin and prints the largest von Mises stress of call_after_solve().  "ramp" then
// runs three time steps of ProblemGpu with the traction scaled by the step number, the matrix
// assumed constant (K and its block IC(0) factor built once) and each step warm-started.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "femgpu_formulation.hpp"

int main(int argc, char **argv) {
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: %s <tetra|tetra10|hexa> <points per side> <out.bin> [jacobi|ic0|ramp]\n", argv[0]);
        return 2;
    }
    const femgpu_elem_kind kind = !std::strcmp(argv[1], "hexa") ? FEMGPU_HEXA
                                  : !std::strcmp(argv[1], "tetra10") ? FEMGPU_TETRA10 : FEMGPU_TETRA;
    const uint32_t n = (uint32_t)std::atoi(argv[2]);
    const uint32_t nb[3] = {n, n, n};
    const double X0[3] = {0, 0, 0}, X1[3] = {1, 1, 1};
    uint32_t nn = 0, ne = 0;
    femgpu::check(femgpu_make_rect(kind, X0, X1, nb, nullptr, nullptr, &nn, &ne), "make_rect");
    const int npe = kind == FEMGPU_TETRA ? 4 : kind == FEMGPU_TETRA10 ? 10 : 8;
    std::vector<double> coords(3 * (size_t)nn);
    std::vector<uint32_t> conn((size_t)ne * npe);
    femgpu::check(femgpu_make_rect(kind, X0, X1, nb, coords.data(), conn.data(), &nn, &ne), "make_rect");
    uint32_t nf = 0;
    femgpu::check(femgpu_skin_faces(kind, nn, ne, conn.data(), nullptr, &nf), "skin_faces");
    const int npf = kind == FEMGPU_TETRA ? 3 : kind == FEMGPU_TETRA10 ? 6 : 4;
    std::vector<uint32_t> skin((size_t)nf * npf);
    femgpu::check(femgpu_skin_faces(kind, nn, ne, conn.data(), skin.data(), &nf), "skin_faces");

    femgpu::FormulationGpu f(kind, coords, conn);
    const bool ramp = argc == 5 && !std::strcmp(argv[4], "ramp");
    if (argc == 5 && (ramp || !std::strcmp(argv[4], "ic0"))) f.options().precond = FEMGPU_BLOCK_IC0;
    f.set_material(210000.0, 0.3);
    std::vector<uint32_t> faces;
    std::vector<double> trac;
    for (uint32_t s = 0; s < nf; ++s) {
        bool on = true;
        for (int a = 0; a < npf; ++a) on &= std::fabs(coords[3 * (size_t)skin[s * npf + a]] - 1.0) < 1e-6;
        if (!on) continue;
        faces.insert(faces.end(), skin.begin() + (size_t)s * npf, skin.begin() + (size_t)(s + 1) * npf);
        trac.insert(trac.end(), {0.0, 0.0, -1.0});
    }
    f.set_traction(faces, trac);
    for (uint32_t i = 0; i < nn; ++i)
        if (std::fabs(coords[3 * (size_t)i]) < 1e-6)
            for (uint32_t d = 0; d < 3; ++d) f.add_constraint(3 * i + d, 0.0, 1e5);
    f.assemble();
    const std::vector<double> F = f.get_sollicitation();
    double fmax = 0;
    for (double v : F) fmax = std::fmax(fmax, std::fabs(v));
    if (!f.solve(1e-10 * fmax)) {
        std::fprintf(stderr, "solve failed: %s\n", femgpu_status_string(f.last_status()));
        return 1;
    }
    const std::vector<double> &u = f.update_variables();
    FILE *o = std::fopen(argv[3], "wb");
    if (!o) return 1;
    std::fwrite(u.data(), 8, u.size(), o);
    std::fclose(o);
    const femgpu::FormulationGpu::AfterSolve as = f.call_after_solve();
    double vm = 0;
    for (uint32_t e = 0; e < ne; ++e) {
        const double *s = &as.sigma[6 * (size_t)e];
        const double j2 = ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) + (s[2] - s[0]) * (s[2] - s[0])) / 6 +
                          s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
        vm = std::fmax(vm, std::sqrt(3 * j2));
    }
    std::printf("%u nodes %u elems: %d iterations, max|r| %.3e, tip uz %.9e, max von Mises %.6e\n", nn, ne,
                f.last_stats().iters, f.last_stats().max_abs_res, u[3 * (size_t)(nn - 1) + 2], vm);
    if (!ramp) return 0;
    f.set_assume_constant_matrix(true);
    f.set_initial_time_step(1.0);
    f.default_iterative_criterium = 1e-12 * fmax;
    femgpu::ProblemGpu problem({&f});
    for (int k = 1; k <= 3; ++k) {
        std::vector<double> tk(trac);
        for (double &v : tk) v *= k;
        f.set_traction(faces, tk);
        problem.iteration();
        const std::vector<double> &uk = f.update_variables();
        std::printf("step %d time %g: %d iterations, tip uz %.9e\n", k, problem.time(), f.last_stats().iters,
                    uk[3 * (size_t)(nn - 1) + 2]);
    }
    return 0;
}
