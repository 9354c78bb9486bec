// Batched uneven problems: n independent elasticity problems assembled and solved by one call.
// Each problem keeps its own pattern, scalars and stop test.
#include <vector>

#include "../../include/femgpu.h"
#include "femgpu_internal.hpp"

struct femgpu_batch {
    std::vector<femgpu_problem *> probs;
};

extern "C" {

int femgpu_batch_create(uint32_t n, const femgpu_mesh *meshes, femgpu_batch **out) {
    if (!out || !meshes || !n) return FEMGPU_BAD_INPUT;
    *out = nullptr;
    auto *b = new femgpu_batch();
    for (uint32_t i = 0; i < n; ++i) {
        femgpu_problem *p = nullptr;
        const int rc = femgpu_create(&meshes[i], &p);
        if (rc) {
            for (auto *q : b->probs) femgpu_destroy(q);
            delete b;
            return rc;
        }
        b->probs.push_back(p);
    }
    *out = b;
    return FEMGPU_OK;
}

int femgpu_batch_set_material(femgpu_batch *b, uint32_t i, const femgpu_material *mat) {
    if (!b || i >= b->probs.size()) return FEMGPU_BAD_INPUT;
    return femgpu_set_material(b->probs[i], mat);
}

int femgpu_batch_set_traction(femgpu_batch *b, uint32_t i, uint32_t n_faces, const uint32_t *faces, const double *trac) {
    if (!b || i >= b->probs.size()) return FEMGPU_BAD_INPUT;
    return femgpu_set_traction(b->probs[i], n_faces, faces, trac);
}

int femgpu_batch_add_constraints(femgpu_batch *b, uint32_t i, uint32_t n, const uint32_t *dofs, const double *values,
                                 double penalty) {
    if (!b || i >= b->probs.size()) return FEMGPU_BAD_INPUT;
    return femgpu_add_constraints(b->probs[i], n, dofs, values, penalty);
}

int femgpu_batch_assemble(femgpu_batch *b) {
    if (!b) return FEMGPU_BAD_INPUT;
    for (auto *p : b->probs) {
        const int rc = femgpu_assemble(p, 1, 1);
        if (rc) return rc;
    }
    return FEMGPU_OK;
}

int femgpu_batch_solve(femgpu_batch *b, const femgpu_solve_opts *opts, femgpu_solve_stats *per, double *total_ms) {
    if (!b) return FEMGPU_BAD_INPUT;
    double tot = 0;
    int worst = FEMGPU_OK;
    for (size_t i = 0; i < b->probs.size(); ++i) {
        femgpu_solve_stats st{};
        const int rc = femgpu_solve_system(b->probs[i], opts, &st);
        if (per) per[i] = st;
        tot += st.solve_ms;
        if (rc && !worst) worst = rc;
    }
    if (total_ms) *total_ms = tot;
    return worst;
}

int femgpu_batch_get_result(femgpu_batch *b, uint32_t i, double *u, size_t n) {
    if (!b || i >= b->probs.size()) return FEMGPU_BAD_INPUT;
    return femgpu_get_result(b->probs[i], u, n);
}

int femgpu_batch_destroy(femgpu_batch *b) {
    if (!b) return FEMGPU_OK;
    for (auto *p : b->probs) femgpu_destroy(p);
    delete b;
    return FEMGPU_OK;
}

}  // extern "C"
