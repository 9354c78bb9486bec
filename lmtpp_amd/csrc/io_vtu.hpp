// VTU result files in the layout of LMT++'s write_mesh_vtk<true> (mesh/write_mesh_vtk.h:434-656).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace femgpu {

// One named field: n_comp doubles per item (node, cell, or a single item for a global field).
struct VtuField {
    std::string name;
    uint32_t n_comp = 0;
    const double *data = nullptr;
};

// VTK cell type of an element kind (Tetra 10, Tetra_10 24, Hexa 12) and its node count.
int vtk_cell_type(int elem_kind, int *n_points);

// Writes point fields in the given order, cell fields in name order, the element group (null:
// 0), global fields (one item each, repeated for every cell), the points and the cells.
bool write_vtu(const std::string &path, int elem_kind, uint32_t n_nodes, const double *coords, uint32_t n_elem,
               const uint32_t *conn, const std::vector<VtuField> &point, std::vector<VtuField> cell,
               const uint32_t *group, const std::vector<VtuField> &global);

}  // namespace femgpu
