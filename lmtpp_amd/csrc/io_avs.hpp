// AVS UCD result files (ASCII) in the layout of LMT++'s write_avs(mesh, name, display_fields,
// Ascii()) (mesh/write_avs.h:389-530).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace femgpu {

// One named field: n_comp doubles per item (node, cell, or a single item for a global field).
// nest is the shape of the value in the reference's mesh data and fixes its spacing in the file:
// 0 a scalar (n_comp 1), 1 a Vec<T,n_comp>, 2 a Vec<Vec<T,n_comp>,1> (epsilon, sigma).
struct AvsField {
    std::string name;
    uint32_t n_comp = 0;
    uint32_t nest = 1;
    const double *data = nullptr;
};

// Writes the header, the nodes, the cells (hex / tet lines; a Tetra_10 cell, which the reference's
// writer has no type for, as its number and material 0 alone), then the node, cell (in name
// order) and global fields that display selects: a field is written when its name is listed, or
// every field when display[0] is "all".
bool write_avs(const std::string &path, int elem_kind, uint32_t n_nodes, const double *coords, uint32_t n_elem,
               const uint32_t *conn, const std::vector<AvsField> &point, std::vector<AvsField> cell,
               const std::vector<AvsField> &global, const std::vector<std::string> &display);

}  // namespace femgpu
