// VTK XML unstructured-grid (.vtu) result files in the layout of LMT++'s binary writer
// write_mesh_vtk<true> (mesh/write_mesh_vtk.h:434-656): point fields, cell fields in name order,
// the element group as ASCII, global fields repeated per cell, points, ASCII cells, and every
// float array base64-encoded with a 4-byte length prefix in one AppendedData block, offsets
// counted in base64 characters.  Host code only.
#include "io_vtu.hpp"

#include <algorithm>
#include <cstdio>
#include <string>

namespace femgpu {

namespace {

const char B64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

// Appends base64 of (uint32 byte count, little endian) + bytes.
void append_b64(const std::string &bytes, std::string &out) {
    const uint32_t n = (uint32_t)bytes.size();
    std::string s(reinterpret_cast<const char *>(&n), 4);
    s += bytes;
    const size_t len = s.size();
    out.reserve(out.size() + (len + 2) / 3 * 4);
    size_t k = 0;
    for (; k + 3 <= len; k += 3) {
        const unsigned a = (unsigned char)s[k], b = (unsigned char)s[k + 1], c = (unsigned char)s[k + 2];
        out += B64[a >> 2];
        out += B64[((a << 4) & 0x30) | (b >> 4)];
        out += B64[((b << 2) & 0x3c) | (c >> 6)];
        out += B64[c & 0x3f];
    }
    if (len - k == 2) {
        const unsigned a = (unsigned char)s[k], b = (unsigned char)s[k + 1];
        out += B64[a >> 2];
        out += B64[((a << 4) & 0x30) | (b >> 4)];
        out += B64[(b << 2) & 0x3c];
        out += '=';
    } else if (len - k == 1) {
        const unsigned a = (unsigned char)s[k];
        out += B64[a >> 2];
        out += B64[(a << 4) & 0x30];
        out += "==";
    }
}

// Components as written: a 2-component field gets a zero third component.
uint32_t written_comp(uint32_t n) { return n + (n == 2); }

// Values of count items; pad: a 2-component item gets a zero third component (point and global
// fields; cell fields keep their component count).
std::string field_bytes(const VtuField &f, size_t count, bool pad) {
    const uint32_t w = pad ? written_comp(f.n_comp) : f.n_comp;
    std::string b;
    b.reserve(count * w * 8);
    const double zero = 0.0;
    for (size_t i = 0; i < count; ++i) {
        b.append(reinterpret_cast<const char *>(f.data + i * f.n_comp), 8 * (size_t)f.n_comp);
        if (w != f.n_comp) b.append(reinterpret_cast<const char *>(&zero), 8);
    }
    return b;
}

void array_header(std::string &x, const std::string &name, uint32_t n_comp, size_t offset) {
    x += "                <DataArray Name='" + name + "' NumberOfComponents='" + std::to_string(n_comp) +
         "' type='Float64' format='appended' offset='" + std::to_string(offset) + "'>\n";
    x += "                </DataArray>\n";
}

void ascii_ints(std::string &x, const std::vector<uint32_t> &v) {
    for (size_t i = 0; i < v.size(); ++i) {
        x += std::to_string(v[i]);
        x += (i % 8) == 7 || i + 1 == v.size() ? '\n' : ' ';
    }
}

}  // namespace

int vtk_cell_type(int elem_kind, int *n_points) {
    switch (elem_kind) {
        case TETRA: *n_points = 4; return 10;
        case TETRA10: *n_points = 10; return 24;
        case HEXA: *n_points = 8; return 12;
        default: *n_points = 0; return 0;
    }
}

bool write_vtu(const std::string &path, int elem_kind, uint32_t n_nodes, const double *coords, uint32_t n_elem,
               const uint32_t *conn, const std::vector<VtuField> &point, std::vector<VtuField> cell,
               const uint32_t *group, const std::vector<VtuField> &global) {
    int npe = 0;
    const int ctype = vtk_cell_type(elem_kind, &npe);
    std::string x, app;
    x += "<?xml version='1.0' ?>\n";
    x += "<VTKFile type='UnstructuredGrid' byte_order='LittleEndian'>\n";
    x += "    <UnstructuredGrid>\n";
    x += "        <Piece NumberOfPoints='" + std::to_string(n_nodes) + "' NumberOfCells='" + std::to_string(n_elem) + "'>\n";
    x += "            <PointData>\n";
    if (n_nodes)
        for (const VtuField &f : point) {
            if (!f.n_comp) continue;
            array_header(x, f.name, written_comp(f.n_comp), app.size());
            append_b64(field_bytes(f, n_nodes, true), app);
        }
    x += "            </PointData> \n";
    x += "            <CellData>\n";
    if (n_elem) {
        std::stable_sort(cell.begin(), cell.end(), [](const VtuField &a, const VtuField &b) { return a.name < b.name; });
        for (const VtuField &f : cell) {
            if (!f.n_comp) continue;
            array_header(x, f.name, f.n_comp, app.size());
            append_b64(field_bytes(f, n_elem, false), app);
        }
        x += "                <DataArray Name='group' NumberOfComponents='1' type='UInt32' format='ascii'>\n";
        for (uint32_t e = 0; e < n_elem; ++e) {
            x += std::to_string(group ? group[e] : 0u);
            x += ' ';
        }
        x += "\n                </DataArray>\n";
    }
    for (const VtuField &f : global) {
        if (!f.n_comp) continue;
        array_header(x, f.name, written_comp(f.n_comp), app.size());
        const std::string one = field_bytes(f, 1, true);
        std::string all;
        all.reserve(one.size() * n_elem);
        for (uint32_t e = 0; e < n_elem; ++e) all += one;
        append_b64(all, app);
    }
    x += "            </CellData> \n";
    x += "            <Points>\n";
    x += "                <DataArray NumberOfComponents='3' type='Float64' format='appended' offset='" +
         std::to_string(app.size()) + "'>\n";
    append_b64(std::string(reinterpret_cast<const char *>(coords), 24 * (size_t)n_nodes), app);
    x += "                </DataArray>\n";
    x += "            </Points> \n";
    std::vector<uint32_t> cv(conn, conn + (size_t)n_elem * npe), off(n_elem), ty(n_elem, (uint32_t)ctype);
    for (uint32_t e = 0; e < n_elem; ++e) off[e] = (uint32_t)((e + 1) * (size_t)npe);
    x += "            <Cells>\n";
    x += "                <DataArray Name='connectivity' type='Int32' format='ascii'>\n";
    ascii_ints(x, cv);
    x += "                </DataArray>\n";
    x += "                <DataArray Name='offsets' type='Int32' format='ascii'>\n";
    ascii_ints(x, off);
    x += "                </DataArray>\n";
    x += "                <DataArray Name='types' type='Int32' format='ascii'>\n";
    ascii_ints(x, ty);
    x += "                </DataArray>\n";
    x += "            </Cells>\n";
    x += "        </Piece>\n";
    x += "    </UnstructuredGrid>\n";
    x += "    <AppendedData encoding='base64'>\n";
    x += "        _";
    FILE *fo = std::fopen(path.c_str(), "wb");
    if (!fo) return false;
    bool ok = std::fwrite(x.data(), 1, x.size(), fo) == x.size();
    ok = ok && std::fwrite(app.data(), 1, app.size(), fo) == app.size();
    const char tail[] = "\n    </AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail, 1, sizeof(tail) - 1, fo) == sizeof(tail) - 1;
    return std::fclose(fo) == 0 && ok;
}

}  // namespace femgpu
