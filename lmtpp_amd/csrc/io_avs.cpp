// AVS UCD ASCII writer (see io_avs.hpp).  Numbers are written as an ostream with default flags
// writes them (printf's %g at precision 6 for doubles), so the bytes equal the reference writer's.
#include "io_avs.hpp"

#include <algorithm>
#include <cstdio>

#include "host_mesh.hpp"

namespace femgpu {
namespace {

class Out {
  public:
    void s(const char *t) { buf_ += t; }
    void s(const std::string &t) { buf_ += t; }
    void u(uint64_t v) {
        char b[24];
        buf_.append(b, (size_t)std::snprintf(b, sizeof b, "%llu", (unsigned long long)v));
    }
    void g(double v) {
        char b[40];
        buf_.append(b, (size_t)std::snprintf(b, sizeof b, "%g", v));
    }
    // A field value as the reference's operator<< prints it: a scalar bare, a Vec with a space
    // after each component, a Vec of one Vec with one more space after the inner one.
    void value(const AvsField &f, size_t item) {
        const double *v = f.data + (size_t)f.n_comp * item;
        if (f.nest == 0) {
            g(v[0]);
            return;
        }
        for (uint32_t c = 0; c < f.n_comp; ++c) {
            g(v[c]);
            s(" ");
        }
        if (f.nest == 2) s(" ");
    }
    // "count c0 c1 ... " then one "name,1" line per field
    void block_header(const std::vector<const AvsField *> &fs, bool pad2) {
        u(fs.size());
        s(" ");
        for (const AvsField *f : fs) {
            u(f->n_comp + (pad2 && f->n_comp == 2));
            s(" ");
        }
        s("\n");
        for (const AvsField *f : fs) {
            s(f->name);
            s(",1\n");
        }
    }
    bool flush(FILE *fp) {
        const bool ok = std::fwrite(buf_.data(), 1, buf_.size(), fp) == buf_.size();
        buf_.clear();
        return ok;
    }
    size_t size() const { return buf_.size(); }

  private:
    std::string buf_;
};

std::vector<const AvsField *> select(const std::vector<AvsField> &fs, const std::vector<std::string> &display) {
    const bool all = !display.empty() && display[0] == "all";
    std::vector<const AvsField *> out;
    for (const AvsField &f : fs)
        if (all || std::find(display.begin(), display.end(), f.name) != display.end()) out.push_back(&f);
    return out;
}

uint64_t comps(const std::vector<const AvsField *> &fs, bool pad2) {
    uint64_t n = 0;
    for (const AvsField *f : fs) n += f->n_comp + (pad2 && f->n_comp == 2);
    return n;
}

}  // namespace

bool write_avs(const std::string &path, int elem_kind, uint32_t n_nodes, const double *coords, uint32_t n_elem,
               const uint32_t *conn, const std::vector<AvsField> &point, std::vector<AvsField> cell,
               const std::vector<AvsField> &global, const std::vector<std::string> &display) {
    const uint32_t npe = nodes_per_elem(elem_kind);
    const char *type = elem_kind == HEXA ? "hex " : elem_kind == TETRA ? "tet " : nullptr;
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    // node fields in the given order (counted with 2 components padded to 3), cell fields in name
    // order, global fields in the given order (padded as the node fields)
    std::stable_sort(cell.begin(), cell.end(), [](const AvsField &a, const AvsField &b) { return a.name < b.name; });
    const std::vector<const AvsField *> pf = n_nodes ? select(point, display) : std::vector<const AvsField *>();
    const std::vector<const AvsField *> cf = n_elem ? select(cell, display) : std::vector<const AvsField *>();
    const std::vector<const AvsField *> gf = select(global, display);
    Out o;
    bool ok = true;
    auto chunk = [&]() {
        if (o.size() > (1u << 20)) ok = o.flush(fp) && ok;
    };
    o.u(n_nodes);
    o.s("  ");
    o.u(n_elem);
    o.s(" ");
    o.u(comps(pf, true));
    o.s(" ");
    o.u(comps(cf, false));
    o.s(" ");
    o.u(comps(gf, true));
    o.s("\n");
    for (uint32_t i = 0; i < n_nodes; ++i) {
        o.u(i + 1ull);
        o.s(" ");
        for (int d = 0; d < 3; ++d) {
            o.g(coords[3 * (size_t)i + d]);
            o.s(" ");
        }
        o.s("\n");
        chunk();
    }
    for (uint32_t k = 0; k < n_elem; ++k) {
        o.u(k + 1ull);
        o.s(" 0 ");
        if (type) {
            o.s(type);
            for (uint32_t j = 0; j < npe; ++j) {
                o.u(conn[(size_t)k * npe + j] + 1ull);
                o.s(" ");
            }
            o.s("\n");
        }
        chunk();
    }
    if (!pf.empty()) {
        o.block_header(pf, true);
        for (uint32_t i = 0; i < n_nodes; ++i) {
            o.u(i + 1ull);
            o.s(" ");
            for (const AvsField *f : pf) {
                o.value(*f, i);
                o.s(" ");
            }
            o.s("\n");
            chunk();
        }
    }
    if (!cf.empty()) {
        o.block_header(cf, false);
        for (uint32_t k = 0; k < n_elem; ++k) {
            o.u(k + 1ull);
            o.s(" ");
            for (const AvsField *f : cf) {
                o.value(*f, k);
                o.s(" ");
            }
            o.s("\n");
            chunk();
        }
    }
    if (!gf.empty()) {
        o.block_header(gf, true);
        for (const AvsField *f : gf) {
            o.value(*f, 0);
            if (f->n_comp == 2) o.s(" 0");
            o.s(" \n");
        }
    }
    ok = o.flush(fp) && ok;
    return std::fclose(fp) == 0 && ok;
}

}  // namespace femgpu
