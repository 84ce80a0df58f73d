// Text files of the reference pipeline, written and read with the same
// layout and the default std::ostream formatting (6 significant digits) the
// reference uses, so outputs can be exchanged with it file for file:
//   uvw / updated_uvw  "u v w[ count]" per line   MIncline_generate_uvw_float.cu:420-427,
//                                                 read at shade10M.cu:155-171 and
//                                                 vissGen10M_shade.cu:351-366
//   shadeM             "tS tE pS pE" per line     shade10M.cu:108-127, read at
//                                                 vissGen10M_shade.cu:376-411
//   F (image)          one real value per line    vissGen10M_shade.cu:563-583
//   Ex (fine cells)    F of each cell's parent    vissGen1M_shade_accurate.cu:567-588
//   uvwMap             "n" then "u v w count"     read at vissGen1M_shade_accurate.cu:248-289
// plus a binary complex64 visibility file (the reference never writes V).
// The readers never write past the caller's capacity (the reference's fixed
// uvw_presize buffers have no bound check, vissGen10M_shade.cu:56,359).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>

#include "io.hpp"

namespace vrs {

int write_uvw_txt(const char *path, const double *uvw, int64_t n, const float *col4) {
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    for (int64_t i = 0; i < n; ++i) {
        f << (float)uvw[3 * i] << " " << (float)uvw[3 * i + 1] << " " << (float)uvw[3 * i + 2];
        if (col4) f << " " << col4[i];
        f << std::endl;
    }
    return f.good() ? 0 : -1;
}

int read_uvw_txt(const char *path, int n_cols, double *uvw, float *col4, int64_t cap,
                 int64_t *n_read) {
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    int64_t n = 0;
    float u, v, w, c = 0.0f;
    while (f >> u >> v >> w) {
        if (n_cols == 4 && !(f >> c)) return -3;  // truncated line
        if (n >= cap) return -2;
        uvw[3 * n] = u;
        uvw[3 * n + 1] = v;
        uvw[3 * n + 2] = w;
        if (col4) col4[n] = c;
        ++n;
    }
    *n_read = n;
    return 0;
}

int write_shade_txt(const char *path, const int32_t *box4, int64_t n) {
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    for (int64_t i = 0; i < n; ++i)
        f << box4[4 * i] << " " << box4[4 * i + 1] << " " << box4[4 * i + 2] << " "
          << box4[4 * i + 3] << std::endl;
    return f.good() ? 0 : -1;
}

// As the reader of vissGen10M_shade.cu:385-399: each pair is ordered S <= E and
// made 0-based (the files are 1-based, written for MATLAB).
int read_shade_txt(const char *path, int32_t *box4, int64_t cap, int64_t *n_read) {
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    int64_t n = 0;
    int a, b, c, d;
    while (f >> a >> b >> c >> d) {
        if (n >= cap) return -2;
        box4[4 * n] = (a > b ? b : a) - 1;
        box4[4 * n + 1] = (a > b ? a : b) - 1;
        box4[4 * n + 2] = (c > d ? d : c) - 1;
        box4[4 * n + 3] = (c > d ? c : d) - 1;
        ++n;
    }
    *n_read = n;
    return 0;
}

int write_image_txt(const char *path, const float *F, int64_t n) {
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    for (int64_t i = 0; i < n; ++i) f << F[i] << std::endl;
    return f.good() ? 0 : -1;
}

int read_image_txt(const char *path, float *F, int64_t cap, int64_t *n_read) {
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    int64_t n = 0;
    float x;
    while (f >> x) {
        if (n >= cap) return -2;
        F[n++] = x;
    }
    *n_read = n;
    return 0;
}

int write_ex_txt(const char *path, const float *F, int64_t n_pix, const int32_t *nx,
                 int64_t n_cells, int32_t nx_base) {
    for (int64_t c = 0; c < n_cells; ++c) {
        const int64_t q = (int64_t)nx[c] - nx_base;
        if (q < 0 || q >= n_pix) return -2;
    }
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    for (int64_t c = 0; c < n_cells; ++c) f << F[(int64_t)nx[c] - nx_base] << std::endl;
    return f.good() ? 0 : -1;
}

// Binary visibilities: "VRSV", u32 version 1, i64 n_maps, i64 n_vis, then
// n_maps * n_vis complex64 (re, im) little-endian.
namespace {
constexpr char kVisMagic[4] = {'V', 'R', 'S', 'V'};
}

int write_vis_bin(const char *path, const float *V, int64_t n_maps, int64_t n_vis) {
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) return -1;
    const uint32_t ver = 1;
    f.write(kVisMagic, 4);
    f.write(reinterpret_cast<const char *>(&ver), 4);
    f.write(reinterpret_cast<const char *>(&n_maps), 8);
    f.write(reinterpret_cast<const char *>(&n_vis), 8);
    f.write(reinterpret_cast<const char *>(V), (std::streamsize)(sizeof(float) * 2 * n_maps * n_vis));
    return f.good() ? 0 : -1;
}

int read_vis_bin(const char *path, float *V, int64_t cap, int64_t *n_maps, int64_t *n_vis) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return -1;
    char magic[4];
    uint32_t ver = 0;
    int64_t nm = 0, nv = 0;
    f.read(magic, 4);
    f.read(reinterpret_cast<char *>(&ver), 4);
    f.read(reinterpret_cast<char *>(&nm), 8);
    f.read(reinterpret_cast<char *>(&nv), 8);
    if (!f.good() || std::memcmp(magic, kVisMagic, 4) != 0 || ver != 1 || nm < 0 || nv < 0)
        return -3;
    *n_maps = nm;
    *n_vis = nv;
    if (!V) return 0;                     // header only
    if (nm * nv > cap) return -2;
    f.read(reinterpret_cast<char *>(V), (std::streamsize)(sizeof(float) * 2 * nm * nv));
    return f.good() ? 0 : -3;
}

// ------------------------------------------------------------------ uvwMap
// The 1 MHz pipeline's cell-count table (vissGen1M_shade_accurate.cu:248-289):
// a first line holding the number of cells, then one "u v w count" line of
// integers per cell.  A line that does not start with four integers is
// skipped; a later line for a cell replaces an earlier one.  Row i's weight is
// 1 / count of its cell, or 1 where the table has no such cell (:352-368).  A
// row's cell is its u v w rounded to FP32 (the reference reads the uvw file
// as float) and truncated toward zero (static_cast<int>).
namespace {
struct Cell {
    int32_t u, v, w;
    bool operator==(const Cell &o) const { return u == o.u && v == o.v && w == o.w; }
};
struct CellHash {
    size_t operator()(const Cell &c) const {
        uint64_t h = (uint64_t)(uint32_t)c.u * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)c.v + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)c.w * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        return (size_t)(h ^ (h >> 31));
    }
};

// static_cast<int>(float) is defined only inside the int range
bool row_cell(const double *r, Cell *c) {
    int32_t out[3];
    for (int k = 0; k < 3; ++k) {
        const float x = (float)r[k];
        if (!std::isfinite(x) || x <= -2147483904.0f || x >= 2147483648.0f) return false;
        out[k] = (int32_t)x;
    }
    *c = Cell{out[0], out[1], out[2]};
    return true;
}
}  // namespace

int uvw_map_weights(const char *path, const double *uvw, int64_t n, float *w, int64_t *n_cells,
                    int64_t *n_skipped, int64_t *n_found) {
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    std::string line;
    if (!std::getline(f, line)) return -3;
    char *end = nullptr;
    const long long declared = std::strtoll(line.c_str(), &end, 10);
    if (end == line.c_str() || declared < 0) return -3;      // no leading count
    std::unordered_map<Cell, float, CellHash> table;
    table.reserve((size_t)std::min<long long>(declared, 1ll << 28));
    int64_t skipped = 0;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        int u, v, ww, count;
        if (in >> u >> v >> ww >> count) {
            if (count <= 0) return -5;
            table[Cell{u, v, ww}] = (float)count;
        } else {
            ++skipped;
        }
    }
    int64_t found = 0;
    for (int64_t i = 0; i < n; ++i) {
        Cell c;
        if (!row_cell(uvw + 3 * i, &c)) return -4;
        const auto it = table.find(c);
        if (it != table.end()) {
            w[i] = 1.0f / it->second;
            ++found;
        } else {
            w[i] = 1.0f;
        }
    }
    if (n_cells) *n_cells = (int64_t)table.size();
    if (n_skipped) *n_skipped = skipped;
    if (n_found) *n_found = found;
    return 0;
}

int write_uvw_map_txt(const char *path, const double *uvw, const int32_t *counts, int64_t n) {
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    f << n << "\n";
    for (int64_t i = 0; i < n; ++i) {
        Cell c;
        if (!row_cell(uvw + 3 * i, &c) || counts[i] <= 0) return -4;
        f << c.u << " " << c.v << " " << c.w << " " << counts[i] << "\n";
    }
    return f.good() ? 0 : -1;
}

}  // namespace vrs
