// host_formats.cpp - COO / compressed-matrix containers, Matrix Market I/O
// and seeded synthetic generators for libspgemm_amd.so.
//
// Container semantics follow the reference's lib/coo.h + src/coo.cpp and
// lib/cm.h + src/cm.cpp; every allocation of coo.data / cm arrays uses
// malloc so that freeCOO / cmFree (which call free) match (defect D13 of
// the reference: new[] released with free()).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spgemm.h"
#include "spg_rng.h"

// ----------------------------------------------------------------- COO
// src/coo.cpp:6-13
extern "C" void initCOO(coo* mat) {
    mat->data = nullptr;
    mat->nnz = 0;
    mat->num_cols = 0;
    mat->num_rows = 0;
    mat->type = ROW_MAJOR;
}

// src/coo.cpp:15-22
extern "C" void freeCOO(coo* mat) {
    free(mat->data);
    initCOO(mat);
}

// src/coo.cpp:38-41
extern "C" void setCOOtype(coo* mat, bool type) { mat->type = type; }

// src/coo.cpp:24-36 (type is copied too)
extern "C" coo copyCOO(coo mat) {
    coo r;
    initCOO(&r);
    r.nnz = mat.nnz;
    r.num_cols = mat.num_cols;
    r.num_rows = mat.num_rows;
    r.type = mat.type;
    if (mat.nnz > 0) {
        r.data = (triplet*)malloc(sizeof(triplet) * (size_t)mat.nnz);
        if (r.data) memcpy(r.data, mat.data, sizeof(triplet) * (size_t)mat.nnz);
        else r.nnz = 0;
    }
    return r;
}

// src/coo.cpp:43-79: (major, minor, value) order.  A 64-bit (major, minor)
// key with the value as tie break gives the same array as the reference's
// qsort (its comparator never returns 0; equal triplets are identical).
extern "C" void sortCOO(coo mat) {
    if (mat.nnz <= 1 || !mat.data) return;
    const bool by_row = mat.type == ROW_MAJOR;
    std::sort(mat.data, mat.data + mat.nnz,
              [by_row](const triplet& x, const triplet& y) {
                  int xm = by_row ? x.ridx : x.cidx, ym = by_row ? y.ridx : y.cidx;
                  if (xm != ym) return xm < ym;
                  int xn = by_row ? x.cidx : x.ridx, yn = by_row ? y.cidx : y.ridx;
                  if (xn != yn) return xn < yn;
                  return x.val < y.val;
              });
}

// src/coo.cpp:82-108: "dim num_cols nnz" then "r c [v]" (0-based); dim 2
// means pattern (value 1.1).  Square, as in the reference.
extern "C" void readCOO(coo* mat, char* path) {
    initCOO(mat);
    FILE* fp = fopen(path, "r");
    if (!fp) return;
    int dim = 0, n = 0, nnz = 0;
    if (fscanf(fp, "%d %d %d", &dim, &n, &nnz) != 3 || nnz < 0 || n < 0) {
        fclose(fp);
        return;
    }
    triplet* d = (triplet*)malloc(sizeof(triplet) * (size_t)(nnz > 0 ? nnz : 1));
    int got = 0;
    for (; got < nnz; ++got) {
        int r, c;
        float v = 1.1f;
        if (fscanf(fp, "%d %d", &r, &c) != 2) break;
        if (dim != 2 && fscanf(fp, "%f", &v) != 1) break;
        d[got].ridx = r;
        d[got].cidx = c;
        d[got].val = v;
    }
    fclose(fp);
    mat->data = d;
    mat->nnz = got;
    mat->num_rows = n;
    mat->num_cols = n;
}

namespace {

bool contains_ci(const std::string& s, const char* w) {
    std::string a = s, b = w;
    for (auto& ch : a) ch = (char)std::tolower((unsigned char)ch);
    for (auto& ch : b) ch = (char)std::tolower((unsigned char)ch);
    // whole-token match so "symmetric" does not match "skew-symmetric"
    size_t pos = 0;
    while ((pos = a.find(b, pos)) != std::string::npos) {
        bool left = pos == 0 || std::isspace((unsigned char)a[pos - 1]);
        bool right = pos + b.size() == a.size() ||
                     std::isspace((unsigned char)a[pos + b.size()]);
        if (left && right) return true;
        pos += 1;
    }
    return false;
}

bool read_line(FILE* fp, std::string& out) {
    out.clear();
    int ch;
    bool any = false;
    while ((ch = fgetc(fp)) != EOF) {
        any = true;
        if (ch == '\n') break;
        out.push_back((char)ch);
    }
    return any;
}

}  // namespace

// src/coo.cpp:110-173 with D11 fixed.
extern "C" int spgReadMTX(coo* mat, const char* path, int flags) {
    initCOO(mat);
    FILE* fp = fopen(path, "r");
    if (!fp) return SPG_EIO;
    std::string line;
    if (!read_line(fp, line) || line.rfind("%%MatrixMarket", 0) != 0) {
        fclose(fp);
        return SPG_EIO;
    }
    const std::string banner = line;
    if (!contains_ci(banner, "coordinate")) {
        fclose(fp);
        return SPG_EIO;  // dense "array" files are not sparse inputs
    }
    const bool pattern = contains_ci(banner, "pattern");
    const bool complex_v = contains_ci(banner, "complex");
    const bool skew = contains_ci(banner, "skew-symmetric");
    const bool sym = contains_ci(banner, "symmetric") || contains_ci(banner, "hermitian");
    const bool compat = (flags & SPG_MTX_COMPAT_VALUES) != 0;
    long long nr = -1, nc = -1, ne = -1;
    while (read_line(fp, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        if (sscanf(line.c_str() + p, "%lld %lld %lld", &nr, &nc, &ne) != 3) {
            fclose(fp);
            return SPG_EIO;
        }
        break;
    }
    if (nr < 0 || nc < 0 || ne < 0 || nr > 0x7fffffff || nc > 0x7fffffff) {
        fclose(fp);
        return SPG_EIO;
    }
    long long cap = (sym || skew) ? 2 * ne : ne;
    if (cap > 0x7fffffff) {
        fclose(fp);
        return SPG_EOVERFLOW;
    }
    triplet* d = (triplet*)malloc(sizeof(triplet) * (size_t)(cap > 0 ? cap : 1));
    if (!d) {
        fclose(fp);
        return SPG_ENOMEM;
    }
    long long got = 0, seen = 0;
    while (seen < ne && read_line(fp, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        const char* s = line.c_str() + p;
        char* end = nullptr;
        long long r = strtoll(s, &end, 10);
        if (end == s) break;
        s = end;
        long long c = strtoll(s, &end, 10);
        if (end == s) break;
        s = end;
        double v = 1.0;
        if (!pattern) {
            v = strtod(s, &end);
            if (end == s) break;
            if (complex_v) { /* real part only */ }
        }
        if (r < 1 || r > nr || c < 1 || c > nc) break;
        float fv = compat ? 1.1f : (float)v;
        d[got].ridx = (int)(r - 1);
        d[got].cidx = (int)(c - 1);
        d[got].val = fv;
        ++got;
        if ((sym || skew) && r != c) {
            d[got].ridx = (int)(c - 1);
            d[got].cidx = (int)(r - 1);
            d[got].val = (skew && !compat) ? -fv : fv;
            ++got;
        }
        ++seen;
    }
    fclose(fp);
    if (seen != ne) {
        free(d);
        return SPG_EIO;
    }
    mat->data = d;
    mat->nnz = (int)got;
    mat->num_rows = (int)nr;
    mat->num_cols = (int)nc;
    return SPG_OK;
}

extern "C" void readMTX(coo* mat, char* path) {
    bool type = mat->type;
    if (spgReadMTX(mat, path, 0) != SPG_OK) initCOO(mat);
    mat->type = type;
}

extern "C" int spgWriteMTX(const spg_csr64* C, const char* path) {
    FILE* fp = fopen(path, "w");
    if (!fp) return SPG_EIO;
    fprintf(fp, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(fp, "%d %d %lld\n", C->num_rows, C->num_cols, (long long)C->nnz);
    for (int32_t i = 0; i < C->num_rows; ++i)
        for (int64_t e = C->ptr[i]; e < C->ptr[i + 1]; ++e)
            fprintf(fp, "%d %d %.9g\n", i + 1, C->idx[e] + 1, (double)C->val[e]);
    int rc = ferror(fp) ? SPG_EIO : SPG_OK;
    fclose(fp);
    return rc;
}

// ----------------------------------------------------------------- cm
// src/cm.cpp:6-14
extern "C" void cmInit(cm* mat) {
    mat->ptr = nullptr;
    mat->idx = nullptr;
    mat->val = nullptr;
    mat->nnz = 0;
    mat->num_rows = 0;
    mat->num_cols = 0;
    mat->type = ROW_MAJOR;
}
// src/cm.cpp:16-30
extern "C" void cmFree(cm* mat) {
    free(mat->ptr);
    free(mat->idx);
    free(mat->val);
    cmInit(mat);
}
extern "C" int cmGetNumRows(cm mat) { return mat.num_rows; }
extern "C" int cmGetNumCols(cm mat) { return mat.num_cols; }
extern "C" int cmGetNNZ(cm mat) { return mat.nnz; }
extern "C" void cmSetNumRows(cm* mat, int num_rows) { mat->num_rows = num_rows; }
extern "C" void cmSetNumCols(cm* mat, int num_cols) { mat->num_cols = num_cols; }
extern "C" void cmSetNNZ(cm* mat, int nnz) { mat->nnz = nnz; }
extern "C" void cmSetType(cm* mat, bool type) { mat->type = type; }

extern "C" void spgFreeCSR64(spg_csr64* mat) {
    free(mat->ptr);
    free(mat->idx);
    free(mat->val);
    mat->ptr = nullptr;
    mat->idx = nullptr;
    mat->val = nullptr;
    mat->nnz = 0;
    mat->num_rows = 0;
    mat->num_cols = 0;
}

// ----------------------------------------------------------------- generators
namespace {

int emit_sorted_edges(coo* out, std::vector<uint64_t>& keys, int nrows,
                      int ncols, spg::Xoshiro& rng) {
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (keys.size() > 0x7fffffffull) return SPG_EOVERFLOW;
    triplet* d = (triplet*)malloc(sizeof(triplet) * (keys.size() ? keys.size() : 1));
    if (!d) return SPG_ENOMEM;
    for (size_t e = 0; e < keys.size(); ++e) {
        d[e].ridx = (int)(keys[e] >> 32);
        d[e].cidx = (int)(keys[e] & 0xffffffffu);
        d[e].val = rng.uniform_pm1();
    }
    out->data = d;
    out->nnz = (int)keys.size();
    out->num_rows = nrows;
    out->num_cols = ncols;
    out->type = ROW_MAJOR;
    return SPG_OK;
}

}  // namespace

extern "C" int spgGenUniformRows(coo* out, int n_rows, int n_cols,
                                 int nnz_per_row, uint64_t seed) {
    initCOO(out);
    if (n_rows < 0 || n_cols <= 0 || nnz_per_row < 0 || nnz_per_row > n_cols)
        return SPG_EINVAL;
    long long total = (long long)n_rows * nnz_per_row;
    if (total > 0x7fffffff) return SPG_EOVERFLOW;
    triplet* d = (triplet*)malloc(sizeof(triplet) * (size_t)(total ? total : 1));
    if (!d) return SPG_ENOMEM;
    spg::Xoshiro rng(seed);
    std::vector<int> cols;
    cols.reserve((size_t)nnz_per_row);
    for (int i = 0; i < n_rows; ++i) {
        // Floyd's sampling of nnz_per_row distinct values in [0, n_cols)
        cols.clear();
        for (int j = n_cols - nnz_per_row; j < n_cols; ++j) {
            int t = (int)rng.below((uint64_t)j + 1);
            if (std::find(cols.begin(), cols.end(), t) == cols.end()) cols.push_back(t);
            else cols.push_back(j);
        }
        std::sort(cols.begin(), cols.end());
        for (int q = 0; q < nnz_per_row; ++q) {
            triplet& t = d[(size_t)i * nnz_per_row + q];
            t.ridx = i;
            t.cidx = cols[(size_t)q];
            t.val = rng.uniform_pm1();
        }
    }
    out->data = d;
    out->nnz = (int)total;
    out->num_rows = n_rows;
    out->num_cols = n_cols;
    out->type = ROW_MAJOR;
    return SPG_OK;
}

extern "C" int spgGenRMAT(coo* out, int row_bits, int col_bits, int64_t n_draws,
                          double a, double b, double c, int clip_rows,
                          int clip_cols, uint64_t seed) {
    initCOO(out);
    if (row_bits < 0 || row_bits > 30 || col_bits < 0 || col_bits > 30 ||
        n_draws < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return SPG_EINVAL;
    const int nrows = clip_rows > 0 ? clip_rows : (1 << row_bits);
    const int ncols = clip_cols > 0 ? clip_cols : (1 << col_bits);
    if (nrows > (1 << row_bits) || ncols > (1 << col_bits)) return SPG_EINVAL;
    const int levels = row_bits > col_bits ? row_bits : col_bits;
    spg::Xoshiro rng(seed);
    std::vector<uint64_t> keys;
    try {
        keys.reserve((size_t)n_draws);
    } catch (...) {
        return SPG_ENOMEM;
    }
    const double ab = a + b, abc = a + b + c;
    for (int64_t e = 0; e < n_draws; ++e) {
        for (;;) {
            uint64_t r = 0, q = 0;
            for (int l = 0; l < levels; ++l) {
                double u = rng.uniform01();
                int rb = u >= ab;                        // quadrants c, d
                int cb = (u >= a && u < ab) || u >= abc; // quadrants b, d
                if (l < row_bits) r = (r << 1) | (uint64_t)rb;
                if (l < col_bits) q = (q << 1) | (uint64_t)cb;
            }
            if ((int64_t)r < nrows && (int64_t)q < ncols) {
                keys.push_back((r << 32) | q);
                break;
            }
        }
    }
    return emit_sorted_edges(out, keys, nrows, ncols, rng);
}

extern "C" int spgGenStencil27(coo* out, int nx, int ny, int nz, uint64_t seed) {
    initCOO(out);
    if (nx <= 0 || ny <= 0 || nz <= 0) return SPG_EINVAL;
    long long n = (long long)nx * ny * nz;
    if (n > 0x7fffffff || n * 27 > 0x7fffffff) return SPG_EOVERFLOW;
    auto cnt = [](int v, int nv) { return (v > 0) + 1 + (v + 1 < nv); };
    long long total = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                total += (long long)cnt(x, nx) * cnt(y, ny) * cnt(z, nz);
    triplet* d = (triplet*)malloc(sizeof(triplet) * (size_t)(total ? total : 1));
    if (!d) return SPG_ENOMEM;
    spg::Xoshiro rng(seed);
    long long e = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int row = x + nx * (y + ny * z);
                // neighbours in ascending column order: dz, then dy, then dx
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            int X = x + dx, Y = y + dy, Z = z + dz;
                            if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz)
                                continue;
                            triplet& t = d[e++];
                            t.ridx = row;
                            t.cidx = X + nx * (Y + ny * Z);
                            if (dx == 0 && dy == 0 && dz == 0) t.val = 26.0f;
                            else t.val = -1.0f + 0.01f * rng.uniform_pm1();
                        }
            }
    out->data = d;
    out->nnz = (int)total;
    out->num_rows = (int)n;
    out->num_cols = (int)n;
    out->type = ROW_MAJOR;
    return SPG_OK;
}

// The expansion kernels keep three counts in 32 bits: a virtual row's slot
// cursor (every product of the row is reserved through it, so its bound must
// fit), the run offsets of a row chunk inside one tile (at most the tile's
// bound) and the product offsets of a chunk of 256 consecutive A entries of
// untiled rows (each entry contributes its B row, which is at most the
// longest B row and at most its untiled row's bound).  Each store position
// is computed as (block << 6) + lane, so the limit keeps 64 spare.  The
// reference's int offsets overflow silently instead (src/mm.cu:173-180).
extern "C" int spgSlotLimitStatus(int64_t max_vrow_bound, int64_t max_untiled_bound,
                                  int64_t max_b_row) {
    const int64_t lim = (int64_t)INT32_MAX - 64;
    if (max_vrow_bound < 0 || max_untiled_bound < 0 || max_b_row < 0) return SPG_EINVAL;
    if (max_vrow_bound > lim) return SPG_EOVERFLOW;
    const int64_t per_entry = std::min(max_b_row, max_untiled_bound);
    if (per_entry > lim / 256) return SPG_EOVERFLOW;
    return SPG_OK;
}
