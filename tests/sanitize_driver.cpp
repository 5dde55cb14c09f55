// sanitize_driver.cpp - runs the host code of the library (Matrix Market
// loader and writer, COO containers and sort, generators, the 32-bit limit
// check) and the CPU oracle under AddressSanitizer and UBSan.  Built and run
// by tests/test_sanitizers.py with g++/gcc (no HIP code is involved):
//
//   sanitize_driver OUTDIR FILE.mtx...
//
// Every input file is read in both value modes, sorted by rows and by
// columns, copied, written back as C-style CSR through spgWriteMTX and read
// again; the oracle runs every restated phase on a small R-MAT product.  Any
// memory error or undefined behaviour aborts the program (the sanitizers are
// built without recovery), and a failed consistency check exits with 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/spgemm.h"
#include "../oracle/oracle.h"

static int failures = 0;
#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

// C = the COO's matrix as CSR (duplicates kept), for spgWriteMTX.
static spg_csr64 csr_of(const coo& a) {
    spg_csr64 c{};
    c.num_rows = a.num_rows;
    c.num_cols = a.num_cols;
    c.nnz = a.nnz;
    c.ptr = (int64_t*)calloc((size_t)a.num_rows + 1, sizeof(int64_t));
    c.idx = (int32_t*)malloc(sizeof(int32_t) * ((size_t)a.nnz + 1));
    c.val = (float*)malloc(sizeof(float) * ((size_t)a.nnz + 1));
    for (int e = 0; e < a.nnz; ++e) c.ptr[a.data[e].ridx + 1] += 1;
    for (int i = 0; i < a.num_rows; ++i) c.ptr[i + 1] += c.ptr[i];
    std::vector<int64_t> cur(c.ptr, c.ptr + a.num_rows + 1);
    for (int e = 0; e < a.nnz; ++e) {
        const int64_t p = cur[a.data[e].ridx]++;
        c.idx[p] = a.data[e].cidx;
        c.val[p] = a.data[e].val;
    }
    return c;
}

static void exercise_file(const char* path, const std::string& outdir, int q) {
    for (int flags = 0; flags <= SPG_MTX_COMPAT_VALUES; ++flags) {
        coo a;
        initCOO(&a);
        const int rc = spgReadMTX(&a, path, flags);
        if (rc != SPG_OK) {  // malformed inputs are expected to fail cleanly
            EXPECT(a.nnz == 0);
            freeCOO(&a);
            continue;
        }
        for (int e = 0; e < a.nnz; ++e)
            EXPECT(a.data[e].ridx >= 0 && a.data[e].ridx < a.num_rows && a.data[e].cidx >= 0 &&
                   a.data[e].cidx < a.num_cols);
        for (int type = ROW_MAJOR; type <= COL_MAJOR; ++type) {
            coo b = copyCOO(a);
            setCOOtype(&b, type);
            sortCOO(b);
            for (int e = 1; e < b.nnz; ++e) {
                const triplet& x = b.data[e - 1];
                const triplet& y = b.data[e];
                const int xm = type == ROW_MAJOR ? x.ridx : x.cidx, ym = type == ROW_MAJOR ? y.ridx : y.cidx;
                EXPECT(xm <= ym);
            }
            freeCOO(&b);
        }
        spg_csr64 c = csr_of(a);
        const std::string out = outdir + "/rt" + std::to_string(q) + "_" + std::to_string(flags) + ".mtx";
        EXPECT(spgWriteMTX(&c, out.c_str()) == SPG_OK);
        coo r;
        initCOO(&r);
        EXPECT(spgReadMTX(&r, out.c_str(), 0) == SPG_OK);
        EXPECT(r.nnz == a.nnz && r.num_rows == a.num_rows && r.num_cols == a.num_cols);
        freeCOO(&r);
        spgFreeCSR64(&c);
        // the reference entry point (no status; empty on error)
        coo l;
        initCOO(&l);
        readMTX(&l, (char*)path);
        freeCOO(&l);
        freeCOO(&a);
    }
    coo rc;
    readCOO(&rc, (char*)path);  // not the "dim n nnz" format: ends empty or partial
    freeCOO(&rc);
}

static void exercise_oracle() {
    coo g;
    initCOO(&g);
    EXPECT(spgGenRMAT(&g, 9, 9, 8 << 9, 0.57, 0.19, 0.19, 500, 480, 3) == SPG_OK);
    const int m = g.num_rows, K = g.num_cols, n = g.num_cols;
    std::vector<orc_triplet> t((size_t)g.nnz);
    for (int e = 0; e < g.nnz; ++e) t[e] = {g.data[e].ridx, g.data[e].cidx, g.data[e].val};
    std::vector<orc_triplet> tc = t;
    orc_sort_coo(g.nnz, tc.data(), ORC_COL_MAJOR);
    std::vector<int32_t> a_ptr(K + 1), a_idx(g.nnz + 1), b_ptr(m + 1), b_idx(g.nnz + 1);
    std::vector<float> a_val(g.nnz + 1), b_val(g.nnz + 1);
    orc_coo2cm(g.nnz, tc.data(), ORC_COL_MAJOR, K, a_ptr.data(), a_idx.data(), a_val.data());
    std::vector<orc_triplet> tr = t;
    orc_sort_coo(g.nnz, tr.data(), ORC_ROW_MAJOR);
    orc_coo2cm(g.nnz, tr.data(), ORC_ROW_MAJOR, m, b_ptr.data(), b_idx.data(), b_val.data());
    std::vector<int32_t> r_ptr(m + 1), r_idx(g.nnz + 1);
    std::vector<float> r_val(g.nnz + 1);
    orc_csc_to_csr(m, K, a_ptr.data(), a_idx.data(), a_val.data(), r_ptr.data(), r_idx.data(),
                   r_val.data());
    EXPECT(r_ptr == b_ptr);  // A*A: CSR(A) from CSC(A) equals CSR(A) from the COO
    std::vector<int64_t> ub(m), base(m + 1);
    const int64_t nprod = orc_upper_bound(m, K, a_ptr.data(), a_idx.data(), b_ptr.data(), ub.data(),
                                          base.data());
    std::vector<int32_t> bins(K), perm(K);
    std::vector<int64_t> boff(ORC_NUM_KBINS + 1);
    orc_kbins(K, a_ptr.data(), b_ptr.data(), nprod, bins.data(), perm.data(), boff.data());
    std::vector<int32_t> icol(nprod + 1);
    std::vector<float> ival(nprod + 1);
    orc_expand(m, K, a_ptr.data(), a_idx.data(), a_val.data(), b_ptr.data(), b_idx.data(),
               b_val.data(), base.data(), icol.data(), ival.data());
    std::vector<int64_t> cnt(m);
    const int64_t nnz = orc_symbolic_from_slots(m, n, base.data(), icol.data(), cnt.data());
    int64_t *cp = nullptr, *gp = nullptr;
    int32_t *ci = nullptr, *gi = nullptr;
    float* cv = nullptr;
    double *gv = nullptr, *gs = nullptr;
    int64_t np2 = 0;
    const int64_t nnz2 = orc_spgemm_restated(m, n, K, a_ptr.data(), a_idx.data(), a_val.data(),
                                             b_ptr.data(), b_idx.data(), b_val.data(), &cp, &ci, &cv,
                                             &np2);
    EXPECT(nnz2 == nnz && np2 == nprod);
    const int64_t nnz3 = orc_gustavson(m, n, r_ptr.data(), r_idx.data(), r_val.data(), b_ptr.data(),
                                       b_idx.data(), b_val.data(), nullptr, m, 1, &gp, &gi, &gv, &gs);
    EXPECT(nnz3 == nnz);
    orc_free(gp);
    orc_free(gi);
    orc_free(gv);
    orc_free(gs);
    const int32_t sel[3] = {0, m / 2, m - 1};
    orc_gustavson(m, n, r_ptr.data(), r_idx.data(), r_val.data(), b_ptr.data(), b_idx.data(),
                  b_val.data(), sel, 3, 0, &gp, &gi, &gv, &gs);
    orc_free(gp);
    orc_free(gi);
    orc_free(gv);
    orc_free(gs);
    orc_free(cp);
    orc_free(ci);
    orc_free(cv);
    for (int threads = 1; threads <= 3; threads += 2) {
        double tms[5];
        int64_t nz = 0;
        double cs = 0;
        EXPECT(orc_cpu_baseline(m, n, K, a_ptr.data(), a_idx.data(), a_val.data(), b_ptr.data(),
                                b_idx.data(), b_val.data(), 0, m, threads, tms, &nz, &cs) == nprod);
        EXPECT(nz == nnz);
    }
    freeCOO(&g);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s OUTDIR FILE.mtx...\n", argv[0]);
        return 2;
    }
    const std::string outdir = argv[1];
    for (int a = 2; a < argc; ++a) exercise_file(argv[a], outdir, a);
    coo u, s;
    EXPECT(spgGenUniformRows(&u, 300, 200, 7, 1) == SPG_OK);
    EXPECT(spgGenStencil27(&s, 4, 5, 6, 2) == SPG_OK);
    freeCOO(&u);
    freeCOO(&s);
    EXPECT(spgGenUniformRows(&u, 3, 2, 5, 1) != SPG_OK);  // more per row than columns
    EXPECT(u.nnz == 0);
    freeCOO(&u);
    EXPECT(spgSlotLimitStatus(100, 10, 10) == SPG_OK);
    EXPECT(spgSlotLimitStatus((int64_t)1 << 40, 0, 0) == SPG_EOVERFLOW);
    exercise_oracle();
    printf("sanitize driver: %d failed checks\n", failures);
    return failures ? 1 : 0;
}
