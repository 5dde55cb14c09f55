// spgemm_outer - command-line driver, the counterpart of the reference's
// bin/outer (src/main.cu:11-45): C = A * B for two Matrix Market files.
//
//   spgemm_outer A.mtx B.mtx [--compat-values] [--simple] [--out C.mtx]
//                [--repeat N]
//
// Phase order as src/main.cu:18-37: read both files, sort A by column and B
// by row, COO2CM(A) (CSC), COO2CM(B) (CSR), InitGEMM, CategorizeGEMM,
// BinGEMM (or SimpleGEMM), MergeGEMM; then C is fetched (the reference
// never returns it) and the reference's two counts are printed: the
// structural nnz and the number of values > 1e-4 (src/mm.cu:438-449).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/spgemm.h"

static int check(int rc, spg_context* ctx, const char* what) {
    if (rc != SPG_OK) {
        fprintf(stderr, "%s failed: %s (%s)\n", what, spgGetErrorString(rc),
                ctx ? spgGetLastError(ctx) : "");
        exit(1);
    }
    return rc;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr,
                "usage: %s A.mtx B.mtx [--compat-values] [--simple] [--out C.mtx] "
                "[--repeat N]\n",
                argv[0]);
        return 2;
    }
    int flags = 0, repeat = 1;
    bool simple = false;
    const char* out = nullptr;
    for (int a = 3; a < argc; ++a) {
        if (!strcmp(argv[a], "--compat-values")) flags |= SPG_MTX_COMPAT_VALUES;
        else if (!strcmp(argv[a], "--simple")) simple = true;
        else if (!strcmp(argv[a], "--out") && a + 1 < argc) out = argv[++a];
        else if (!strcmp(argv[a], "--repeat") && a + 1 < argc) repeat = atoi(argv[++a]);
    }
    coo A, B;
    initCOO(&A);
    initCOO(&B);
    check(spgReadMTX(&A, argv[1], flags), nullptr, "read A");
    check(spgReadMTX(&B, argv[2], flags), nullptr, "read B");
    setCOOtype(&A, COL_MAJOR);
    setCOOtype(&B, ROW_MAJOR);
    sortCOO(A);
    sortCOO(B);
    spg_context* ctx = nullptr;
    check(spgCreate(&ctx, 0), nullptr, "spgCreate");
    check(spgCOO2CM(ctx, A, nullptr), ctx, "COO2CM(A)");
    check(spgCOO2CM(ctx, B, nullptr), ctx, "COO2CM(B)");
    int64_t nprod = 0;
    double best_ms = 1e300;
    for (int r = 0; r < repeat; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        check(spgInitGEMM(ctx, nullptr, &nprod), ctx, "InitGEMM");
        if (simple) {
            check(spgSimpleGEMM(ctx), ctx, "SimpleGEMM");
        } else {
            check(spgCategorizeGEMM(ctx, nullptr), ctx, "CategorizeGEMM");
            check(spgBinGEMM(ctx), ctx, "BinGEMM");
        }
        check(spgMergeGEMM(ctx), ctx, "MergeGEMM");
        check(spgSynchronize(ctx), ctx, "Synchronize");
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms < best_ms) best_ms = ms;
    }
    spg_csr64 C;
    check(spgGetC(ctx, &C), ctx, "GetC");
    long long nnzCf = 0;
    for (int64_t e = 0; e < C.nnz; ++e) nnzCf += C.val[e] > 0.0001f;
    printf("nprod %lld\n", (long long)nprod);
    printf("%lld %lld\n", (long long)C.nnz, nnzCf);
    printf("time_ms %.3f GFLOP/s %.2f\n", best_ms, 2.0 * (double)nprod / (best_ms * 1e6));
    if (out) check(spgWriteMTX(&C, out), ctx, "write C");
    spgFreeCSR64(&C);
    spgDestroy(ctx);
    freeCOO(&A);
    freeCOO(&B);
    return 0;
}
