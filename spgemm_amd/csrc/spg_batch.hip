// spg_batch.hip - many independent products in one pass of the phase
// kernels (no reference counterpart; the reference multiplies one product
// per process run).
//
// The batch is stacked block-diagonally: A = diag(A_0 .. A_{N-1}) as one
// CSC, B = diag(B_0 .. B_{N-1}) as one CSR, so that C = diag(C_0 ..
// C_{N-1}) and every product's (i, k, j) triples stay inside its own block.
// Product b's columns start at a multiple of 2^13 (the largest tile width),
// and every C row and every k carries its product's column window, which
// the tiling, the split points of B rows and the rank merge use in place
// of B's full column range.  One run of the phase kernels (each a
// persistent kernel over the work of all products) then computes the whole
// batch; spgBatchMultiply can record that run once and capture it into a
// HIP graph that replays the batch with a single launch.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "spg_internal.h"

using namespace spg;

namespace spg {

void graph_release(spg_context* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_pred = nullptr;
    c->graph_pred_version = -1;
}

}  // namespace spg

namespace {

constexpr int64_t kWindowAlign = 1LL << kTileMaxLog2;

int check_cm(spg_context* c, const cm& x, bool csc, int b) {
    const int nmajor = csc ? x.num_cols : x.num_rows;
    if ((x.type == COL_MAJOR) != csc || nmajor < 0 || x.num_rows < 0 || x.num_cols < 0 ||
        x.nnz < 0 || !x.ptr || (x.nnz > 0 && (!x.idx || !x.val)) || x.ptr[0] != 0 ||
        x.ptr[nmajor] != x.nnz)
        return fail(c, SPG_EINVAL,
                    std::string("batch product ") + std::to_string(b) +
                        (csc ? ": A must be a valid CSC (COL_MAJOR)" : ": B must be a valid CSR (ROW_MAJOR)"));
    return SPG_OK;
}

int run_batch(spg_context* c, const spg_predictor* p, spg_learned_stats* st) {
    if (c->lat_on) SPG_TRY(launch_latency_start(c));
    SPG_TRY(spgInitGEMM(c, nullptr, nullptr));
    if (c->lat_on) SPG_TRY(launch_vprod(c));
    SPG_TRY(spgCategorizeGEMM(c, nullptr));
    SPG_TRY(spgBinGEMM(c));
    if (p) {
        SPG_TRY(spgPredictRows(c, p, nullptr));
        return spgLearnedMergeGEMM(c, st);
    }
    SPG_TRY(spgSymbolicGEMM(c, nullptr));
    return spgMergeGEMM(c);
}

}  // namespace

extern "C" {

int spgBatchSetOperands(spg_context* c, int count, const cm* A, const cm* B) {
    if (!c) return SPG_EINVAL;
    if (count < 1 || !A || !B) return fail(c, SPG_EINVAL, "batch needs count >= 1 and A, B arrays");
    int64_t m = 0, K = 0, ncol = 0, nnzA = 0, nnzB = 0;
    std::vector<int64_t> row_off(count + 1), k_off(count + 1), col_off(count);
    std::vector<int32_t> ncols(count);
    for (int b = 0; b < count; ++b) {
        SPG_TRY(check_cm(c, A[b], true, b));
        SPG_TRY(check_cm(c, B[b], false, b));
        if (A[b].num_cols != B[b].num_rows)
            return fail(c, SPG_EINVAL, "batch product " + std::to_string(b) + ": A.num_cols != B.num_rows");
        row_off[b] = m;
        k_off[b] = K;
        col_off[b] = ncol;
        ncols[b] = B[b].num_cols;
        m += A[b].num_rows;
        K += A[b].num_cols;
        ncol += (B[b].num_cols + kWindowAlign - 1) / kWindowAlign * kWindowAlign;
        nnzA += A[b].nnz;
        nnzB += B[b].nnz;
    }
    row_off[count] = m;
    k_off[count] = K;
    if (m > INT_MAX || K > INT_MAX || ncol > INT_MAX || nnzA > INT_MAX || nnzB > INT_MAX)
        return fail(c, SPG_EOVERFLOW, "stacked batch exceeds 32-bit rows, columns or nnz");
    // stacked CSC of A (row indices shifted) and CSR of B (column indices shifted)
    std::vector<int> ap((size_t)K + 1), ai((size_t)nnzA), bp((size_t)K + 1), bi((size_t)nnzB);
    std::vector<float> av((size_t)nnzA), bv((size_t)nnzB);
    std::vector<int2> rwin((size_t)std::max<int64_t>(m, 1)), kwin((size_t)std::max<int64_t>(K, 1));
    std::vector<int32_t> rprod((size_t)std::max<int64_t>(m, 1));
    int64_t ea = 0, eb = 0;
    for (int b = 0; b < count; ++b) {
        const cm& a = A[b];
        const cm& x = B[b];
        const int2 win = make_int2((int)col_off[b], ncols[b]);
        for (int k = 0; k < a.num_cols; ++k) {
            ap[(size_t)(k_off[b] + k)] = (int)(ea + a.ptr[k]);
            bp[(size_t)(k_off[b] + k)] = (int)(eb + x.ptr[k]);
            kwin[(size_t)(k_off[b] + k)] = win;
        }
        for (int e = 0; e < a.nnz; ++e) {
            ai[(size_t)(ea + e)] = a.idx[e] + (int)row_off[b];
            av[(size_t)(ea + e)] = a.val[e];
        }
        for (int e = 0; e < x.nnz; ++e) {
            bi[(size_t)(eb + e)] = x.idx[e] + (int)col_off[b];
            bv[(size_t)(eb + e)] = x.val[e];
        }
        for (int i = 0; i < a.num_rows; ++i) {
            rwin[(size_t)(row_off[b] + i)] = win;
            rprod[(size_t)(row_off[b] + i)] = b;
        }
        ea += a.nnz;
        eb += x.nnz;
    }
    ap[(size_t)K] = (int)ea;
    bp[(size_t)K] = (int)eb;
    const int n = (int)(col_off[count - 1] + ncols[count - 1]);
    cm bigA{ap.data(), ai.data(), av.data(), (int)nnzA, (int)m, (int)K, COL_MAJOR};
    cm bigB{bp.data(), bi.data(), bv.data(), (int)nnzB, (int)K, n, ROW_MAJOR};
    SPG_TRY(spgSetOperand(c, SPG_OPERAND_A, &bigA));  // validates ranges and order
    SPG_TRY(spgSetOperand(c, SPG_OPERAND_B, &bigB));
    SPG_TRY(ensure(c, c->rwinb, sizeof(int2) * rwin.size()));
    SPG_TRY(ensure(c, c->kwinb, sizeof(int2) * kwin.size()));
    SPG_TRY(ensure(c, c->rprod, sizeof(int32_t) * rprod.size()));
    SPG_TRY(ensure(c, c->lat_t, sizeof(uint64_t) * ((size_t)count + 1)));
    SPG_HIP(hipMemcpyAsync(c->rprod.p, rprod.data(), sizeof(int32_t) * rprod.size(),
                           hipMemcpyHostToDevice, c->stream));
    SPG_HIP(hipMemcpyAsync(c->rwinb.p, rwin.data(), sizeof(int2) * rwin.size(),
                           hipMemcpyHostToDevice, c->stream));
    SPG_HIP(hipMemcpyAsync(c->kwinb.p, kwin.data(), sizeof(int2) * kwin.size(),
                           hipMemcpyHostToDevice, c->stream));
    SPG_HIP(hipStreamSynchronize(c->stream));
    c->rwin = c->rwinb.p;
    c->kwin = c->kwinb.p;
    c->batch_count = count;
    c->b_row_off = row_off;
    c->b_k_off = k_off;
    c->b_col_off = col_off;
    c->b_ncols = ncols;
    return SPG_OK;
}

int spgBatchGetLayout(spg_context* c, int64_t* row_off, int64_t* col_off) {
    if (!c) return SPG_EINVAL;
    if (c->batch_count <= 0) return fail(c, SPG_EORDER, "spgBatchSetOperands must run first");
    if (row_off)
        for (int b = 0; b <= c->batch_count; ++b) row_off[b] = c->b_row_off[(size_t)b];
    if (col_off)
        for (int b = 0; b < c->batch_count; ++b) col_off[b] = c->b_col_off[(size_t)b];
    return SPG_OK;
}

int spgBatchMultiply(spg_context* c, const spg_predictor* p, int flags, spg_learned_stats* st) {
    if (!c) return SPG_EINVAL;
    if (c->batch_count <= 0) return fail(c, SPG_EORDER, "spgBatchSetOperands must run first");
    if (flags & ~(SPG_BATCH_GRAPH | SPG_BATCH_LATENCY)) return fail(c, SPG_EINVAL, "unknown batch flags");
    SPG_HIP(hipSetDevice(c->device));
    spg_learned_stats local{};
    spg_learned_stats* sp = p ? &local : nullptr;
    const bool lat = (flags & SPG_BATCH_LATENCY) != 0;
    if (!(flags & SPG_BATCH_GRAPH)) {
        c->lat_on = lat;
        const int rc = run_batch(c, p, sp);
        c->lat_on = false;
        SPG_TRY(rc);
        c->lat_valid = lat;
        if (st && p) *st = local;
        return SPG_OK;
    }
    if (!c->stream) return fail(c, SPG_EINVAL, "graph capture needs a non-default stream");
    const bool valid = c->graph_exec && c->graph_learned == (p != nullptr) && c->graph_pred == p &&
                       (!p || predictor_version(p) == c->graph_pred_version) && c->graph_lat == lat;
    if (valid) {
        SPG_HIP(hipGraphLaunch(c->graph_exec, c->stream));
        c->stage = 5;
        c->lat_valid = lat;
        if (st && p) *st = c->graph_stats;  // the counts of the recorded run
        return SPG_OK;
    }
    graph_release(c);
    // 1) an ordinary run whose host readbacks are recorded
    c->plan.clear();
    c->plan_mode = 1;
    c->lat_on = lat;
    int rc = run_batch(c, p, sp);
    c->plan_mode = 0;
    if (rc != SPG_OK) {
        c->lat_on = false;
        return rc;
    }
    if (p) c->graph_stats = local;
    // 2) the same sequence captured, with the readbacks replayed from the log
    SPG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->plan_mode = 2;
    c->plan_pos = 0;
    spg_learned_stats scratch_stats{};
    rc = run_batch(c, p, p ? &scratch_stats : nullptr);
    c->plan_mode = 0;
    c->lat_on = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != SPG_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc != SPG_OK ? rc : fail(c, SPG_EHIP, "batch graph capture failed");
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return fail(c, SPG_EHIP, "batch graph instantiation failed");
    }
    c->graph = g;
    c->graph_exec = x;
    c->graph_learned = p != nullptr;
    c->graph_pred = p;
    c->graph_pred_version = p ? predictor_version(p) : -1;
    c->graph_lat = lat;
    c->lat_valid = lat;
    c->stage = 5;  // C of the recorded run is in place
    if (st && p) *st = local;
    return SPG_OK;
}

int spgBatchGetLatency(spg_context* c, double* ms) {
    if (!c || !ms) return SPG_EINVAL;
    if (c->batch_count <= 0) return fail(c, SPG_EORDER, "spgBatchSetOperands must run first");
    if (!c->lat_valid)
        return fail(c, SPG_EORDER, "the last spgBatchMultiply did not record completion times");
    SPG_HIP(hipSetDevice(c->device));
    int khz = 0;
    SPG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    if (khz <= 0) return fail(c, SPG_EHIP, "no wall-clock rate");
    std::vector<uint64_t> t((size_t)c->batch_count + 1);
    SPG_HIP(hipMemcpyAsync(t.data(), c->lat_t.p, sizeof(uint64_t) * t.size(), hipMemcpyDeviceToHost,
                           c->stream));
    SPG_HIP(hipStreamSynchronize(c->stream));
    const uint64_t t0 = t[(size_t)c->batch_count];
    for (int b = 0; b < c->batch_count; ++b)
        ms[b] = t[(size_t)b] > t0 ? (double)(t[(size_t)b] - t0) / (double)khz : 0.0;
    return SPG_OK;
}

int spgBatchGetC(spg_context* c, int b, spg_csr64* out) {
    if (!c || !out) return SPG_EINVAL;
    if (c->batch_count <= 0) return fail(c, SPG_EORDER, "spgBatchSetOperands must run first");
    if (b < 0 || b >= c->batch_count) return fail(c, SPG_EINVAL, "batch index out of range");
    if (c->stage < 5) return fail(c, SPG_EORDER, "spgBatchMultiply must run first");
    SPG_HIP(hipSetDevice(c->device));
    const int64_t r0 = c->b_row_off[(size_t)b], r1 = c->b_row_off[(size_t)b + 1];
    const int mb = (int)(r1 - r0);
    out->ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)mb + 1));
    if (!out->ptr) return fail(c, SPG_ENOMEM, "host C");
    out->idx = nullptr;
    out->val = nullptr;
    SPG_HIP(hipMemcpyAsync(out->ptr, (const int64_t*)c->crow.p + r0,
                           sizeof(int64_t) * ((size_t)mb + 1), hipMemcpyDeviceToHost, c->stream));
    SPG_HIP(hipStreamSynchronize(c->stream));
    const int64_t e0 = out->ptr[0], nnz = out->ptr[mb] - e0;
    out->idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
    out->val = (float*)malloc(sizeof(float) * (size_t)(nnz + 1));
    if (!out->idx || !out->val) {
        spgFreeCSR64(out);
        return fail(c, SPG_ENOMEM, "host C");
    }
    if (nnz > 0) {
        SPG_HIP(hipMemcpyAsync(out->idx, (const int32_t*)c->ccol.p + e0, sizeof(int32_t) * nnz,
                               hipMemcpyDeviceToHost, c->stream));
        SPG_HIP(hipMemcpyAsync(out->val, (const float*)c->cval.p + e0, sizeof(float) * nnz,
                               hipMemcpyDeviceToHost, c->stream));
        SPG_HIP(hipStreamSynchronize(c->stream));
    }
    const int32_t co = (int32_t)c->b_col_off[(size_t)b];
    for (int64_t q = 0; q < nnz; ++q) out->idx[q] -= co;
    for (int i = 0; i <= mb; ++i) out->ptr[i] -= e0;
    out->nnz = nnz;
    out->num_rows = mb;
    out->num_cols = c->b_ncols[(size_t)b];
    return SPG_OK;
}

}  // extern "C"
