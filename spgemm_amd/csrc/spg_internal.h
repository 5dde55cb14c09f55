// spg_internal.h - context and launch-wrapper declarations shared by the
// HIP translation units of libspgemm_amd.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spgemm.h"

namespace spg {

// Upper limits of the merge structures held in LDS.
constexpr int kBlock = 256;             // threads per workgroup
constexpr int kPT = 512;                // threads of the persistent kernels
constexpr int kHashMaxLog2 = 13;        // largest LDS hash: 8192 slots
constexpr int kHashMaxCount = 1 << (kHashMaxLog2 - 1);  // <= 4096 keys
constexpr int kTileMaxLog2 = 13;        // dense tile: 8192 FP64 accumulators
constexpr int kHashMinLog2 = 6;         // smallest hash: 64 slots
// Merge/symbolic classes of a virtual row.
constexpr int kNumHashClasses = kHashMaxLog2 - kHashMinLog2 + 1;  // 64..8192
constexpr int kNumVClasses = kNumHashClasses + 5;  // + tile (large), rank, tile small/medium, rank (wave)
// The context's work counters (64-bit): slots below kCtrPoolBase have fixed
// uses; above it, kCtrPoolSlots slots 128 bytes apart are handed to the
// merge kernels of one merge phase, zeroed together when the phase starts.
constexpr int kCtrWords = 1024;
constexpr int kCtrPoolBase = 64;
constexpr int kCtrPoolStride = 16;
constexpr int kCtrPoolSlots = (kCtrWords - kCtrPoolBase) / kCtrPoolStride;

// Grow-only device buffer.  fixed: memory the caller owns (an output of the
// allocation-free entry point), never replaced; too small is an error.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool fixed = false;
};

struct Operand {
    const int* ptr = nullptr;
    const int* idx = nullptr;
    const float* val = nullptr;
    int num_rows = 0, num_cols = 0, nnz = 0;
    bool set = false;
    bool owned = false;
    DevBuf bptr, bidx, bval;  // storage when owned
};

struct Timer {
    hipEvent_t ev[2 * SPG_NUM_PHASES] = {};
    bool recorded[SPG_NUM_PHASES] = {};
};

}  // namespace spg

struct spg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // two more streams (default priority) on which merge classes run beside
    // the dense-tile class, forked from and joined to `stream` by events;
    // created on first use (never in the allocation-free mode)
    hipStream_t aux[2] = {};
    hipEvent_t fork_ev = nullptr, join_ev[2] = {};
    bool aux_tried = false, aux_ok = false;
    spg_config cfg{};
    std::string err;

    spg::Operand A, B;  // A: CSC (ptr over columns), B: CSR (ptr over rows)
    // Batched products (block-diagonal A and B): the column window (lo,
    // width) of every C row and of every k, or null for one product (all of
    // B's columns).  Window starts are multiples of 2^kTileMaxLog2.
    spg::DevBuf rwinb, kwinb;
    void* rwin = nullptr;
    void* kwin = nullptr;

    // shape of the product
    int32_t m = 0, K = 0, n = 0;
    int32_t T = 1;        // column tiles of a long row
    int32_t W = 0;        // tile width
    int64_t V = 0;        // virtual rows
    int64_t nprod = 0;
    int64_t nnzC = 0;
    bool have_long = false;

    // phase state
    int stage = 0;  // 0 none, 1 init, 2 categorize, 3 expanded, 4 symbolic, 5 merged

    // init products
    spg::DevBuf ub, ntile, vstart, vrow_row, vlo, vwid, bsplit, vub, vbase, cursor;
    spg::DevBuf counters;  // work counters of persistent kernels
    spg::DevBuf arow_ptr, arow_k, arow_v;  // A transposed: row i -> its (k, a_ik)
    spg::DevBuf tilebits;          // column bitmap of every tile-width virtual row
    spg::DevBuf feat;              // predictor features of every row (float4)
    int num_cu = 256;
    // categorize products
    spg::DevBuf kbin, kperm, kbin_off, kitems, kitem_off;
    int64_t kbin_off_h[SPG_NUM_KBINS + 1] = {};
    int64_t kitems_bin_h[SPG_NUM_KBINS + 1] = {};  // item offsets at bin starts
    // intermediates, in groups of 64 products: each 512-byte group holds the
    // 64 columns, then the 64 values (product s: word ((s >> 6) << 7) + (s & 63)
    // of the column half, 64 words later its value)
    spg::DevBuf inter;
    // B's (column, value) pairs interleaved, for the row-chunk expansion's
    // gathers (one 8-byte load per product)
    spg::DevBuf bpair;
    // slot of the first product of every A entry of an untiled row (row-wise
    // expansion by entry chunks), -1 for the entries of tiled rows
    spg::DevBuf ebase;
    // symbolic / merge
    spg::DevBuf vclass, vlist, vclass_off, vcnt, vout, crow, ccol, cval;
    // learned merge: row predictions, per-virtual-row capacities, padded
    // output (pbase/pcol/pval) and the fallback output (fbase/fcol/fval)
    spg::DevBuf pred, vcap, pbase, fbase, pcol, pval, fcol, fval;
    bool have_pred = false;
    // scratch for scans and partitions
    spg::DevBuf scratch;
    size_t scratch_top = 0;
    spg::DevBuf pinned_small;  // host-pinned staging for small readbacks

    spg::Timer timer;
    double phase_ms[SPG_NUM_PHASES] = {};

    // Allocation-free mode (spgMultiplyNoAlloc): every device buffer is a
    // range of one fixed per-device array (arena, bump-allocated from
    // arena_top and never freed), and readbacks go through host_stage
    // (pageable, kReadbackMax bytes) instead of pinned memory.
    bool noalloc = false;
    char* arena = nullptr;
    size_t arena_cap = 0, arena_top = 0;
    unsigned char* host_stage = nullptr;

    // Replay log.  A recorded run appends every host readback (sizes, class
    // offsets) to plan; a replayed run (the one captured into a HIP graph)
    // takes them from plan instead of reading the device, so it contains no
    // synchronization.  Valid while the operands' structure, the config and
    // the predictor are unchanged.
    int plan_mode = 0;  // 0: off, 1: record, 2: replay
    int ctr_pool = 0;   // counter slots handed out in this merge phase
    std::vector<unsigned char> plan;
    size_t plan_pos = 0;

    // Batch of products stacked block-diagonally (spgBatchSetOperands):
    // product b owns C rows [row_off[b], row_off[b+1]), k [k_off[b],
    // k_off[b+1]) and columns [col_off[b], col_off[b] + ncols[b]).
    int batch_count = 0;
    std::vector<int64_t> b_row_off, b_k_off, b_col_off;
    std::vector<int32_t> b_ncols;
    // completion time of every product (SPG_BATCH_LATENCY): lat_t holds
    // batch_count wall-clock times and then the run's start time; rprod is
    // the product of every stacked row, vprod of every virtual row
    bool lat_on = false;     // the run in progress records them
    bool lat_valid = false;  // the last batch run recorded them
    spg::DevBuf lat_t, rprod, vprod;
    // captured batch run
    bool graph_lat = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_learned = false;
    const spg_predictor* graph_pred = nullptr;
    long long graph_pred_version = -1;
    spg_learned_stats graph_stats{};
};

#ifndef SPG_IGROUP_LOG2
#define SPG_IGROUP_LOG2 6
#endif
namespace spg {
constexpr int kIGroupLog2 = SPG_IGROUP_LOG2;  // products per group (log2)
constexpr long long kIGroup = 1LL << kIGroupLog2;
// Word index of intermediate product s in the grouped layout (its column in
// inter_cols, its value at the same index of inter_vals).
__host__ __device__ __forceinline__ long long islot(long long s) {
    return ((s >> kIGroupLog2) << (kIGroupLog2 + 1)) + (s & (kIGroup - 1));
}
// Words of the grouped buffer for np products (whole groups, one spare
// group: 16-byte loads of a row's last group stay inside).
inline size_t inter_words(long long np) {
    return (size_t)((np + kIGroup - 1) / kIGroup + 1) * (size_t)(2 * kIGroup);
}
inline int* inter_cols(spg_context* c) { return (int*)c->inter.p; }
inline float* inter_vals(spg_context* c) { return (float*)c->inter.p + kIGroup; }
// Narrow virtual rows (width <= the tile width W, so every column lies
// within 2^13 of the row's origin lo) keep their products' columns as the
// low 16 bits, product s at element s of a separate array placed after the
// grouped words (64 products per 128-byte line); their values stay in the
// grouped layout and the column half of their groups is never touched.  The
// column is lo + ((c16 - lo) & 0xFFFF).  Elements past nprod pad the 8-byte
// loads of a row's last four products.
inline size_t inter16_elems(long long np) { return (size_t)np + 256; }
inline size_t inter_bytes(long long np) {
    return sizeof(int32_t) * inter_words(np) + sizeof(uint16_t) * inter16_elems(np);
}
inline uint16_t* inter_cols16(spg_context* c) {
    return (uint16_t*)((int32_t*)c->inter.p + inter_words(c->nprod));
}
}  // namespace spg

namespace spg {

constexpr size_t kReadbackMax = 4096;  // bytes of one host readback

// Device memory helpers (spg_api.hip).
int ensure(spg_context* c, DevBuf& b, size_t bytes);
void* scratch_alloc(spg_context* c, size_t bytes);
void scratch_reset(spg_context* c);
int fail(spg_context* c, int code, const std::string& msg);
// Records the start (end = 0) or end (end = 1) event of a phase timer.
void tmark(spg_context* c, int phase, int end);
// Copies bytes (<= 4096) from device memory to h, synchronizing the
// context stream; recorded into / replayed from the context's replay log.
int readback(spg_context* c, const void* d, size_t bytes, void* h);
// Drops a captured batch graph (operands, config or predictor changed).
void graph_release(spg_context* c);
// Forgets every device buffer of the context without freeing it (the
// allocation-free mode's buffers all live in its arena).
void forget_buffers(spg_context* c);
// Fill kernel (used instead of hipMemsetAsync so the path stays capturable).
template <typename T>
int fill(spg_context* c, T* p, T v, long long n);

// ---- launch wrappers (spg_device.hip); all enqueue on c->stream ----
// exclusive scan: out[0..n] with out[n] = total
int scan_i32_to_i64(spg_context* c, const int32_t* in, int64_t* out, int64_t n);
int scan_i64(spg_context* c, const int64_t* in, int64_t* out, int64_t n);
int scan_i32(spg_context* c, const int32_t* in, int32_t* out, int64_t n);
// stable partition of n keys in [-1, nb) (-1 dropped) into perm; offsets
// (nb+1 entries, device) receive bucket starts.
int stable_partition(spg_context* c, const int32_t* key, int64_t n, int nb,
                     int32_t* perm, int64_t* offsets);

int launch_coo2cm(spg_context* c, const triplet* d_trip, int nnz, int type,
                  int nmajor, int* ptr, int* idx, float* val);
int launch_init(spg_context* c);
// Largest vrow bound, largest untiled-row bound, longest B row and nprod
// into d_out[0..3] (after launch_init_vrows).
int launch_limits(spg_context* c, unsigned long long* d_out);
size_t transpose_scratch(const spg_context* c, int m, long long nnz);  // launch_init's transpose
int launch_categorize(spg_context* c);
int launch_expand_binned(spg_context* c);
int launch_expand_simple(spg_context* c);
int launch_expand_rows(spg_context* c);
int launch_symbolic(spg_context* c);
int launch_merge(spg_context* c);
int launch_learned_merge(spg_context* c, spg_learned_stats* st);
int launch_rowptr(spg_context* c);
// Runs the MLP over c's row features into d_pred (spg_predict.hip).
int predictor_run(spg_context* c, const spg_predictor* p, float scale, float* d_pred);
int predictor_device(const spg_predictor* p);
long long predictor_version(const spg_predictor* p);
int launch_regroup_intermediate(spg_context* c, int64_t* d_rowptr);
// Batch completion times: clears the per-product times and stamps the
// start; the product of every virtual row (after InitGEMM).
int launch_latency_start(spg_context* c);
int launch_vprod(spg_context* c);

}  // namespace spg

#define SPG_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return ::spg::fail(c, SPG_EHIP,                                  \
                               std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

#define SPG_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != SPG_OK) return r_; \
    } while (0)
