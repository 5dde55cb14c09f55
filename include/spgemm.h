/*
 * spgemm.h - C-ABI of the MI355X outer-product SpGEMM (libspgemm_amd.so).
 *
 * Drop-in boundary for the reference fast-squid/spGEMM hot path.  Every
 * declaration names the reference interface it replaces.  The reference's
 * boundary is lib/mm.h:17-25 (GPU phases), lib/coo.h:35-42 (COO container
 * and loaders) and lib/cm.h:23-31 (compressed-matrix container), all
 * extern "C".
 *
 * Differences from the reference, all deliberate:
 *  - GPU phases take an explicit context (no file-scope device globals,
 *    src/mm.cu:19-41) and return an int status instead of calling exit(-1)
 *    (src/mm.cu:6-14).
 *  - C is returned to the caller (the reference never hands it back,
 *    src/mm.cu:152-192,420-449) with 64-bit row offsets (nnz(C) and the
 *    intermediate count can exceed 2^31-1, src/mm.cu:27-28,173-180).
 *  - Values are real FP32 sums (src/mm.cu:33,368,386 truncated to int).
 *
 * No HIP or torch types appear in the signatures: streams are void*, device
 * arrays are plain pointers.
 */
#ifndef SPGEMM_AMD_H_
#define SPGEMM_AMD_H_

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lib/coo.h:13-14 */
#define ROW_MAJOR 0
#define COL_MAJOR 1

/* lib/coo.h:16-22: 12-byte AoS triplet, same layout. */
typedef struct triplet {
    int ridx;
    int cidx;
    float val;
} triplet;

/* lib/coo.h:24-33: same layout. */
typedef struct coo {
    triplet* data;
    int nnz;
    int num_cols;
    int num_rows;
    bool type;
} coo;

/* lib/cm.h:10-21: same layout (int32 ptr, for the operands). */
typedef struct CompressedMatrix {
    int* ptr;
    int* idx;
    float* val;
    int nnz;
    int num_rows;
    int num_cols;
    bool type;
} cm;

/* Result matrix C in CSR with 64-bit row offsets (replaces the int32
 * c_ptr_base of src/mm.cu:28).  Host arrays owned by the caller, released
 * with spgFreeCSR64. */
typedef struct spg_csr64 {
    int64_t* ptr;
    int32_t* idx;
    float* val;
    int64_t nnz;
    int32_t num_rows;
    int32_t num_cols;
} spg_csr64;

/* Status codes (the reference has none: src/mm.cu:6-14 exits). */
#define SPG_OK 0
#define SPG_EINVAL 1          /* bad argument or shape mismatch            */
#define SPG_ENOMEM 2          /* device or host allocation failed          */
#define SPG_EHIP 3            /* HIP runtime error                         */
#define SPG_EORDER 4          /* phase called before its prerequisites     */
#define SPG_EIO 5             /* file could not be read or parsed          */
#define SPG_ENODEV 6          /* no usable GPU                             */
#define SPG_EOVERFLOW 7       /* a size exceeds a representable range      */

/* Operand slots. */
#define SPG_OPERAND_A 0   /* left operand, held as CSC (COL_MAJOR)  */
#define SPG_OPERAND_B 1   /* right operand, held as CSR (ROW_MAJOR) */

/* Number of k-bins of the categorize phase: bin 0 = heavy
 * (a_len*b_len > 50*floor(nprod/K)), bins 1..7 = b_len in [2^(b-1), 2^b),
 * bin 8 = b_len >= 128.  Every k with a non-empty B row is binned. */
#define SPG_NUM_KBINS 9

/* Phase timer slots for spgGetPhaseTimes. */
#define SPG_PHASE_INIT 0
#define SPG_PHASE_CATEGORIZE 1
#define SPG_PHASE_EXPAND 2
#define SPG_PHASE_SYMBOLIC 3
#define SPG_PHASE_MERGE 4
#define SPG_PHASE_PREDICT 5   /* learned path: inference + capacities       */
#define SPG_PHASE_COMPACT 6   /* learned path: padded rows -> CSR           */
#define SPG_NUM_PHASES 7

/* readMTX flags. */
#define SPG_MTX_COMPAT_VALUES 1  /* every value := 1.1 (src/coo.cpp:151-168) */

/* ---------------- host containers (lib/coo.h:35-42, lib/cm.h:23-31) ------ */

/* lib/coo.h:35 / src/coo.cpp:6-13 */
void initCOO(coo* mat);
/* lib/coo.h:36 / src/coo.cpp:15-22 (releases with the matching
 * allocator, defect D13 fixed) */
void freeCOO(coo* mat);
/* lib/coo.h:37 / src/coo.cpp:38-41 */
void setCOOtype(coo* mat, bool type);
/* lib/coo.h:38 / src/coo.cpp:24-36 (also copies type, D13) */
coo copyCOO(coo mat);
/* lib/coo.h:39 / src/coo.cpp:73-79: sort by (major, minor, value); major
 * is the row for ROW_MAJOR, the column for COL_MAJOR. */
void sortCOO(coo mat);
/* lib/coo.h:41 / src/coo.cpp:82-108: "dim n nnz" header then triplets. */
void readCOO(coo* mat, char* path);
/* lib/coo.h:42 / src/coo.cpp:110-173, with defect D11 fixed: Matrix
 * Market size line read as rows cols nnz, values kept, pattern files,
 * symmetric / skew-symmetric / hermitian expansion, any line length.  On
 * error mat is left empty (nnz 0). */
void readMTX(coo* mat, char* path);
/* readMTX with flags (SPG_MTX_COMPAT_VALUES) and a status. */
int spgReadMTX(coo* mat, const char* path, int flags);
/* Writes C as a Matrix Market "coordinate real general" file. */
int spgWriteMTX(const spg_csr64* C, const char* path);

/* lib/cm.h:23-31 / src/cm.cpp */
void cmInit(cm* mat);
void cmFree(cm* mat);
int cmGetNumRows(cm mat);
int cmGetNumCols(cm mat);
int cmGetNNZ(cm mat);
void cmSetNumRows(cm* mat, int num_rows);
void cmSetNumCols(cm* mat, int num_cols);
void cmSetNNZ(cm* mat, int nnz);
void cmSetType(cm* mat, bool type);
void spgFreeCSR64(spg_csr64* mat);

/* ---------------- seeded synthetic inputs (no reference counterpart) ---- */

/* n_rows x n_cols, exactly nnz_per_row distinct columns per row drawn
 * uniformly without replacement, values uniform in [-1, 1]. */
int spgGenUniformRows(coo* out, int n_rows, int n_cols, int nnz_per_row,
                      uint64_t seed);
/* R-MAT with 2^row_bits rows and 2^col_bits columns, n_draws edges drawn
 * with quadrant probabilities (a, b, c, 1-a-b-c), duplicates merged, values
 * uniform in [-1, 1]; rows/cols >= clip_rows/clip_cols are rejected when
 * clip_* > 0 (for non power-of-two sizes). */
int spgGenRMAT(coo* out, int row_bits, int col_bits, int64_t n_draws,
               double a, double b, double c, int clip_rows, int clip_cols,
               uint64_t seed);
/* 27-point stencil on an nx*ny*nz grid: diagonal 26, off-diagonals
 * -1 + uniform[-0.01, 0.01]. */
int spgGenStencil27(coo* out, int nx, int ny, int nz, uint64_t seed);

/* ---------------- GPU context ------------------------------------------- */

typedef struct spg_context spg_context;

/* Tunables (the reference's compile-time DENSE_NUM/NUM_BINS/heavy factor,
 * src/mm.cu:17,205,262, become runtime settings). */
typedef struct spg_config {
    int32_t tile_log2;         /* column tile width of long rows = 2^tile_log2 */
    int32_t item_products;     /* target products per expansion work item      */
    int64_t long_row_ub;       /* rows with more products are column-tiled     */
    int32_t xcd_swizzle;       /* 1: map consecutive work items to one XCD     */
    int32_t timing;            /* 1: record per-phase HIP events               */
    int32_t expand_mode;       /* SPG_EXPAND_ROWS (default) or SPG_EXPAND_KBINS */
    float pred_margin;         /* learned path: capacity = pred*(1+margin)+slack */
    int32_t pred_slack;        /* learned path: slots added to every capacity    */
    int32_t pred_tight;        /* learned path: 0 = capacity is the merge class's
                                  maximum (fewest overflows), 1 = capacity is the
                                  inflated prediction (least memory)            */
    int32_t heavy_factor;      /* k is heavy (bin 0) when a_len*b_len >
                                  heavy_factor * floor(nprod / K); the
                                  reference's constant is 50 (src/mm.cu:205)   */
    int32_t rank_min_count;    /* wide rows with more distinct columns than this
                                  merge through a bitmap of their window (rank),
                                  smaller ones through an LDS hash               */
    int32_t transpose_mode;    /* InitGEMM's A-by-rows transpose: SPG_TRANSPOSE_SORT
                                  (counting sort in LDS, no global atomics; more
                                  than 2^26 rows use the atomic one),
                                  SPG_TRANSPOSE_ATOMIC (one atomic per entry, as
                                  src/mm.cu:138-148 counts rows) or
                                  SPG_TRANSPOSE_AUTO (the sort from 2^20 entries
                                  on, the atomic one below: fewer launches)     */
    int32_t rowchunk_ub;       /* SPG_EXPAND_KBINS: tiled rows whose upper bound
                                  is below this (0: none) are expanded row by
                                  row, in chunks of entries, instead of by k
                                  (as SPG_EXPAND_ROWS expands every tiled row):
                                  their tiles are filled front to back           */
} spg_config;

#define SPG_TRANSPOSE_AUTO 0
#define SPG_TRANSPOSE_ATOMIC 1
#define SPG_TRANSPOSE_SORT 2

/* Expansion mappings of BinGEMM.  KBINS groups the (i,k) pairs by k as the
 * reference does (one work item per k-chunk, k binned by nnz(B row k);
 * src/mm.cu:288-327); ROWS groups them by C row: A's entries in row order,
 * in chunks (tiled rows: one run of slots per chunk and tile, reserved with
 * one atomic; untiled rows: slots known from the row bases), so that every
 * store writes consecutive slots.  The order of the products inside a row
 * is not fixed in either mapping (slot reservations follow the schedule). */
#define SPG_EXPAND_KBINS 0
#define SPG_EXPAND_ROWS 1

int spgCreate(spg_context** ctx, int device);
int spgDestroy(spg_context* ctx);
/* All work of the context goes to this stream (hipStream_t as void*). */
int spgSetStream(spg_context* ctx, void* stream);
int spgGetConfig(spg_context* ctx, spg_config* cfg);
int spgSetConfig(spg_context* ctx, const spg_config* cfg);
const char* spgGetErrorString(int status);
/* Message of the last failure seen by this context. */
const char* spgGetLastError(spg_context* ctx);
int spgSynchronize(spg_context* ctx);
int spgDeviceCount(int* count);

/* lib/mm.h:17 cudaCOO2CM / src/mm.cu:65-136.  Uploads a COO sorted by
 * sortCOO.  mat.type COL_MAJOR fills operand A as CSC (idx = row),
 * ROW_MAJOR fills operand B as CSR (idx = col); the prefix sum runs on the
 * GPU.  If out is not NULL it receives host copies (free with cmFree). */
int spgCOO2CM(spg_context* ctx, coo mat, cm* out);
/* Upload an already compressed host matrix into an operand slot
 * (A: CSC, B: CSR). */
int spgSetOperand(spg_context* ctx, int which, const cm* mat);
/* Borrow device arrays as an operand (no copy; caller keeps ownership and
 * keeps them alive until the context is destroyed or the slot is reset). */
int spgSetOperandDevice(spg_context* ctx, int which, const int* d_ptr,
                        const int* d_idx, const float* d_val, int num_rows,
                        int num_cols, int nnz);

/* lib/mm.h:18 cudaInitGEMM / src/mm.cu:138-193.  Upper bound of every C
 * row (sum of nnz(B row k) over A(i,k)), long-row column tiling, slot
 * bases and the intermediate buffers.  C (may be NULL) receives the
 * shape, type ROW_MAJOR and nnz = nprod clamped to INT_MAX as the
 * reference reports it; nprod (may be NULL) receives the exact count. */
int spgInitGEMM(spg_context* ctx, cm* C, int64_t* nprod);
/* lib/mm.h:19 cudaCategorizeGEMM / src/mm.cu:196-286: k-bins and the
 * expansion work table.  bin_counts (may be NULL) has SPG_NUM_KBINS. */
int spgCategorizeGEMM(spg_context* ctx, int64_t* bin_counts);
/* lib/mm.h:20 cudaSimpleGEMM / src/mm.cu:332-364: expansion with one
 * mapping for every k (the plain method). */
int spgSimpleGEMM(spg_context* ctx);
/* lib/mm.h:21 cudaBinGEMM / src/mm.cu:288-327: expansion with per-bin
 * mappings. */
int spgBinGEMM(spg_context* ctx);
/* Exact symbolic count (no reference counterpart): distinct columns of
 * every C row; sizes C.  nnz (may be NULL) receives nnz(C). */
int spgSymbolicGEMM(spg_context* ctx, int64_t* nnz);
/* lib/mm.h:25 cudaMergeGEMM / src/mm.cu:366-450: merges the intermediate
 * products of every row into sorted CSR on the GPU (no per-row copies). */
int spgMergeGEMM(spg_context* ctx);
/* Copies C to the host (one bulk copy per array). */
int spgGetC(spg_context* ctx, spg_csr64* out);
/* Device pointers of C (valid until the next product on the context). */
int spgGetCDevice(spg_context* ctx, const int64_t** ptr, const int32_t** idx,
                  const float** val, int64_t* nnz);
/* Runs InitGEMM, CategorizeGEMM, BinGEMM, SymbolicGEMM, MergeGEMM. */
int spgMultiply(spg_context* ctx);

/* Inspection for tests and tools. */
int spgGetUpperBound(spg_context* ctx, int64_t* ub);
int spgGetKBins(spg_context* ctx, int32_t* bin_of_k, int32_t* perm,
                int64_t* offsets);
int spgGetRowCounts(spg_context* ctx, int64_t* cnt);
/* Intermediate products regrouped per C row: row_ptr (m+1 entries),
 * cols/vals (nprod entries) in an unspecified order inside a row. */
int spgGetIntermediateByRow(spg_context* ctx, int64_t* row_ptr,
                            int32_t* cols, float* vals);
/* Milliseconds per phase (SPG_NUM_PHASES) of the last product, from HIP
 * events on the context stream; requires cfg.timing. */
int spgGetPhaseTimes(spg_context* ctx, double* ms);
/* The 32-bit limits of the expansion (defect D10 of src/mm.cu:173-180, whose
 * int offsets wrap silently): SPG_OK when the largest virtual-row bound and
 * 256 times min(longest B row, largest untiled-row bound) both stay below
 * 2^31 - 64, else SPG_EOVERFLOW.  InitGEMM applies it to the maxima it
 * reduces on the GPU and fails with SPG_EOVERFLOW before any expansion. */
int spgSlotLimitStatus(int64_t max_vrow_bound, int64_t max_untiled_bound,
                       int64_t max_b_row);

/* ---------------- allocation-free product (no reference counterpart) ----
 * For callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc): spgMultiply's phases on the caller's
 * device arrays, with every intermediate buffer taken from a fixed device
 * array of SPG_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds
 * until the process ends.  No hipMalloc, hipFree or pinned allocation is made
 * (a failing call may build its message on the host heap).  A (m x K) is CSC
 * (a_ptr has K+1 entries), B (K x n) is CSR (b_ptr has K+1 entries, columns
 * ascending in each row).  C is written to c_ptr (m+1 int64), c_idx and c_val
 * (c_capacity entries each); *c_nnz receives nnz(C).  Returns SPG_EOVERFLOW
 * with *c_nnz = the required capacity when C does not fit (c_ptr then
 * undefined), SPG_ENOMEM when the product needs more scratch than the
 * maximum (roughly 8 bytes per intermediate product plus 1 KB per virtual
 * row and 16 bytes per entry of A), and returns after the work has finished
 * on stream (hipStream_t as void*, may be NULL).  Calls on one GPU share its
 * scratch and race when they overlap (several host threads, or streams).
 * err (may be NULL) receives the failure message. */
#define SPG_NOALLOC_SCRATCH_BYTES (256u << 20)
int spgMultiplyNoAlloc(int device, void* stream, const int* a_ptr, const int* a_idx,
                       const float* a_val, int m, int K, int nnz_a, const int* b_ptr,
                       const int* b_idx, const float* b_val, int n, int nnz_b,
                       int64_t* c_ptr, int32_t* c_idx, float* c_val, int64_t c_capacity,
                       int64_t* c_nnz, char* err, size_t err_len);

/* ---------------- learned allocation (no reference counterpart) ---------
 * A 4-64-64-1 FP32 MLP predicts nnz(C row i) from four features of the
 * row: log1p of nnz(A row i), of ub_i, of max_k nnz(B row k) and of the
 * standard deviation of nnz(B row k) over the row's k.  The learned merge
 * sizes every row's output from the prediction instead of the symbolic
 * pass, detects rows that did not fit, merges those again with their upper
 * bound as capacity (exact fallback), and compacts the rows into CSR.  The
 * result is identical to spgMergeGEMM's. */
typedef struct spg_predictor spg_predictor;

typedef struct spg_learned_stats {
    int64_t vrows;           /* virtual rows with at least one product        */
    int64_t overflow_vrows;  /* rows whose prediction was too small (re-merged) */
    int64_t unrecovered;     /* rows still not fitting after the fallback (0)  */
    int64_t padded_slots;    /* sum of predicted capacities                    */
    int64_t fallback_slots;  /* sum of fallback capacities                     */
    int64_t nnz;             /* nnz(C)                                         */
} spg_learned_stats;

int spgPredictorCreate(spg_predictor** p, int device, uint64_t seed);
int spgPredictorDestroy(spg_predictor* p);
const char* spgPredictorGetLastError(spg_predictor* p);
int spgPredictorNumParams(void);
/* Adds the non-empty rows of ctx's last product as training samples
 * (features and exact counts); needs SymbolicGEMM or a merge on ctx. */
int spgPredictorAddSamples(spg_predictor* p, spg_context* ctx);
/* Adam on the mean squared error of y against log1p(cnt) - log1p(s), with
 * s = n * (1 - exp(-ub / n)) the count of ub uniform products, over all
 * samples; batch is rounded down to a multiple of 64.  Deterministic.  The
 * feature normalization is fixed from the samples at the first call. */
int spgPredictorTrain(spg_predictor* p, int epochs, int batch, float lr,
                      float* final_loss);
/* Mean squared error over all samples and its gradient (NumParams floats)
 * at the current weights; no update.  For checking the training kernels. */
int spgPredictorGradient(spg_predictor* p, float* grad, float* loss);
/* w: spgPredictorNumParams() floats; norm8: feature means then sds. */
int spgPredictorGetWeights(spg_predictor* p, float* w, float* norm8);
int spgPredictorSetWeights(spg_predictor* p, const float* w, const float* norm8);
int spgPredictorSave(spg_predictor* p, const char* path);
int spgPredictorLoad(spg_predictor* p, const char* path);

/* Predicted nnz of every C row of ctx's product (after InitGEMM); kept on
 * the context for spgLearnedMergeGEMM.  pred (may be NULL) receives m floats. */
int spgPredictRows(spg_context* ctx, const spg_predictor* p, float* pred);
/* Replaces the context's row predictions with m caller values (testing and
 * external predictors). */
int spgSetRowPredictions(spg_context* ctx, const float* pred);
/* Merge with predicted capacities (after BinGEMM/SimpleGEMM and
 * spgPredictRows/spgSetRowPredictions); st may be NULL. */
int spgLearnedMergeGEMM(spg_context* ctx, spg_learned_stats* st);
/* InitGEMM, CategorizeGEMM, BinGEMM, PredictRows, LearnedMergeGEMM. */
int spgMultiplyLearned(spg_context* ctx, const spg_predictor* p,
                       spg_learned_stats* st);

/* ---------------- batched products (no reference counterpart) -----------
 * count independent products C_b = A_b * B_b run as one: the operands are
 * stacked block-diagonally (product b's columns start at a multiple of
 * 8192) and every phase kernel runs once over the work of all products. */
#define SPG_BATCH_GRAPH 1  /* record the run once, then replay it as one HIP graph */
#define SPG_BATCH_LATENCY 2  /* record every product's completion time on the GPU */

/* A[b]: CSC (COL_MAJOR), B[b]: CSR (ROW_MAJOR), host arrays (copied). */
int spgBatchSetOperands(spg_context* ctx, int count, const cm* A, const cm* B);
/* row_off (count+1 entries): product b's C rows in the stacked C;
 * col_off (count): its column offset there.  Either may be NULL. */
int spgBatchGetLayout(spg_context* ctx, int64_t* row_off, int64_t* col_off);
/* Plain path (p == NULL) or learned path over the whole batch.  With
 * SPG_BATCH_GRAPH the first call runs normally while recording its host
 * readbacks and captures the same sequence into a HIP graph; later calls
 * with the same operands, config and predictor (weights unchanged) replay
 * it with one hipGraphLaunch and no host synchronization (call
 * spgSynchronize before reading C).  st (learned path; may be NULL). */
int spgBatchMultiply(spg_context* ctx, const spg_predictor* p, int flags,
                     spg_learned_stats* st);
/* Host copy of product b's C (columns relative to its block). */
int spgBatchGetC(spg_context* ctx, int b, spg_csr64* out);
/* After a spgBatchMultiply with SPG_BATCH_LATENCY: ms[b] (count entries) =
 * milliseconds from the start of the run to the moment the GPU wrote
 * product b's last C row (device wall clock; the kernel writing a row's
 * final entries stamps its product), 0 for a product without entries.
 * Synchronizes the context stream. */
int spgBatchGetLatency(spg_context* ctx, double* ms);

/* ---------------- autotuner (no reference counterpart) ------------------
 * Coordinate descent over tile_log2, long_row_ub, item_products,
 * expand_mode and heavy_factor on the operands set on ctx (one product or a
 * batch; the learned path when p != NULL): each candidate is timed as the
 * median wall time of reps full runs, a candidate replaces the incumbent
 * when it is more than 1 % faster, and seed fixes the order in which the
 * tunables are visited (two sweeps).  The winner is set on ctx and, if
 * table_path is not NULL, appended to that table as one JSON line keyed by
 * the workload's shape.  best / best_ms may be NULL. */
int spgTune(spg_context* ctx, const spg_predictor* p, int reps, uint64_t seed,
            const char* table_path, spg_config* best, double* best_ms);
/* Applies the table's last entry for the shape of the operands on ctx (and
 * the path: learned != 0), if any; *found (may be NULL) tells. */
int spgTuneLookup(spg_context* ctx, const char* table_path, int learned, int* found);

#ifdef __cplusplus
}
#endif
#endif
