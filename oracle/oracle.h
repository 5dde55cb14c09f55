/*
 * oracle.h - CPU oracle for the outer-product SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (spgemm_amd/) links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the correctness
 * reference or as the timed CPU baseline.
 *
 * The functions restate the reference method (fast-squid/spGEMM) in plain C:
 *   src/coo.cpp:43-79    COO sort comparators and sortCOO
 *   src/mm.cu:43-136     COO -> CSC/CSR conversion (coo2cm + host prefix)
 *   src/mm.cu:138-193    per-row upper bound (initGEMM) and slot bases
 *   src/mm.cu:196-286    k-binning (inspectGEMM / categorizeGEMM)
 *   src/mm.cu:288-364    outer-product expansion (binGEMM / simpleGEMM)
 *   src/mm.cu:366-450    per-row merge and compaction (mergeGEMM)
 * with the reference defects D1-D10 fixed: products are FP32 (a*b rounded to
 * float), accumulation of a column is in FP64 and the sum is rounded to FP32
 * once, first touch is by occupancy (never by value), bins cover every k with
 * a non-empty B row, every offset and count is 64-bit.
 *
 * An independent textbook row-wise Gustavson product (FP64 products and FP64
 * sums) is the FP64 value reference.
 */
#ifndef SPGEMM_ORACLE_H_
#define SPGEMM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 12-byte layout as the reference's triplet (lib/coo.h:16-22). */
typedef struct orc_triplet {
    int32_t ridx;
    int32_t cidx;
    float val;
} orc_triplet;

#define ORC_ROW_MAJOR 0
#define ORC_COL_MAJOR 1

/* Number of k-bins: bin 0 heavy, bins 1..7 for b_len in [2^(b-1), 2^b),
 * bin 8 for b_len >= 128 (the reference never bins these, defect D6). */
#define ORC_NUM_KBINS 9
#define ORC_HEAVY_FACTOR 50

/* Fault injection used by the oracle self-tests to prove that the pinning
 * checks catch plausible mistakes.  0 = no fault. */
#define ORC_FAULT_NONE 0
#define ORC_FAULT_DROP_LAST_K 1     /* merge ignores the last k of each row   */
#define ORC_FAULT_TRANSPOSE_B 2     /* uses B's column index as its row index */
#define ORC_FAULT_OFF_BY_ONE_UB 3   /* upper bound counts b_len(k)+1          */
#define ORC_FAULT_VALUE_SIGN 4      /* negates the product a*b                */
#define ORC_FAULT_SORT_MINOR 5      /* sort_coo orders by minor before major  */
#define ORC_FAULT_PTR_UNSHIFTED 6   /* coo2cm counts major i at ptr[i], not i+1 */
#define ORC_FAULT_CSR_ROW_IDX 7     /* csc_to_csr stores the row, not k       */
void orc_set_fault(int fault);
int orc_get_fault(void);

/* src/coo.cpp:43-79: qsort by (major, minor, val). type selects the major. */
void orc_sort_coo(int64_t nnz, orc_triplet* t, int type);

/* src/mm.cu:43-63 + 100-108: sorted COO -> compressed (CSR if ROW_MAJOR,
 * CSC if COL_MAJOR).  ptr has nmajor+1 entries. */
void orc_coo2cm(int64_t nnz, const orc_triplet* t, int type, int32_t nmajor,
                int32_t* ptr, int32_t* idx, float* val);

/* CSC(A) -> CSR(A), stable (k ascending inside each row). */
void orc_csc_to_csr(int32_t m, int32_t K, const int32_t* a_ptr,
                    const int32_t* a_idx, const float* a_val,
                    int32_t* r_ptr, int32_t* r_idx, float* r_val);

/* src/mm.cu:138-148,173-180: ub[i] = sum over A(i,k) of nnz(B row k);
 * base = exclusive scan (m+1 entries).  Returns nprod = base[m]. */
int64_t orc_upper_bound(int32_t m, int32_t K, const int32_t* a_ptr,
                        const int32_t* a_idx, const int32_t* b_ptr,
                        int64_t* ub, int64_t* base);

/* src/mm.cu:196-257 with D5/D6/D7 fixed.  bin_of_k[k] in 0..8 or -1 when
 * b_len(k) == 0.  perm lists k grouped by bin, k ascending inside a bin;
 * offsets has ORC_NUM_KBINS+1 entries.  Returns the number of binned k. */
int32_t orc_kbins(int32_t K, const int32_t* a_ptr, const int32_t* b_ptr,
                  int64_t nprod, int32_t* bin_of_k, int32_t* perm,
                  int64_t* offsets);

/* src/mm.cu:288-312 / 332-354: expansion into row slots.  Slot order is
 * canonical: inside row i, products are ordered by k ascending, then by the
 * position inside B row k.  icol/ival have nprod entries. */
void orc_expand(int32_t m, int32_t K, const int32_t* a_ptr,
                const int32_t* a_idx, const float* a_val,
                const int32_t* b_ptr, const int32_t* b_idx,
                const float* b_val, const int64_t* base,
                int32_t* icol, float* ival);

/* Exact symbolic count from the expanded slots: cnt[i] = number of distinct
 * columns in row i's slots (numerical zeros kept). Returns sum. */
int64_t orc_symbolic_from_slots(int32_t m, int32_t n, const int64_t* base,
                                const int32_t* icol, int64_t* cnt);

/* src/mm.cu:366-404 + 420-435: per-row merge with occupancy first touch,
 * FP64 accumulation, FP32 result, columns emitted ascending; then CSR.
 * c_ptr has m+1 entries; c_idx/c_val are sized by the caller from
 * orc_symbolic_from_slots. */
void orc_merge(int32_t m, int32_t n, const int64_t* base, const int32_t* icol,
               const float* ival, int64_t* c_ptr, int32_t* c_idx,
               float* c_val);

/* Whole restated method in one call (ub -> expand -> symbolic -> merge).
 * Allocates C (malloc); the caller frees with orc_free. Returns nnz(C). */
int64_t orc_spgemm_restated(int32_t m, int32_t n, int32_t K,
                            const int32_t* a_ptr, const int32_t* a_idx,
                            const float* a_val, const int32_t* b_ptr,
                            const int32_t* b_idx, const float* b_val,
                            int64_t** c_ptr, int32_t** c_idx, float** c_val,
                            int64_t* nprod_out);

/* Textbook row-wise Gustavson (independent of the method above) on A in
 * CSR.  prod32 = 0: FP64 products a*b and FP64 sums (the FP64 reference);
 * prod32 = 1: FP32 products then FP64 sums rounded to FP32 (must equal the
 * restated method).  rows == NULL means every row; otherwise only the
 * listed rows, and C is the concatenation of those rows.  Values are
 * returned in c_val64 (always) and c_sabs (sum of |a*b| per entry, may be
 * NULL).  Returns nnz. */
int64_t orc_gustavson(int32_t m, int32_t n, const int32_t* r_ptr,
                      const int32_t* r_idx, const float* r_val,
                      const int32_t* b_ptr, const int32_t* b_idx,
                      const float* b_val, const int32_t* rows,
                      int32_t nrows_sel, int prod32, int64_t** c_ptr,
                      int32_t** c_idx, double** c_val64, double** c_sabs);

/* Timed CPU baseline: the restated method (upper bound, bins, k-driven
 * expansion with deterministic slots, per-row merge, compaction) on the
 * row block [row_begin, row_end) with nthreads POSIX threads.  Output C is
 * discarded after a checksum.  times_ms receives 5 phase times
 * (ub, bins, expand, merge, compact).  Returns nprod of the block. */
int64_t orc_cpu_baseline(int32_t m, int32_t n, int32_t K,
                         const int32_t* a_ptr, const int32_t* a_idx,
                         const float* a_val, const int32_t* b_ptr,
                         const int32_t* b_idx, const float* b_val,
                         int32_t row_begin, int32_t row_end, int nthreads,
                         double* times_ms, int64_t* nnz_out,
                         double* checksum_out);

void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
