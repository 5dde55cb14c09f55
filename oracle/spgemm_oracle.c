/*
 * spgemm_oracle.c - CPU oracle for the outer-product SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built by oracle/Makefile into
 * oracle/_build/liboracle.so; loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg.  The product never links it.
 *
 * Build with -ffp-contract=off so that a*b is rounded to float before any
 * addition (the reference computes a_val*b_val in FP32, src/mm.cu:308).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int g_fault = ORC_FAULT_NONE;
void orc_set_fault(int fault) { g_fault = fault; }
int orc_get_fault(void) { return g_fault; }
void orc_free(void* p) { free(p); }

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------ */
/* src/coo.cpp:43-56 (cooCompareByRow) and :58-71 (cooCompareByCol).  The
 * reference never returns 0; equal triplets are indistinguishable, so a
 * total order on (major, minor, val) gives the same sorted array.        */
static int cmp_by_row(const void* a, const void* b) {
    const orc_triplet* x = (const orc_triplet*)a;
    const orc_triplet* y = (const orc_triplet*)b;
    if (x->ridx != y->ridx) return x->ridx < y->ridx ? -1 : 1;
    if (x->cidx != y->cidx) return x->cidx < y->cidx ? -1 : 1;
    if (x->val > y->val) return 1;
    if (x->val < y->val) return -1;
    return 0;
}
static int cmp_by_col(const void* a, const void* b) {
    const orc_triplet* x = (const orc_triplet*)a;
    const orc_triplet* y = (const orc_triplet*)b;
    if (x->cidx != y->cidx) return x->cidx < y->cidx ? -1 : 1;
    if (x->ridx != y->ridx) return x->ridx < y->ridx ? -1 : 1;
    if (x->val > y->val) return 1;
    if (x->val < y->val) return -1;
    return 0;
}

/* src/coo.cpp:73-79 (ORC_FAULT_SORT_MINOR swaps the comparators: minor
 * index first) */
void orc_sort_coo(int64_t nnz, orc_triplet* t, int type) {
    const int by_row = (type == ORC_ROW_MAJOR) != (g_fault == ORC_FAULT_SORT_MINOR);
    qsort(t, (size_t)nnz, sizeof(orc_triplet), by_row ? cmp_by_row : cmp_by_col);
}

/* src/mm.cu:43-63 (coo2cm: idx/val copied in COO order, ptr histogram)
 * and src/mm.cu:100-108,122-130 (host prefix sum). */
void orc_coo2cm(int64_t nnz, const orc_triplet* t, int type, int32_t nmajor,
                int32_t* ptr, int32_t* idx, float* val) {
    memset(ptr, 0, sizeof(int32_t) * ((size_t)nmajor + 1));
    for (int64_t e = 0; e < nnz; ++e) {
        int32_t major = type == ORC_ROW_MAJOR ? t[e].ridx : t[e].cidx;
        int32_t minor = type == ORC_ROW_MAJOR ? t[e].cidx : t[e].ridx;
        idx[e] = minor;
        val[e] = t[e].val;
        ptr[major + (g_fault == ORC_FAULT_PTR_UNSHIFTED ? 0 : 1)] += 1;
    }
    for (int32_t i = 0; i < nmajor; ++i) ptr[i + 1] += ptr[i];
}

void orc_csc_to_csr(int32_t m, int32_t K, const int32_t* a_ptr,
                    const int32_t* a_idx, const float* a_val,
                    int32_t* r_ptr, int32_t* r_idx, float* r_val) {
    memset(r_ptr, 0, sizeof(int32_t) * ((size_t)m + 1));
    for (int32_t e = 0; e < a_ptr[K]; ++e) r_ptr[a_idx[e] + 1] += 1;
    for (int32_t i = 0; i < m; ++i) r_ptr[i + 1] += r_ptr[i];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)m + 1));
    memcpy(cur, r_ptr, sizeof(int32_t) * ((size_t)m + 1));
    for (int32_t k = 0; k < K; ++k) {
        for (int32_t e = a_ptr[k]; e < a_ptr[k + 1]; ++e) {
            int32_t p = cur[a_idx[e]]++;
            r_idx[p] = g_fault == ORC_FAULT_CSR_ROW_IDX ? a_idx[e] : k;
            r_val[p] = a_val[e];
        }
    }
    free(cur);
}

static inline int64_t blen_of(const int32_t* b_ptr, int32_t k) {
    int64_t l = (int64_t)b_ptr[k + 1] - b_ptr[k];
    if (g_fault == ORC_FAULT_OFF_BY_ONE_UB) l += 1;
    return l;
}

/* src/mm.cu:138-148 (ub via atomicAdd per A entry) + :173-180 (scan). */
int64_t orc_upper_bound(int32_t m, int32_t K, const int32_t* a_ptr,
                        const int32_t* a_idx, const int32_t* b_ptr,
                        int64_t* ub, int64_t* base) {
    memset(ub, 0, sizeof(int64_t) * (size_t)m);
    for (int32_t k = 0; k < K; ++k) {
        int64_t bl = blen_of(b_ptr, k);
        for (int32_t e = a_ptr[k]; e < a_ptr[k + 1]; ++e) ub[a_idx[e]] += bl;
    }
    int64_t s = 0;
    for (int32_t i = 0; i < m; ++i) {
        if (base) base[i] = s;
        s += ub[i];
    }
    if (base) base[m] = s;
    return s;
}

/* src/mm.cu:196-257.  The heavy threshold is 50 * floor(nprod / K) in
 * 64-bit (D7); the range is k < K for both passes (D5); b_len >= 128 gets
 * bin 8 instead of being dropped (D6). */
static int32_t kbin_of(int64_t a_len, int64_t b_len, int64_t thresh) {
    if (b_len <= 0) return -1;
    if (a_len * b_len > thresh) return 0;
    if (b_len >= 128) return 8;
    int32_t b = 0;
    while ((1LL << (b + 1)) <= b_len) ++b; /* floor(log2 b_len) */
    return b + 1;
}

int32_t orc_kbins(int32_t K, const int32_t* a_ptr, const int32_t* b_ptr,
                  int64_t nprod, int32_t* bin_of_k, int32_t* perm,
                  int64_t* offsets) {
    int64_t thresh = K > 0 ? (nprod / K) * ORC_HEAVY_FACTOR : 0;
    int64_t cnt[ORC_NUM_KBINS];
    memset(cnt, 0, sizeof(cnt));
    for (int32_t k = 0; k < K; ++k) {
        int32_t b = kbin_of((int64_t)a_ptr[k + 1] - a_ptr[k],
                            (int64_t)b_ptr[k + 1] - b_ptr[k], thresh);
        bin_of_k[k] = b;
        if (b >= 0) cnt[b]++;
    }
    offsets[0] = 0;
    for (int b = 0; b < ORC_NUM_KBINS; ++b) offsets[b + 1] = offsets[b] + cnt[b];
    int64_t cur[ORC_NUM_KBINS];
    for (int b = 0; b < ORC_NUM_KBINS; ++b) cur[b] = offsets[b];
    for (int32_t k = 0; k < K; ++k)
        if (bin_of_k[k] >= 0) perm[cur[bin_of_k[k]]++] = k;
    return (int32_t)offsets[ORC_NUM_KBINS];
}

static inline float prod32(float a, float b) {
    volatile float p = a * b; /* rounded to FP32, never fused */
    float r = p;
    if (g_fault == ORC_FAULT_VALUE_SIGN) r = -r;
    return r;
}

/* src/mm.cu:288-312 / 332-354.  k-driven outer product: for each k, for each
 * A(i,k) in column k, B row k scaled by a is appended to row i's slots.
 * Walking k in ascending order makes the slot order canonical. */
void orc_expand(int32_t m, int32_t K, const int32_t* a_ptr,
                const int32_t* a_idx, const float* a_val,
                const int32_t* b_ptr, const int32_t* b_idx,
                const float* b_val, const int64_t* base,
                int32_t* icol, float* ival) {
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * ((size_t)m + 1));
    memcpy(cur, base, sizeof(int64_t) * (size_t)m);
    for (int32_t k = 0; k < K; ++k) {
        int32_t b0 = b_ptr[k], b1 = b_ptr[k + 1];
        for (int32_t e = a_ptr[k]; e < a_ptr[k + 1]; ++e) {
            int32_t i = a_idx[e];
            float a = a_val[e];
            int64_t s = cur[i];
            for (int32_t j = b0; j < b1; ++j, ++s) {
                int32_t col = b_idx[j];
                if (g_fault == ORC_FAULT_TRANSPOSE_B) col = k;
                icol[s] = col;
                ival[s] = prod32(a, b_val[j]);
            }
            cur[i] = s;
        }
    }
    free(cur);
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return x < y ? -1 : (x > y);
}

int64_t orc_symbolic_from_slots(int32_t m, int32_t n, const int64_t* base,
                                const int32_t* icol, int64_t* cnt) {
    unsigned char* occ = (unsigned char*)calloc((size_t)n + 1, 1);
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int64_t total = 0;
    for (int32_t i = 0; i < m; ++i) {
        int64_t c = 0;
        for (int64_t s = base[i]; s < base[i + 1]; ++s) {
            int32_t col = icol[s];
            if (!occ[col]) { occ[col] = 1; list[c++] = col; }
        }
        for (int64_t t = 0; t < c; ++t) occ[list[t]] = 0;
        cnt[i] = c;
        total += c;
    }
    free(occ);
    free(list);
    return total;
}

/* Merge of one row's slots [s0, s1): occupancy first touch (fixes D2),
 * FP64 accumulation (fixes D1), ascending emit.  Returns the count. */
static int64_t merge_row(int64_t s0, int64_t s1, const int32_t* icol,
                         const float* ival, double* acc, unsigned char* occ,
                         int32_t* list, int32_t* out_idx, float* out_val,
                         int64_t skip_from) {
    int64_t c = 0;
    for (int64_t s = s0; s < s1; ++s) {
        if (s >= skip_from) break;
        int32_t col = icol[s];
        if (!occ[col]) { occ[col] = 1; acc[col] = 0.0; list[c++] = col; }
        acc[col] += (double)ival[s];
    }
    qsort(list, (size_t)c, sizeof(int32_t), cmp_i32);
    for (int64_t t = 0; t < c; ++t) {
        int32_t col = list[t];
        if (out_idx) out_idx[t] = col;
        if (out_val) out_val[t] = (float)acc[col];
        occ[col] = 0;
    }
    return c;
}

/* For ORC_FAULT_DROP_LAST_K: slot where the last k of row i starts. */
static int64_t last_k_start(int64_t s0, int64_t s1, const int32_t* icol) {
    /* products of one k are a run of strictly increasing columns (B rows
     * are sorted); the last run starts after the last descent. */
    int64_t st = s0;
    for (int64_t s = s0 + 1; s < s1; ++s)
        if (icol[s] <= icol[s - 1]) st = s;
    return st;
}

void orc_merge(int32_t m, int32_t n, const int64_t* base, const int32_t* icol,
               const float* ival, int64_t* c_ptr, int32_t* c_idx,
               float* c_val) {
    double* acc = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    unsigned char* occ = (unsigned char*)calloc((size_t)n + 1, 1);
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    c_ptr[0] = 0;
    for (int32_t i = 0; i < m; ++i) {
        int64_t skip = base[i + 1];
        if (g_fault == ORC_FAULT_DROP_LAST_K && base[i + 1] > base[i])
            skip = last_k_start(base[i], base[i + 1], icol);
        int64_t c = merge_row(base[i], base[i + 1], icol, ival, acc, occ, list,
                              c_idx + c_ptr[i], c_val + c_ptr[i], skip);
        c_ptr[i + 1] = c_ptr[i] + c;
    }
    free(acc);
    free(occ);
    free(list);
}

int64_t orc_spgemm_restated(int32_t m, int32_t n, int32_t K,
                            const int32_t* a_ptr, const int32_t* a_idx,
                            const float* a_val, const int32_t* b_ptr,
                            const int32_t* b_idx, const float* b_val,
                            int64_t** c_ptr, int32_t** c_idx, float** c_val,
                            int64_t* nprod_out) {
    int64_t* ub = (int64_t*)malloc(sizeof(int64_t) * ((size_t)m + 1));
    int64_t* base = (int64_t*)malloc(sizeof(int64_t) * ((size_t)m + 1));
    int64_t nprod = orc_upper_bound(m, K, a_ptr, a_idx, b_ptr, ub, base);
    int64_t nprod_reported = nprod;
    if (g_fault == ORC_FAULT_OFF_BY_ONE_UB) {
        /* the faulty bound is reported; the buffers use the true one so the
         * expansion never leaves unwritten slots behind */
        int saved = g_fault;
        g_fault = ORC_FAULT_NONE;
        nprod = orc_upper_bound(m, K, a_ptr, a_idx, b_ptr, ub, base);
        g_fault = saved;
    }
    int32_t* icol = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nprod + 1));
    float* ival = (float*)malloc(sizeof(float) * ((size_t)nprod + 1));
    orc_expand(m, K, a_ptr, a_idx, a_val, b_ptr, b_idx, b_val, base, icol,
               ival);
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * ((size_t)m + 1));
    int64_t nnz = orc_symbolic_from_slots(m, n, base, icol, cnt);
    *c_ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)m + 1));
    *c_idx = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nnz + 1));
    *c_val = (float*)malloc(sizeof(float) * ((size_t)nnz + 1));
    orc_merge(m, n, base, icol, ival, *c_ptr, *c_idx, *c_val);
    nnz = (*c_ptr)[m];
    if (nprod_out) *nprod_out = nprod_reported;
    free(ub);
    free(base);
    free(icol);
    free(ival);
    free(cnt);
    return nnz;
}

/* Textbook Gustavson: row i of C = sum over A(i,k) of A(i,k) * B(k,:). */
int64_t orc_gustavson(int32_t m, int32_t n, const int32_t* r_ptr,
                      const int32_t* r_idx, const float* r_val,
                      const int32_t* b_ptr, const int32_t* b_idx,
                      const float* b_val, const int32_t* rows,
                      int32_t nrows_sel, int prod32_mode, int64_t** c_ptr,
                      int32_t** c_idx, double** c_val64, double** c_sabs) {
    int32_t nr = rows ? nrows_sel : m;
    /* size pass */
    unsigned char* occ = (unsigned char*)calloc((size_t)n + 1, 1);
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int64_t* ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nr + 1));
    ptr[0] = 0;
    for (int32_t q = 0; q < nr; ++q) {
        int32_t i = rows ? rows[q] : q;
        int64_t c = 0;
        for (int32_t e = r_ptr[i]; e < r_ptr[i + 1]; ++e) {
            int32_t k = r_idx[e];
            for (int32_t j = b_ptr[k]; j < b_ptr[k + 1]; ++j) {
                int32_t col = b_idx[j];
                if (!occ[col]) { occ[col] = 1; list[c++] = col; }
            }
        }
        for (int64_t t = 0; t < c; ++t) occ[list[t]] = 0;
        ptr[q + 1] = ptr[q] + c;
    }
    int64_t nnz = ptr[nr];
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nnz + 1));
    double* v64 = (double*)malloc(sizeof(double) * ((size_t)nnz + 1));
    double* sab = c_sabs ? (double*)malloc(sizeof(double) * ((size_t)nnz + 1))
                         : NULL;
    double* acc = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    double* aab = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    for (int32_t q = 0; q < nr; ++q) {
        int32_t i = rows ? rows[q] : q;
        int64_t c = 0;
        for (int32_t e = r_ptr[i]; e < r_ptr[i + 1]; ++e) {
            int32_t k = r_idx[e];
            double a = r_val[e];
            for (int32_t j = b_ptr[k]; j < b_ptr[k + 1]; ++j) {
                int32_t col = b_idx[j];
                if (!occ[col]) {
                    occ[col] = 1;
                    acc[col] = 0.0;
                    aab[col] = 0.0;
                    list[c++] = col;
                }
                double p = prod32_mode ? (double)prod32(r_val[e], b_val[j])
                                       : a * (double)b_val[j];
                acc[col] += p;
                aab[col] += fabs(p);
            }
        }
        qsort(list, (size_t)c, sizeof(int32_t), cmp_i32);
        int64_t o = ptr[q];
        for (int64_t t = 0; t < c; ++t) {
            int32_t col = list[t];
            idx[o + t] = col;
            v64[o + t] = prod32_mode ? (double)(float)acc[col] : acc[col];
            if (sab) sab[o + t] = aab[col];
            occ[col] = 0;
        }
    }
    free(occ);
    free(list);
    free(acc);
    free(aab);
    *c_ptr = ptr;
    *c_idx = idx;
    *c_val64 = v64;
    if (c_sabs) *c_sabs = sab;
    return nnz;
}

/* ------------------------------------------------------------------ */
/* CPU baseline: the restated method, multi-threaded, on a row block.   */

typedef struct {
    int tid, nthreads;
    int32_t m, n, K, r0, r1;
    const int32_t *a_ptr, *a_idx, *b_ptr, *b_idx;
    const float *a_val, *b_val;
    /* per-column sub-range of A entries whose row lies in [r0, r1) */
    int32_t *lo, *hi;
    int64_t* ub;      /* block rows */
    int64_t* base;    /* block rows + 1 */
    int64_t* tcur;    /* nthreads x nrows_block cursor matrix */
    int32_t* icol;
    float* ival;
    int64_t* cnt;
    int64_t* c_ptr;
    int32_t* c_idx;
    float* c_val;
    int phase;
    double checksum;
} bl_args;

static int32_t lower_bound_i32(const int32_t* a, int32_t lo, int32_t hi,
                               int32_t v) {
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void* bl_worker(void* p) {
    bl_args* a = (bl_args*)p;
    int32_t nb = a->r1 - a->r0;
    int32_t K = a->K;
    int32_t k0 = (int32_t)((int64_t)K * a->tid / a->nthreads);
    int32_t k1 = (int32_t)((int64_t)K * (a->tid + 1) / a->nthreads);
    int32_t q0 = (int32_t)((int64_t)nb * a->tid / a->nthreads);
    int32_t q1 = (int32_t)((int64_t)nb * (a->tid + 1) / a->nthreads);
    if (a->phase == 0) {
        /* per-column row range + per-thread per-row counts (upper bound) */
        int64_t* mine = a->tcur + (int64_t)a->tid * nb;
        memset(mine, 0, sizeof(int64_t) * (size_t)nb);
        for (int32_t k = k0; k < k1; ++k) {
            int32_t lo = lower_bound_i32(a->a_idx, a->a_ptr[k], a->a_ptr[k + 1], a->r0);
            int32_t hi = lower_bound_i32(a->a_idx, lo, a->a_ptr[k + 1], a->r1);
            a->lo[k] = lo;
            a->hi[k] = hi;
            int64_t bl = (int64_t)a->b_ptr[k + 1] - a->b_ptr[k];
            for (int32_t e = lo; e < hi; ++e) mine[a->a_idx[e] - a->r0] += bl;
        }
    } else if (a->phase == 1) {
        /* expansion over this thread's k range at deterministic slots */
        int64_t* cur = a->tcur + (int64_t)a->tid * nb;
        for (int32_t k = k0; k < k1; ++k) {
            int32_t b0 = a->b_ptr[k], b1 = a->b_ptr[k + 1];
            for (int32_t e = a->lo[k]; e < a->hi[k]; ++e) {
                int32_t r = a->a_idx[e] - a->r0;
                float av = a->a_val[e];
                int64_t s = cur[r];
                for (int32_t j = b0; j < b1; ++j, ++s) {
                    a->icol[s] = a->b_idx[j];
                    a->ival[s] = prod32(av, a->b_val[j]);
                }
                cur[r] = s;
            }
        }
    } else if (a->phase == 2) {
        /* per-row merge into row-local storage at slot bases (in place) */
        double* acc = (double*)malloc(sizeof(double) * ((size_t)a->n + 1));
        unsigned char* occ = (unsigned char*)calloc((size_t)a->n + 1, 1);
        int32_t* list = (int32_t*)malloc(sizeof(int32_t) * ((size_t)a->n + 1));
        for (int32_t q = q0; q < q1; ++q) {
            int64_t s0 = a->base[q], s1 = a->base[q + 1];
            /* emit in place over the row's own slots (count <= ub) */
            int32_t* tmp_idx = list; /* merge_row sorts list then writes */
            int64_t c = 0;
            for (int64_t s = s0; s < s1; ++s) {
                int32_t col = a->icol[s];
                if (!occ[col]) { occ[col] = 1; acc[col] = 0.0; tmp_idx[c++] = col; }
                acc[col] += (double)a->ival[s];
            }
            qsort(tmp_idx, (size_t)c, sizeof(int32_t), cmp_i32);
            for (int64_t t = 0; t < c; ++t) {
                int32_t col = tmp_idx[t];
                a->icol[s0 + t] = col;
                a->ival[s0 + t] = (float)acc[col];
                occ[col] = 0;
            }
            a->cnt[q] = c;
        }
        free(acc);
        free(occ);
        free(list);
    } else if (a->phase == 3) {
        /* compaction into CSR */
        double cs = 0.0;
        for (int32_t q = q0; q < q1; ++q) {
            int64_t c = a->cnt[q];
            memcpy(a->c_idx + a->c_ptr[q], a->icol + a->base[q], sizeof(int32_t) * (size_t)c);
            memcpy(a->c_val + a->c_ptr[q], a->ival + a->base[q], sizeof(float) * (size_t)c);
            for (int64_t t = 0; t < c; ++t) cs += a->c_val[a->c_ptr[q] + t];
        }
        a->checksum = cs;
    }
    return NULL;
}

static void bl_run(bl_args* args, int nthreads, int phase) {
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) {
        args[t].phase = phase;
        if (nthreads > 1) pthread_create(&th[t], NULL, bl_worker, &args[t]);
    }
    if (nthreads == 1) { bl_worker(&args[0]); return; }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

int64_t orc_cpu_baseline(int32_t m, int32_t n, int32_t K,
                         const int32_t* a_ptr, const int32_t* a_idx,
                         const float* a_val, const int32_t* b_ptr,
                         const int32_t* b_idx, const float* b_val,
                         int32_t row_begin, int32_t row_end, int nthreads,
                         double* times_ms, int64_t* nnz_out,
                         double* checksum_out) {
    (void)m;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    int32_t nb = row_end - row_begin;
    bl_args* args = (bl_args*)calloc((size_t)nthreads, sizeof(bl_args));
    int32_t* lo = (int32_t*)malloc(sizeof(int32_t) * ((size_t)K + 1));
    int32_t* hi = (int32_t*)malloc(sizeof(int32_t) * ((size_t)K + 1));
    int64_t* tcur = (int64_t*)malloc(sizeof(int64_t) * (size_t)nthreads * ((size_t)nb + 1));
    int64_t* ub = (int64_t*)calloc((size_t)nb + 1, sizeof(int64_t));
    int64_t* base = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nb + 1));
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nb + 1));
    int64_t* c_ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nb + 1));
    for (int t = 0; t < nthreads; ++t) {
        bl_args* a = &args[t];
        a->tid = t; a->nthreads = nthreads;
        a->m = m; a->n = n; a->K = K; a->r0 = row_begin; a->r1 = row_end;
        a->a_ptr = a_ptr; a->a_idx = a_idx; a->a_val = a_val;
        a->b_ptr = b_ptr; a->b_idx = b_idx; a->b_val = b_val;
        a->lo = lo; a->hi = hi; a->ub = ub; a->base = base; a->tcur = tcur;
        a->cnt = cnt; a->c_ptr = c_ptr;
    }
    double t0 = now_ms();
    /* 1. upper bound (per-thread partial counts, then reduced + scanned) */
    bl_run(args, nthreads, 0);
    for (int32_t q = 0; q < nb; ++q) {
        int64_t s = 0;
        for (int t = 0; t < nthreads; ++t) s += tcur[(int64_t)t * nb + q];
        ub[q] = s;
    }
    base[0] = 0;
    for (int32_t q = 0; q < nb; ++q) base[q + 1] = base[q] + ub[q];
    int64_t nprod = base[nb];
    double t1 = now_ms();
    /* 2. bins over k (the reference's categorize phase) */
    int32_t* bin_of_k = (int32_t*)malloc(sizeof(int32_t) * ((size_t)K + 1));
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * ((size_t)K + 1));
    int64_t offs[ORC_NUM_KBINS + 1];
    orc_kbins(K, a_ptr, b_ptr, nprod, bin_of_k, perm, offs);
    /* per-thread starting cursors: thread t writes after threads < t */
    for (int32_t q = 0; q < nb; ++q) {
        int64_t s = base[q];
        for (int t = 0; t < nthreads; ++t) {
            int64_t c = tcur[(int64_t)t * nb + q];
            tcur[(int64_t)t * nb + q] = s;
            s += c;
        }
    }
    double t2 = now_ms();
    /* 3. expansion */
    int32_t* icol = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nprod + 1));
    float* ival = (float*)malloc(sizeof(float) * ((size_t)nprod + 1));
    for (int t = 0; t < nthreads; ++t) { args[t].icol = icol; args[t].ival = ival; }
    double t2b = now_ms();
    bl_run(args, nthreads, 1);
    double t3 = now_ms();
    /* 4. merge */
    bl_run(args, nthreads, 2);
    double t4 = now_ms();
    /* 5. compaction */
    c_ptr[0] = 0;
    for (int32_t q = 0; q < nb; ++q) c_ptr[q + 1] = c_ptr[q] + cnt[q];
    int64_t nnz = c_ptr[nb];
    int32_t* c_idx = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nnz + 1));
    float* c_val = (float*)malloc(sizeof(float) * ((size_t)nnz + 1));
    for (int t = 0; t < nthreads; ++t) { args[t].c_idx = c_idx; args[t].c_val = c_val; }
    double t4b = now_ms();
    bl_run(args, nthreads, 3);
    double t5 = now_ms();
    double cs = 0.0;
    for (int t = 0; t < nthreads; ++t) cs += args[t].checksum;
    if (times_ms) {
        times_ms[0] = t1 - t0;
        times_ms[1] = t2 - t1;
        times_ms[2] = t3 - t2b;
        times_ms[3] = t4 - t3;
        times_ms[4] = t5 - t4b;
    }
    if (nnz_out) *nnz_out = nnz;
    if (checksum_out) *checksum_out = cs;
    free(bin_of_k); free(perm); free(icol); free(ival); free(c_idx); free(c_val);
    free(lo); free(hi); free(tcur); free(ub); free(base); free(cnt); free(c_ptr);
    free(args);
    return nprod;
}
