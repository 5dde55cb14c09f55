// spg_device.hip - CDNA4 (gfx950) kernels of the outer-product SpGEMM path
// and their launch wrappers.
//
// Layout in HBM (all per context, grow-only):
//   A  CSC  a_ptr[K+1] a_idx[nnzA] (row)  a_val[nnzA]
//   B  CSR  b_ptr[K+1] b_idx[nnzB] (col)  b_val[nnzB]
//   virtual rows: a C row i with more than long_row_ub products is split
//   into T column tiles of width W = 2^tile_log2; every other row is one
//   virtual row.  vstart[m+1] maps rows to their first virtual row, vrow_row
//   maps back, vub/vbase give each virtual row's slot range.
//   intermediates  one buffer of 512-byte groups of 64 products: the 64
//   columns (int32, two 128-byte lines), then the 64 values (f32); product
//   s sits at word islot(s) of the column pointer and of the value pointer
//   (64 words later).  The column-only symbolic pass reads whole lines of
//   columns (with 16-product groups, which put a product's column and value
//   in one line, it fetched the values too: 6.1 vs 3.8 ms on R-MAT scale
//   18), and one wave store of 64 consecutive products writes two whole
//   lines.  Products are grouped by virtual row (so a C row's products are
//   contiguous too)
//   C  CSR  crow[m+1] int64, ccol[nnzC] int32, cval[nnzC] f32
//
// Value contract: a product is the FP32 product a*b (rounded once, as
// src/mm.cu:308 computes it); the products of one C entry are summed in
// FP64 and the sum is rounded to FP32 once.  Structure is the exact pattern
// (numerical zeros kept), columns ascending.
#include "spg_internal.h"

#include <algorithm>
#include <climits>

namespace spg {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// A value every lane of the wave holds, moved to a scalar register.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long long uni64(long long x) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)x);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(x >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << lane_id()) - 1ull;
}

// Four columns of a narrow row, relative to its origin lo (0 .. width-1),
// from the 16-bit column array (q: products b..b+3, one 8-byte load).
__device__ __forceinline__ void rel4(uint2 q, int lo, int (&c)[4]) {
    const unsigned l = (unsigned)lo;
    c[0] = (int)((q.x - l) & 0xFFFFu);
    c[1] = (int)(((q.x >> 16) - l) & 0xFFFFu);
    c[2] = (int)((q.y - l) & 0xFFFFu);
    c[3] = (int)(((q.y >> 16) - l) & 0xFFFFu);
}

// Exclusive scan over a 256-thread workgroup.  sh needs kBlock/64 slots.
template <typename T>
__device__ __forceinline__ T block_exscan(T x, T* total, T* sh) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    T inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    T wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        T s = sh[i];
        if (i < w) wpre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wpre + inc - x;
}

// Map blockIdx.x so that consecutive logical blocks run on one XCD (the
// dispatcher places block b on XCD b % 8).  Bijective for any grid size.
__device__ __forceinline__ long long xcd_swizzle(long long bid, long long nb) {
    const long long per = nb / 8, rem = nb % 8;
    const long long x = bid % 8, q = bid / 8;
    return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

template <typename T>
__global__ void k_fill(T* p, T v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
__global__ void k_copy_n(T* __restrict__ dst, const T* __restrict__ src, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <typename T>
int fill(spg_context* c, T* p, T v, long long n) {
    if (n <= 0) return SPG_OK;
    long long nb = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, p, v, n);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

// ------------------------------------------------------------------ scans
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_scan_block(const TI* __restrict__ in,
                                                        TO* __restrict__ out,
                                                        TO* __restrict__ bsum,
                                                        long long n) {
    __shared__ TO tile[kScanTile];
    __shared__ TO sh[kBlock / 64];
    const long long base = (long long)blockIdx.x * kScanTile;
    // coalesced load into LDS
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        long long g = base + r * kBlock + threadIdx.x;
        tile[r * kBlock + threadIdx.x] = g < n ? (TO)in[g] : (TO)0;
    }
    __syncthreads();
    TO v[kScanItems];
    TO s = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        v[r] = tile[threadIdx.x * kScanItems + r];
        s += v[r];
    }
    TO tot;
    TO pre = block_exscan<TO>(s, &tot, sh);
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        tile[threadIdx.x * kScanItems + r] = pre;
        pre += v[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        long long g = base + r * kBlock + threadIdx.x;
        if (g < n) out[g] = tile[r * kBlock + threadIdx.x];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename TO>
__global__ __launch_bounds__(kBlock) void k_scan_add(TO* __restrict__ out,
                                                      const TO* __restrict__ boff,
                                                      long long n) {
    const TO o = boff[blockIdx.x];
    const long long base = (long long)blockIdx.x * kScanTile;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        long long g = base + r * kBlock + threadIdx.x;
        if (g < n) out[g] += o;
    }
}

template <typename TO>
__global__ void k_copy1(TO* dst, const TO* src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

template <typename TI, typename TO>
int scan_impl(spg_context* c, const TI* in, TO* out, long long n) {
    if (n <= 0) return fill<TO>(c, out, (TO)0, 1);
    const long long nb = (n + kScanTile - 1) / kScanTile;
    TO* bsum = (TO*)scratch_alloc(c, sizeof(TO) * (size_t)(nb + 1));
    if (!bsum) return fail(c, SPG_ENOMEM, "scan scratch");
    hipLaunchKernelGGL((k_scan_block<TI, TO>), dim3((unsigned)nb), dim3(kBlock), 0,
                       c->stream, in, out, bsum, n);
    SPG_HIP(hipGetLastError());
    if (nb == 1) {
        hipLaunchKernelGGL(k_copy1<TO>, dim3(1), dim3(64), 0, c->stream, out + n, bsum);
        SPG_HIP(hipGetLastError());
        return SPG_OK;
    }
    TO* boff = (TO*)scratch_alloc(c, sizeof(TO) * (size_t)(nb + 1));
    if (!boff) return fail(c, SPG_ENOMEM, "scan scratch");
    SPG_TRY((scan_impl<TO, TO>(c, bsum, boff, nb)));
    hipLaunchKernelGGL(k_scan_add<TO>, dim3((unsigned)nb), dim3(kBlock), 0, c->stream,
                       out, boff, n);
    SPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_copy1<TO>, dim3(1), dim3(64), 0, c->stream, out + n, boff + nb);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

int scan_i32_to_i64(spg_context* c, const int32_t* in, int64_t* out, int64_t n) {
    return scan_impl<int32_t, long long>(c, in, (long long*)out, n);
}
int scan_i64(spg_context* c, const int64_t* in, int64_t* out, int64_t n) {
    return scan_impl<long long, long long>(c, (const long long*)in, (long long*)out, n);
}
int scan_i32(spg_context* c, const int32_t* in, int32_t* out, int64_t n) {
    return scan_impl<int32_t, int32_t>(c, in, out, n);
}

// ------------------------------------------------------- stable partition
// Keys in [-1, nb), nb <= kMaxBuckets; -1 is dropped.  Each workgroup owns
// kPartTile consecutive keys; ranks inside a wave come from ballots, so the
// output keeps the input order inside every bucket.
constexpr int kMaxBuckets = 16;
constexpr int kPartIters = 8;
constexpr int kPartTile = kBlock * kPartIters;

__global__ __launch_bounds__(kBlock) void k_part_count(const int32_t* __restrict__ key,
                                                        long long n, int nb,
                                                        long long nblk,
                                                        long long* __restrict__ cnt) {
    __shared__ int sc[kMaxBuckets];
    if (threadIdx.x < kMaxBuckets) sc[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kPartTile;
    for (int r = 0; r < kPartIters; ++r) {
        long long g = base + r * kBlock + threadIdx.x;
        if (g < n) {
            int k = key[g];
            if (k >= 0 && k < nb) atomicAdd(&sc[k], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < nb) cnt[(long long)threadIdx.x * nblk + blockIdx.x] = sc[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_part_scatter(const int32_t* __restrict__ key,
                                                          long long n, int nb,
                                                          long long nblk,
                                                          const long long* __restrict__ off,
                                                          int32_t* __restrict__ perm) {
    __shared__ int wc[kBlock / 64][kMaxBuckets];
    __shared__ long long run[kMaxBuckets];
    const int w = threadIdx.x >> 6;
    if (threadIdx.x < kMaxBuckets)
        run[threadIdx.x] = threadIdx.x < nb ? off[(long long)threadIdx.x * nblk + blockIdx.x] : 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kPartTile;
    for (int r = 0; r < kPartIters; ++r) {
        long long g = base + r * kBlock + threadIdx.x;
        int k = g < n ? key[g] : -1;
        if (k >= nb) k = -1;
        int rank = 0;
        for (int b = 0; b < nb; ++b) {
            unsigned long long msk = __ballot(k == b);
            if (k == b) rank = __popcll(msk & lanemask_lt());
            if (lane_id() == 0) wc[w][b] = __popcll(msk);
        }
        __syncthreads();
        if (k >= 0) {
            long long pos = run[k] + rank;
            for (int q = 0; q < w; ++q) pos += wc[q][k];
            perm[pos] = (int32_t)g;
        }
        __syncthreads();
        if (threadIdx.x < nb) {
            long long s = 0;
            for (int q = 0; q < kBlock / 64; ++q) s += wc[q][threadIdx.x];
            run[threadIdx.x] += s;
        }
        __syncthreads();
    }
}

__global__ void k_part_offsets(const long long* __restrict__ off, int nb, long long nblk,
                               long long* __restrict__ offsets) {
    int b = threadIdx.x;
    if (b <= nb) offsets[b] = off[(long long)b * nblk];  // off has nb*nblk+1 entries
}

int stable_partition(spg_context* c, const int32_t* key, int64_t n, int nb,
                     int32_t* perm, int64_t* offsets) {
    if (nb > kMaxBuckets) return fail(c, SPG_EINVAL, "too many buckets");
    const long long nblk = std::max<long long>(1, (n + kPartTile - 1) / kPartTile);
    long long* cnt = (long long*)scratch_alloc(c, sizeof(long long) * (size_t)(nb * nblk + 1));
    long long* off = (long long*)scratch_alloc(c, sizeof(long long) * (size_t)(nb * nblk + 1));
    if (!cnt || !off) return fail(c, SPG_ENOMEM, "partition scratch");
    if (n <= 0) {
        SPG_TRY(fill<long long>(c, (long long*)offsets, 0, nb + 1));
        return SPG_OK;
    }
    hipLaunchKernelGGL(k_part_count, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream,
                       key, (long long)n, nb, nblk, cnt);
    SPG_HIP(hipGetLastError());
    SPG_TRY((scan_impl<long long, long long>(c, cnt, off, nb * nblk)));
    hipLaunchKernelGGL(k_part_scatter, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream,
                       key, (long long)n, nb, nblk, off, perm);
    SPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_part_offsets, dim3(1), dim3(64), 0, c->stream, off, nb, nblk,
                       (long long*)offsets);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

// ------------------------------------------------------------ COO -> CM
// src/mm.cu:43-63: minor index and value copied in COO order, a histogram
// of the major index; the prefix sum then runs on the GPU (the reference
// copies the histogram to the host and back, src/mm.cu:100-108).
__global__ void k_coo2cm(const triplet* __restrict__ t, int nnz, int type,
                         int* __restrict__ cnt, int* __restrict__ idx,
                         float* __restrict__ val) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    triplet x = t[e];
    int major = type == ROW_MAJOR ? x.ridx : x.cidx;
    idx[e] = type == ROW_MAJOR ? x.cidx : x.ridx;
    val[e] = x.val;
    atomicAdd(&cnt[major], 1);
}

int launch_coo2cm(spg_context* c, const triplet* d_trip, int nnz, int type, int nmajor,
                  int* ptr, int* idx, float* val) {
    int* cnt = (int*)scratch_alloc(c, sizeof(int) * ((size_t)nmajor + 1));
    if (!cnt) return fail(c, SPG_ENOMEM, "coo2cm scratch");
    SPG_TRY(fill<int>(c, cnt, 0, nmajor + 1));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_coo2cm, dim3((nnz + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, d_trip, nnz, type, cnt, idx, val);
        SPG_HIP(hipGetLastError());
    }
    return scan_i32(c, cnt, ptr, nmajor);
}

// Grid of a persistent kernel: every workgroup that can be resident.
static unsigned persistent_grid(spg_context* c, long long nlist, int per_cu) {
    long long g = (long long)c->num_cu * per_cu;
    if (g > nlist) g = nlist;
    return (unsigned)(g > 0 ? g : 1);
}

// ------------------------------------------------------------------ init
// src/mm.cu:138-148 accumulates the upper bound of C row i with one atomic
// per A(i,k).  Here A is first transposed into row lists (the k of every
// entry of row i), and every row then sums its bounds without atomics: hub
// rows of a power-law A would otherwise serialize thousands of atomics on
// one counter (and T times that for the per-tile bounds of long rows).
// Eight columns of A per wave, eight lanes per column (a short column
// takes one or two strided passes), and each column longer than kShortSeg
// walked by the whole wave in turn: short columns neither idle most of a
// wave nor chain many dependent atomics on one lane, and hub columns do not
// serialize.
constexpr int kShortSeg = 64;
constexpr int kGroup = 8;

__global__ __launch_bounds__(kBlock) void k_row_hist(const int* __restrict__ a_ptr,
                                                      const int* __restrict__ a_idx, int K,
                                                      int* __restrict__ rcnt) {
    const int lane = lane_id(), sub = lane & (kGroup - 1);
    const long long k = ((long long)blockIdx.x * kBlock + threadIdx.x) / kGroup;
    int e0 = 0, e1 = 0;
    if (k < K) {
        e0 = a_ptr[k];
        e1 = a_ptr[k + 1];
    }
    const bool lng = e1 - e0 > kShortSeg;
    if (!lng)
        for (int e = e0 + sub; e < e1; e += kGroup) atomicAdd(&rcnt[a_idx[e]], 1);
    unsigned long long mask = __ballot(lng && sub == 0);
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int b0 = __shfl(e0, l, 64), b1 = __shfl(e1, l, 64);
        for (int e = b0 + lane; e < b1; e += 64) atomicAdd(&rcnt[a_idx[e]], 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_row_fill(const int* __restrict__ a_ptr,
                                                      const int* __restrict__ a_idx,
                                                      const float* __restrict__ a_val, int K,
                                                      int* __restrict__ rcur,
                                                      int* __restrict__ rk,
                                                      float* __restrict__ rv) {
    const int lane = lane_id(), sub = lane & (kGroup - 1);
    const long long k = ((long long)blockIdx.x * kBlock + threadIdx.x) / kGroup;
    int e0 = 0, e1 = 0;
    if (k < K) {
        e0 = a_ptr[k];
        e1 = a_ptr[k + 1];
    }
    const bool lng = e1 - e0 > kShortSeg;
    if (!lng) {
        for (int e = e0 + sub; e < e1; e += kGroup) {
            const int pos = atomicAdd(&rcur[a_idx[e]], 1);
            rk[pos] = (int)k;
            rv[pos] = a_val[e];
        }
    }
    unsigned long long mask = __ballot(lng && sub == 0);
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int b0 = __shfl(e0, l, 64), b1 = __shfl(e1, l, 64);
        const int kk = (int)(((long long)blockIdx.x * kBlock + (threadIdx.x - lane + l)) / kGroup);
        for (int e = b0 + lane; e < b1; e += 64) {
            const int pos = atomicAdd(&rcur[a_idx[e]], 1);
            rk[pos] = kk;
            rv[pos] = a_val[e];
        }
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_max(T x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_min(T x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
    return x;
}

// ub_i = sum of nnz(B row k) over the row's k (the flop upper bound), the
// column tiling of the row, and the predictor's row features: log1p of
// nnz(A row i), ub_i, max_k nnz(B row k) and the standard deviation of the
// nnz(B row k).  Eight rows per wave, eight lanes per short row, each long
// row by the whole wave (as in k_row_hist).
__device__ __forceinline__ void row_bound_out(long long i, int r, long long s, long long s2,
                                              int mx, long long long_ub, int T,
                                              const int2* __restrict__ rwin, int tile_log2,
                                              long long* __restrict__ ub,
                                              int* __restrict__ ntile,
                                              float4* __restrict__ feat) {
    ub[i] = s;
    // tiles of the row's column window (batched products: its block)
    const int Ti = rwin ? (rwin[i].y + (1 << tile_log2) - 1) >> tile_log2 : T;
    ntile[i] = (Ti > 1 && s > long_ub) ? Ti : 1;
    double var = 0.0;
    if (r > 0) {
        const double mean = (double)s / r;
        var = (double)s2 / r - mean * mean;
        if (var < 0.0) var = 0.0;
    }
    feat[i] = make_float4(log1pf((float)r), log1pf((float)s), log1pf((float)mx),
                          log1pf((float)sqrt(var)));
}

__global__ __launch_bounds__(kBlock) void k_row_bound(const int* __restrict__ rptr,
                                                       const int* __restrict__ rk,
                                                       const int* __restrict__ b_ptr, int m,
                                                       long long long_ub, int T,
                                                       const int2* __restrict__ rwin,
                                                       int tile_log2,
                                                       long long* __restrict__ ub,
                                                       int* __restrict__ ntile,
                                                       float4* __restrict__ feat) {
    const int lane = lane_id(), sub = lane & (kGroup - 1);
    const long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) / kGroup;
    int e0 = 0, e1 = 0;
    if (i < m) {
        e0 = rptr[i];
        e1 = rptr[i + 1];
    }
    const bool lng = e1 - e0 > kShortSeg;
    {
        long long s = 0, s2 = 0;
        int mx = 0;
        if (!lng) {
            // two entries per lane and pass, their loads issued together (the
            // second, past the row, reads the row's last entry and is masked)
            for (int e = e0 + sub; e < e1; e += 2 * kGroup) {
                const int ka = rk[e], kb = rk[min(e + kGroup, e1 - 1)];
                const int a0 = b_ptr[ka], a1 = b_ptr[ka + 1], c0 = b_ptr[kb], c1 = b_ptr[kb + 1];
                const long long ba = a1 - a0, bb = e + kGroup < e1 ? c1 - c0 : 0;
                s += ba + bb;
                s2 += ba * ba + bb * bb;
                mx = max(mx, (int)max(ba, bb));
            }
        }
#pragma unroll
        for (int o = kGroup / 2; o > 0; o >>= 1) {  // within the 8-lane group
            s += __shfl_xor(s, o, 64);
            s2 += __shfl_xor(s2, o, 64);
            mx = max(mx, __shfl_xor(mx, o, 64));
        }
        if (i < m && !lng && sub == 0)
            row_bound_out(i, e1 - e0, s, s2, mx, long_ub, T, rwin, tile_log2, ub, ntile, feat);
    }
    unsigned long long mask = __ballot(lng && sub == 0);
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int b0 = __shfl(e0, l, 64), b1 = __shfl(e1, l, 64);
        long long s = 0, s2 = 0;
        int mx = 0;
        for (int e = b0 + lane; e < b1; e += 4 * 64) {  // four entries per lane and pass
            int k[4], p0[4], p1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) k[u] = rk[min(e + u * 64, b1 - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                p0[u] = b_ptr[k[u]];
                p1[u] = b_ptr[k[u] + 1];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const long long b = e + u * 64 < b1 ? p1[u] - p0[u] : 0;
                s += b;
                s2 += b * b;
                mx = max(mx, (int)b);
            }
        }
        s = wave_sum(s);
        s2 = wave_sum(s2);
        mx = wave_max(mx);
        if (lane == l)
            row_bound_out(i, e1 - e0, s, s2, mx, long_ub, T, rwin, tile_log2, ub, ntile, feat);
    }
}

// Column range of every virtual row: the row's window [lo, lo + w) (all of
// B's columns for one product; the product's block in a batch), or one
// W-wide tile of it.
__global__ void k_fill_vrows(const int* __restrict__ vstart, const long long* __restrict__ ub,
                             int m, int n, const int2* __restrict__ rwin, int tile_log2,
                             int* __restrict__ vrow_row, long long* __restrict__ vub,
                             int* __restrict__ vlo, int* __restrict__ vwid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int s = vstart[i], e = vstart[i + 1];
    const int wlo = rwin ? rwin[i].x : 0, ww = rwin ? rwin[i].y : n;
    for (int v = s; v < e; ++v) {
        vrow_row[v] = i;
        if (e - s == 1) {
            vub[v] = ub[i];
            vlo[v] = wlo;
            vwid[v] = ww;
        } else {
            const int lo = (v - s) << tile_log2;
            vlo[v] = wlo + lo;
            vwid[v] = min(1 << tile_log2, ww - lo);
        }
    }
}

// Split points of every B row at the tile boundaries lo + t*W of its column
// window (t = 0..T; past the window's last tile: the row's end).
__global__ void k_bsplit(const int* __restrict__ b_ptr, const int* __restrict__ b_idx, int K,
                         int T, const int2* __restrict__ kwin, int tile_log2,
                         int* __restrict__ bsplit) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K * (T + 1)) return;
    const int k = (int)(q / (T + 1)), t = (int)(q - (long long)k * (T + 1));
    int lo = b_ptr[k], hi = b_ptr[k + 1];
    const int Tk = kwin ? (kwin[k].y + (1 << tile_log2) - 1) >> tile_log2 : T;
    if (t >= Tk) {
        bsplit[q] = hi;
        return;
    }
    const long long bound = (kwin ? (long long)kwin[k].x : 0LL) + ((long long)t << tile_log2);
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (b_idx[mid] < bound) lo = mid + 1;
        else hi = mid;
    }
    bsplit[q] = lo;
}

// Products of every tile of a long row: one workgroup per row; the lanes of
// a wave are (entry slot, tile) pairs, 64 / P entries side by side for a row
// of at most P tiles (P a power of two), and sum the tile pieces of their
// share of the row's B rows (coalesced reads of the split points, four
// entries in flight per lane); the slots' and waves' partial sums are then
// added by shuffles and in LDS.
__global__ void k_tiled_key(int m, const int* __restrict__ vstart, int* __restrict__ key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) key[i] = vstart[i + 1] - vstart[i] > 1 ? 0 : -1;
}

// One workgroup per tiled row (rows list[0 .. *count), persistent).
__global__ __launch_bounds__(kBlock) void k_row_tile_bound(const int* __restrict__ rptr,
                                                            const int* __restrict__ rk,
                                                            const int* __restrict__ bsplit,
                                                            const int* __restrict__ vstart,
                                                            const int* __restrict__ list,
                                                            const long long* __restrict__ count,
                                                            int T, long long* __restrict__ vub) {
    __shared__ long long part[kBlock / 64][64];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    constexpr int NWV = kBlock / 64;
    const long long nr = *count;
    // the row's descriptor, and the next row's loaded while this one is summed
    long long q = blockIdx.x;
    int ni = q < nr ? list[q] : 0;
    int nvs = 0, nvs1 = 0, ne0 = 0, ne1 = 0;
    if (q < nr) {
        nvs = vstart[ni];
        nvs1 = vstart[ni + 1];
        ne0 = rptr[ni];
        ne1 = rptr[ni + 1];
    }
    for (; q < nr; q += gridDim.x) {
        const int vs = nvs;
        const int Ti = nvs1 - vs;  // the row's tiles (T + 1: split-point stride)
        const int e0 = ne0, e1 = ne1;
        if (q + gridDim.x < nr) {
            ni = list[q + gridDim.x];
            nvs = vstart[ni];
            nvs1 = vstart[ni + 1];
            ne0 = rptr[ni];
            ne1 = rptr[ni + 1];
        }
        // a wave covers G = 64 / P entries at a time, P = the tiles rounded
        // up to a power of two (at most 64): lane = (entry slot, tile)
        const int lp = Ti >= 64 ? 6 : 32 - __clz(Ti - 1);
        const int P = 1 << lp, G = 64 >> lp;
        const int g = lane >> lp, tl = lane & (P - 1);
        const int stp = NWV * G;
        for (int t0 = 0; t0 < Ti; t0 += P) {
            const int t = t0 + tl;
            long long s = 0;
            if (t < Ti) {
                int e = e0 + wv * G + g;
                for (; e + 3 * stp < e1; e += 4 * stp) {
                    int kk[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) kk[u] = rk[e + u * stp];
                    int d[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int* bs = bsplit + (long long)kk[u] * (T + 1);
                        d[u] = bs[t + 1] - bs[t];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) s += d[u];
                }
                for (; e < e1; e += stp) {
                    const int* bs = bsplit + (long long)rk[e] * (T + 1);
                    s += bs[t + 1] - bs[t];
                }
            }
            for (int o = P; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);  // over the entry slots
            part[wv][lane] = s;
            __syncthreads();
            if (wv == 0 && g == 0 && t < Ti) {
                long long tot = 0;
#pragma unroll
                for (int u = 0; u < NWV; ++u) tot += part[u][lane];
                vub[vs + t] = tot;
            }
            __syncthreads();
        }
    }
}

// A transposed into rows without global atomics (a counting sort in two
// levels, all counters in LDS).  Random global atomics run at a fixed rate
// (about 27e9/s on this part, whatever their scope), so one per entry, as
// the row histogram and fill below do, costs 2 ms per 53M entries.  Here:
// (0) the column of every entry; (1) per tile of E entries, an LDS
// histogram of row buckets (2^BL consecutive rows, at most 4096 buckets),
// scanned bucket-major over the tiles, and (2) the tile's entries scattered
// to their bucket's range; (3) one workgroup per bucket counts its rows in
// LDS, writes their row pointers (bucket start + prefix) and places every
// entry.  The order of the entries inside a row is arbitrary, as with the
// atomic fill.
constexpr int kTrMaxBuckets = 4096;
constexpr int kTrMaxBucketLog2 = 14;  // rows per bucket <= 16384 (64 KB of LDS counters)
constexpr long long kTrStageMin = 4096;  // mean entries per bucket from which buckets are staged
constexpr size_t kTrStageLds = 144 << 10;  // LDS of a staged bucket workgroup, at most

__global__ __launch_bounds__(kBlock) void k_tr_kidx(const int* __restrict__ a_ptr, int K,
                                                     int* __restrict__ kidx) {
    const int lane = lane_id(), sub = lane & (kGroup - 1);
    const long long k = ((long long)blockIdx.x * kBlock + threadIdx.x) / kGroup;
    int e0 = 0, e1 = 0;
    if (k < K) {
        e0 = a_ptr[k];
        e1 = a_ptr[k + 1];
    }
    const bool lng = e1 - e0 > kShortSeg;
    if (!lng)
        for (int e = e0 + sub; e < e1; e += kGroup) kidx[e] = (int)k;
    unsigned long long mask = __ballot(lng && sub == 0);
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int b0 = __shfl(e0, l, 64), b1 = __shfl(e1, l, 64);
        const int kk = (int)(((long long)blockIdx.x * kBlock + (threadIdx.x - lane + l)) / kGroup);
        for (int e = b0 + lane; e < b1; e += 64) kidx[e] = kk;
    }
}

__global__ __launch_bounds__(kBlock) void k_tr_count(const int* __restrict__ a_idx,
                                                      long long nnz, int E, int BL, int NB,
                                                      long long ntiles, int* __restrict__ H) {
    __shared__ int hist[kTrMaxBuckets];
    for (int b = threadIdx.x; b < NB; b += kBlock) hist[b] = 0;
    __syncthreads();
    const long long e0 = (long long)blockIdx.x * E, e1 = min(e0 + E, nnz);
    for (long long e = e0 + threadIdx.x; e < e1; e += 4 * kBlock) {
        int r[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = a_idx[min(e + u * kBlock, e1 - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e + u * kBlock < e1) atomicAdd(&hist[r[u] >> BL], 1);
    }
    __syncthreads();
    // tile-major (one contiguous row per tile), transposed to bucket-major
    // for the scan: written bucket-major, each tile's counts were NB
    // scattered 4-byte stores
    for (int b = threadIdx.x; b < NB; b += kBlock) H[(long long)blockIdx.x * NB + b] = hist[b];
}

__global__ __launch_bounds__(kBlock) void k_tr_scatter(
    const int* __restrict__ a_idx, const float* __restrict__ a_val, const int* __restrict__ kidx,
    long long nnz, int E, int BL, int NB, long long ntiles, const int* __restrict__ O,
    int* __restrict__ tr, int* __restrict__ tk, float* __restrict__ tv) {
    __shared__ int cur[kTrMaxBuckets];
    // the tile's offsets, tile-major (a contiguous row)
    for (int b = threadIdx.x; b < NB; b += kBlock) cur[b] = O[(long long)blockIdx.x * NB + b];
    __syncthreads();
    const long long e0 = (long long)blockIdx.x * E, e1 = min(e0 + E, nnz);
    for (long long e = e0 + threadIdx.x; e < e1; e += 4 * kBlock) {
        int r[4], k[4];
        float x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long long q = min(e + u * kBlock, e1 - 1);
            r[u] = a_idx[q];
            k[u] = kidx[q];
            x[u] = a_val[q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (e + u * kBlock < e1) {
                const int pos = atomicAdd(&cur[r[u] >> BL], 1);
                tr[pos] = r[u];
                tk[pos] = k[u];
                tv[pos] = x[u];
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_tr_bucket(
    const int* __restrict__ tr, const int* __restrict__ tk, const float* __restrict__ tv,
    const int* __restrict__ O, long long ntiles, int NB, int BL, int m, int* __restrict__ rptr,
    int* __restrict__ rk, float* __restrict__ rv) {
    __shared__ int rc[1 << kTrMaxBucketLog2];
    __shared__ int sh[kBlock / 64];
    const int b = blockIdx.x;
    const int nrow = 1 << BL, r0 = b << BL;
    const int p0 = O[(long long)b * ntiles], p1 = O[(long long)(b + 1) * ntiles];
    for (int i = threadIdx.x; i < nrow; i += kBlock) rc[i] = 0;
    __syncthreads();
    // four entries per thread and pass, their loads issued together (a
    // position past the bucket reads its last entry and is masked)
    for (int p = p0 + threadIdx.x; p < p1; p += 4 * kBlock) {
        int r[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = tr[min(p + u * kBlock, p1 - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (p + u * kBlock < p1) atomicAdd(&rc[r[u] - r0], 1);
    }
    __syncthreads();
    // exclusive scan of the bucket's row counts: consecutive counters per thread
    const int per = (nrow + kBlock - 1) / kBlock;
    const int i0 = min(nrow, threadIdx.x * per), i1 = min(nrow, i0 + per);
    int mine = 0;
    for (int i = i0; i < i1; ++i) mine += rc[i];
    int tot;
    int run = block_exscan<int>(mine, &tot, sh);
    for (int i = i0; i < i1; ++i) {
        const int cnt = rc[i];
        rc[i] = run;  // the row's cursor (relative to the bucket)
        if (r0 + i < m) rptr[r0 + i] = p0 + run;
        run += cnt;
    }
    if (b == NB - 1 && threadIdx.x == 0) rptr[m] = p1;
    __syncthreads();
    for (int p = p0 + threadIdx.x; p < p1; p += 4 * kBlock) {
        int r[4], k[4];
        float x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = min(p + u * kBlock, p1 - 1);
            r[u] = tr[q];
            k[u] = tk[q];
            x[u] = tv[q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (p + u * kBlock < p1) {
                const int pos = atomicAdd(&rc[r[u] - r0], 1);
                rk[p0 + pos] = k[u];
                rv[p0 + pos] = x[u];
            }
        }
    }
}

// k_tr_bucket for large buckets: the entries are placed in LDS (2^BL row
// counters, then SW staged entries, sized at launch) and the bucket is
// written front to back.  Placed straight into global memory, the entries
// of the 512 buckets in flight (tens of MB of partly written lines) left
// L2 before their lines were complete: 2.8 GB written for 0.44 GB on the
// 256-product batch.  A bucket of more than SW entries places them
// directly.
__global__ __launch_bounds__(kBlock) void k_tr_bucket_staged(
    const int* __restrict__ tr, const int* __restrict__ tk, const float* __restrict__ tv,
    const int* __restrict__ O, long long ntiles, int NB, int BL, int m, int SW,
    int* __restrict__ rptr, int* __restrict__ rk, float* __restrict__ rv) {
    extern __shared__ int lds[];
    __shared__ int sh[kBlock / 64];
    const int nrow = 1 << BL;
    int* rc = lds;
    int2* stage = (int2*)(lds + max(nrow, 4));
    const int b = blockIdx.x, r0 = b << BL;
    const int p0 = O[(long long)b * ntiles], p1 = O[(long long)(b + 1) * ntiles];
    const bool staged = p1 - p0 <= SW;  // uniform
    for (int i = threadIdx.x; i < nrow; i += kBlock) rc[i] = 0;
    __syncthreads();
    for (int p = p0 + threadIdx.x; p < p1; p += 4 * kBlock) {
        int r[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = tr[min(p + u * kBlock, p1 - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (p + u * kBlock < p1) atomicAdd(&rc[r[u] - r0], 1);
    }
    __syncthreads();
    const int per = (nrow + kBlock - 1) / kBlock;
    const int i0 = min(nrow, threadIdx.x * per), i1 = min(nrow, i0 + per);
    int mine = 0;
    for (int i = i0; i < i1; ++i) mine += rc[i];
    int tot;
    int run = block_exscan<int>(mine, &tot, sh);
    for (int i = i0; i < i1; ++i) {
        const int cnt = rc[i];
        rc[i] = run;
        if (r0 + i < m) rptr[r0 + i] = p0 + run;
        run += cnt;
    }
    if (b == NB - 1 && threadIdx.x == 0) rptr[m] = p1;
    __syncthreads();
    for (int p = p0 + threadIdx.x; p < p1; p += 4 * kBlock) {
        int r[4], k[4];
        float x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = min(p + u * kBlock, p1 - 1);
            r[u] = tr[q];
            k[u] = tk[q];
            x[u] = tv[q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (p + u * kBlock < p1) {
                const int pos = atomicAdd(&rc[r[u] - r0], 1);
                if (staged) {
                    stage[pos] = make_int2(k[u], __float_as_int(x[u]));
                } else {
                    rk[p0 + pos] = k[u];
                    rv[p0 + pos] = x[u];
                }
            }
        }
    }
    if (!staged) return;
    __syncthreads();
    for (int q = threadIdx.x; q < p1 - p0; q += kBlock) {
        const int2 e = stage[q];
        rk[p0 + q] = e.x;
        rv[p0 + q] = __int_as_float(e.y);
    }
}

// out[c][r] = in[r][c] for an R x C matrix of ints, through 32 x 32 tiles
// in LDS (both sides coalesced).
__global__ __launch_bounds__(256) void k_transpose_i32(const int* __restrict__ in, int R, int C,
                                                       int* __restrict__ out) {
    __shared__ int tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        const int r = r0 + ty + j, c = c0 + tx;
        if (r < R && c < C) tile[ty + j][tx] = in[(long long)r * C + c];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        const int c = c0 + ty + j, r = r0 + tx;
        if (c < C && r < R) out[(long long)c * R + r] = tile[tx][ty + j];
    }
}

// Bucket parameters of the transpose, or false when the rows are too many
// (more than 4096 buckets of 16384) and the atomic transpose is used.
static bool tr_params(int m, long long nnz, int* BL, int* NB, int* E, long long* ntiles) {
    int lg = 0;
    while ((1LL << lg) < (long long)std::max(m, 1)) ++lg;
    const int bl = std::max(0, lg - 12);
    if (bl > kTrMaxBucketLog2) return false;
    *BL = bl;
    *NB = (int)(((long long)std::max(m, 1) + (1LL << bl) - 1) >> bl);
    long long e = 16384;
    if ((nnz + e - 1) / e > 4096) e = (((nnz + 4095) / 4096) + kBlock - 1) / kBlock * kBlock;
    *E = (int)e;
    *ntiles = std::max(1LL, (nnz + e - 1) / e);
    return true;
}

// Whether InitGEMM transposes A by the counting sort: SORT always (when
// the rows allow it), AUTO from 2^20 entries on (below, the atomic
// transpose's fewer launches win), never with ATOMIC.
static bool use_sort(const spg_context* c, long long nnz) {
    return c->cfg.transpose_mode == SPG_TRANSPOSE_SORT ||
           (c->cfg.transpose_mode == SPG_TRANSPOSE_AUTO && nnz >= (1LL << 20));
}

// Scratch the transpose needs (bytes), 0 when it is not used.
size_t transpose_scratch(const spg_context* c, int m, long long nnz) {
    int BL, NB, E;
    long long nt;
    if (!use_sort(c, nnz) || nnz <= 0 || !tr_params(m, nnz, &BL, &NB, &E, &nt)) return 0;
    return (size_t)16 * (size_t)nnz + (size_t)12 * (size_t)NB * (size_t)nt + ((size_t)8 << 20);
}

// Upper bounds, tiling and features of every row (A by rows is built).
static int launch_row_bound(spg_context* c) {
    const int m = c->m;
    if (m > 0) {
        const long long nbm = ((long long)m * kGroup + kBlock - 1) / kBlock;  // 8 lanes per row
        hipLaunchKernelGGL(k_row_bound, dim3((unsigned)nbm), dim3(kBlock), 0, c->stream,
                           (const int*)c->arow_ptr.p, (const int*)c->arow_k.p, c->B.ptr, m,
                           (long long)c->cfg.long_row_ub, c->T, (const int2*)c->rwin,
                           c->cfg.tile_log2, (long long*)c->ub.p, (int*)c->ntile.p,
                           (float4*)c->feat.p);
        SPG_HIP(hipGetLastError());
    }
    return scan_i32(c, (const int32_t*)c->ntile.p, (int32_t*)c->vstart.p, m);
}

static int launch_transpose(spg_context* c, bool* done) {
    *done = false;
    const int m = c->m, K = c->K;
    const long long nnz = c->A.nnz;
    int BL, NB, E;
    long long nt;
    if (!use_sort(c, nnz) || nnz <= 0 || m <= 0 || !tr_params(m, nnz, &BL, &NB, &E, &nt))
        return SPG_OK;
    int* kidx = (int*)scratch_alloc(c, sizeof(int) * (size_t)nnz);
    int* tr = (int*)scratch_alloc(c, sizeof(int) * (size_t)nnz);
    int* tk = (int*)scratch_alloc(c, sizeof(int) * (size_t)nnz);
    float* tv = (float*)scratch_alloc(c, sizeof(float) * (size_t)nnz);
    int* H = (int*)scratch_alloc(c, sizeof(int) * (size_t)NB * (size_t)nt);
    int* O = (int*)scratch_alloc(c, sizeof(int) * ((size_t)NB * (size_t)nt + 1));
    int* HT = (int*)scratch_alloc(c, sizeof(int) * (size_t)NB * (size_t)nt);  // tile-major H, then O
    if (!kidx || !tr || !tk || !tv || !H || !O || !HT)
        return fail(c, SPG_ENOMEM, "transpose scratch");
    const long long nbk = ((long long)K * kGroup + kBlock - 1) / kBlock;  // 8 lanes per column
    hipLaunchKernelGGL(k_tr_kidx, dim3((unsigned)nbk), dim3(kBlock), 0, c->stream, c->A.ptr, K,
                       kidx);
    SPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tr_count, dim3((unsigned)nt), dim3(kBlock), 0, c->stream, c->A.idx, nnz,
                       E, BL, NB, nt, HT);
    SPG_HIP(hipGetLastError());
    // [tile][bucket] -> [bucket][tile] for the bucket-major scan, and back
    const dim3 tb(256);
    hipLaunchKernelGGL(k_transpose_i32, dim3((unsigned)((NB + 31) / 32), (unsigned)((nt + 31) / 32)),
                       tb, 0, c->stream, (const int*)HT, (int)nt, NB, H);
    SPG_HIP(hipGetLastError());
    SPG_TRY(scan_i32(c, H, O, (long long)NB * nt));
    hipLaunchKernelGGL(k_transpose_i32, dim3((unsigned)((nt + 31) / 32), (unsigned)((NB + 31) / 32)),
                       tb, 0, c->stream, (const int*)O, NB, (int)nt, HT);
    SPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tr_scatter, dim3((unsigned)nt), dim3(kBlock), 0, c->stream, c->A.idx,
                       c->A.val, (const int*)kidx, nnz, E, BL, NB, nt, (const int*)HT, tr, tk, tv);
    SPG_HIP(hipGetLastError());
    // buckets of many entries are placed in LDS and written front to back
    const long long avg = nnz / NB;
    const size_t rc_bytes = sizeof(int) * (size_t)std::max(1 << BL, 4);
    int SW = 0;
    if (avg >= kTrStageMin) {
        SW = 16384;
        while (SW > 1024 && rc_bytes + 8 * (size_t)SW > kTrStageLds) SW >>= 1;
        if (rc_bytes + 8 * (size_t)SW > kTrStageLds) SW = 0;
    }
    if (SW > 0)
        hipLaunchKernelGGL(k_tr_bucket_staged, dim3((unsigned)NB), dim3(kBlock),
                           rc_bytes + 8 * (size_t)SW, c->stream, (const int*)tr, (const int*)tk,
                           (const float*)tv, (const int*)O, nt, NB, BL, m, SW,
                           (int*)c->arow_ptr.p, (int*)c->arow_k.p, (float*)c->arow_v.p);
    else
        hipLaunchKernelGGL(k_tr_bucket, dim3((unsigned)NB), dim3(kBlock), 0, c->stream,
                           (const int*)tr, (const int*)tk, (const float*)tv, (const int*)O, nt, NB,
                           BL, m, (int*)c->arow_ptr.p, (int*)c->arow_k.p, (float*)c->arow_v.p);
    SPG_HIP(hipGetLastError());
    *done = true;
    return SPG_OK;
}

int launch_init(spg_context* c) {
    const int m = c->m, K = c->K;
    int* rptr = (int*)c->arow_ptr.p;
    int* rk = (int*)c->arow_k.p;
    bool transposed = false;
    SPG_TRY(launch_transpose(c, &transposed));
    if (transposed) return launch_row_bound(c);
    int* rcnt = (int*)scratch_alloc(c, sizeof(int) * ((size_t)m + 1));
    if (!rcnt) return fail(c, SPG_ENOMEM, "init scratch");
    SPG_TRY(fill<int>(c, rcnt, 0, m + 1));
    const long long nbk = ((long long)K * kGroup + kBlock - 1) / kBlock;  // 8 lanes per column
    if (K > 0) {
        hipLaunchKernelGGL(k_row_hist, dim3((unsigned)nbk), dim3(kBlock), 0, c->stream, c->A.ptr,
                           c->A.idx, K, rcnt);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(scan_i32(c, rcnt, rptr, m));
    // rcnt becomes the fill cursor (a copy of the row starts)
    hipLaunchKernelGGL(k_copy_n<int>, dim3((m + kBlock) / kBlock), dim3(kBlock), 0, c->stream,
                       rcnt, (const int*)rptr, (long long)m + 1);
    SPG_HIP(hipGetLastError());
    if (K > 0) {
        hipLaunchKernelGGL(k_row_fill, dim3((unsigned)nbk), dim3(kBlock), 0, c->stream, c->A.ptr,
                           c->A.idx, c->A.val, K, rcnt, rk, (float*)c->arow_v.p);
        SPG_HIP(hipGetLastError());
    }
    return launch_row_bound(c);
}

int launch_init_vrows(spg_context* c) {
    const int m = c->m, K = c->K;
    if (m > 0) {
        hipLaunchKernelGGL(k_fill_vrows, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, (const int*)c->vstart.p, (const long long*)c->ub.p, m,
                           c->n, (const int2*)c->rwin, c->cfg.tile_log2, (int*)c->vrow_row.p,
                           (long long*)c->vub.p, (int*)c->vlo.p, (int*)c->vwid.p);
        SPG_HIP(hipGetLastError());
    }
    if (c->have_long && K > 0) {
        long long nq = (long long)K * (c->T + 1);
        hipLaunchKernelGGL(k_bsplit, dim3((unsigned)((nq + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, c->stream, c->B.ptr, c->B.idx, K, c->T, (const int2*)c->kwin,
                           c->cfg.tile_log2, (int*)c->bsplit.p);
        SPG_HIP(hipGetLastError());
        // the tiled rows, compacted on the device (their count stays there)
        hipLaunchKernelGGL(k_tiled_key, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, m, (const int*)c->vstart.p, (int*)c->vclass.p);
        SPG_HIP(hipGetLastError());
        SPG_TRY(stable_partition(c, (const int32_t*)c->vclass.p, m, 1, (int32_t*)c->vlist.p,
                                 (int64_t*)c->vclass_off.p));
        hipLaunchKernelGGL(k_row_tile_bound, dim3(persistent_grid(c, m, 8)), dim3(kBlock), 0,
                           c->stream, (const int*)c->arow_ptr.p, (const int*)c->arow_k.p,
                           (const int*)c->bsplit.p, (const int*)c->vstart.p,
                           (const int*)c->vlist.p, (const long long*)c->vclass_off.p + 1, c->T,
                           (long long*)c->vub.p);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(scan_i64(c, (const int64_t*)c->vub.p, (int64_t*)c->vbase.p, c->V));
    return fill<int>(c, (int*)c->cursor.p, 0, c->V);
}

// The sizes the expansion's 32-bit slot cursors and chunk offsets depend on
// (checked on the host by spgSlotLimitStatus): out[0] the largest virtual-row
// bound, out[1] the largest bound of an untiled row, out[2] the longest B
// row; out[3] receives nprod (vbase[V]).  out[0..2] start at zero.
__global__ __launch_bounds__(kBlock) void k_limits(long long V, const long long* __restrict__ vub,
                                                   const int* __restrict__ vrow_row,
                                                   const int* __restrict__ vstart, int K,
                                                   const int* __restrict__ b_ptr,
                                                   const long long* __restrict__ vbase,
                                                   unsigned long long* __restrict__ out) {
    long long mv = 0, mu = 0, mb = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const long long u = vub[v];
        const int i = vrow_row[v];
        mv = max(mv, u);
        if (vstart[i + 1] - vstart[i] == 1) mu = max(mu, u);
    }
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < K; k += stride)
        mb = max(mb, (long long)(b_ptr[k + 1] - b_ptr[k]));
    // one atomic per workgroup and value (atomics of a whole grid on one
    // address serialize)
    __shared__ long long part[3][kBlock / 64];
    mv = wave_max(mv);
    mu = wave_max(mu);
    mb = wave_max(mb);
    if (lane_id() == 0) {
        part[0][threadIdx.x >> 6] = mv;
        part[1][threadIdx.x >> 6] = mu;
        part[2][threadIdx.x >> 6] = mb;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        long long x = 0;
        for (int w = 0; w < kBlock / 64; ++w) x = max(x, part[threadIdx.x][w]);
        if (x) atomicMax(&out[threadIdx.x], (unsigned long long)x);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[3] = (unsigned long long)vbase[V];
}

int launch_limits(spg_context* c, unsigned long long* d_out) {
    SPG_TRY(fill<unsigned long long>(c, d_out, 0ull, 4));
    const long long work = std::max<long long>(c->V, c->K);
    const unsigned g = persistent_grid(c, (work + kBlock - 1) / kBlock, 1);
    hipLaunchKernelGGL(k_limits, dim3(g), dim3(kBlock), 0, c->stream, c->V,
                       (const long long*)c->vub.p, (const int*)c->vrow_row.p,
                       (const int*)c->vstart.p, c->K, c->B.ptr, (const long long*)c->vbase.p,
                       d_out);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

// -------------------------------------------------------------- categorize
// src/mm.cu:196-257 with D5 (one range), D6 (b_len >= 128 -> bin 8) and D7
// (64-bit threshold) fixed.
__global__ void k_kbin(const int* __restrict__ a_ptr, const int* __restrict__ b_ptr, int K,
                       long long nprod, int heavy_factor, int* __restrict__ kbin) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const long long alen = a_ptr[k + 1] - a_ptr[k];
    const long long blen = b_ptr[k + 1] - b_ptr[k];
    const long long thresh = (nprod / K) * heavy_factor;  // the reference's factor: 50
    int bin;
    if (blen <= 0) bin = -1;
    else if (alen * blen > thresh) bin = 0;
    else if (blen >= 128) bin = 8;
    else bin = 32 - __clz((int)blen);  // floor(log2 blen) + 1
    kbin[k] = bin;
}

// Work items of k (the "workload table"): entry chunks of about P products;
// a B row longer than P is also cut into P-sized column chunks.
__device__ __forceinline__ long long items_of(long long alen, long long blen, int P) {
    if (alen == 0 || blen == 0) return 0;
    if (blen <= P) {
        long long epi = P / blen;
        if (epi < 1) epi = 1;
        return (alen + epi - 1) / epi;
    }
    return alen * ((blen + P - 1) / P);
}

__global__ void k_kitems(const int* __restrict__ kperm, long long nk,
                         const int* __restrict__ a_ptr, const int* __restrict__ b_ptr, int P,
                         int* __restrict__ kitems) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nk) return;
    const int k = kperm[p];
    kitems[p] = (int)items_of(a_ptr[k + 1] - a_ptr[k], b_ptr[k + 1] - b_ptr[k], P);
}

__global__ void k_gather_bin_items(const long long* __restrict__ kbin_off,
                                   const long long* __restrict__ kitem_off,
                                   long long* __restrict__ out) {
    int b = threadIdx.x;
    if (b <= SPG_NUM_KBINS) out[b] = kitem_off[kbin_off[b]];
}

int launch_categorize(spg_context* c) {
    const int K = c->K;
    if (K > 0) {
        hipLaunchKernelGGL(k_kbin, dim3((K + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                           c->A.ptr, c->B.ptr, K, (long long)c->nprod, c->cfg.heavy_factor,
                           (int*)c->kbin.p);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(stable_partition(c, (const int32_t*)c->kbin.p, K, SPG_NUM_KBINS,
                             (int32_t*)c->kperm.p, (int64_t*)c->kbin_off.p));
    long long off_h[SPG_NUM_KBINS + 1];
    SPG_TRY(readback(c, c->kbin_off.p, sizeof(off_h), off_h));
    for (int b = 0; b <= SPG_NUM_KBINS; ++b) c->kbin_off_h[b] = off_h[b];
    const long long nk = c->kbin_off_h[SPG_NUM_KBINS];
    if (nk > 0) {
        hipLaunchKernelGGL(k_kitems, dim3((unsigned)((nk + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, c->stream, (const int*)c->kperm.p, nk, c->A.ptr, c->B.ptr,
                           c->cfg.item_products, (int*)c->kitems.p);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(scan_i32_to_i64(c, (const int32_t*)c->kitems.p, (int64_t*)c->kitem_off.p, nk));
    long long* dtmp = (long long*)scratch_alloc(c, sizeof(long long) * (SPG_NUM_KBINS + 1));
    if (!dtmp) return fail(c, SPG_ENOMEM, "categorize scratch");
    hipLaunchKernelGGL(k_gather_bin_items, dim3(1), dim3(64), 0, c->stream,
                       (const long long*)c->kbin_off.p, (const long long*)c->kitem_off.p, dtmp);
    SPG_HIP(hipGetLastError());
    SPG_TRY(readback(c, dtmp, sizeof(off_h), off_h));
    for (int b = 0; b <= SPG_NUM_KBINS; ++b) c->kitems_bin_h[b] = off_h[b];
    return SPG_OK;
}

// Persistent loop helper: the workgroup takes list entries from a global
// counter (largest rows come first in the list, so the tail is short).
__device__ __forceinline__ long long next_item(unsigned long long* ctr, long long* sj) {
    __syncthreads();
    if (threadIdx.x == 0) *sj = (long long)atomicAdd(ctr, 1ull);
    __syncthreads();
    return *sj;
}

// -------------------------------------------------------------- expansion
// src/mm.cu:288-312 (binGEMM) / 332-354 (simpleGEMM): for the A entries of
// one column k and B row k, every product a_ik * b_kj goes to a slot of C
// row i.  Slots are reserved per (entry, virtual row) with one atomic, as
// the reference reserves per row (src/mm.cu:304); lanes then write runs of
// consecutive slots.  A group of G lanes (G = next power of two >= the B row
// piece, at most 64) serves one A entry, so short B rows do not idle lanes.
constexpr int kTBCap = 2048;     // LDS slot bases of long-row tiles per batch
// Workgroups of the binned expansion per CU: its LDS (24.6 KB) allows six,
// and its registers are held to six waves per SIMD (measured: 6 beats 5, 7
// and 8, which spill or shrink the slot tables)
constexpr int kExpandWaves = 6;
constexpr int kItemChunk = 8;    // work items taken per counter increment

struct ExpandArgs {
    const int* a_ptr;
    const int* a_idx;
    const float* a_val;
    const int* b_ptr;
    const int* b_idx;
    const float* b_val;
    const int* kperm;
    const long long* kitem_off;
    long long np;               // positions in kperm (binned k)
    long long nitems;           // work items
    int P;                      // products per work item
    const int* vstart;
    const long long* vbase;
    const int* vwid;
    int* cursor;
    const int* bsplit;
    const int2* kwin;  // column window of every k (batched products), or null
    int T, tile_log2;
    int swizzle;
    int* icol;
    uint16_t* icol16;  // columns of narrow rows (width <= the tile width)
    float* ival;
    const int* rcskip;  // rows expanded in chunks (key >= 0) are skipped, or null
};

struct ExpandLds {
    long long sbase[kBlock];
    float sa[kBlock];
    int sflag[kBlock];
    long long tbase[kTBCap];
    int sbs[kTBCap / 2 + 1];
    int sh[kBlock / 64];
};

// Expands A entries [e_begin, e_begin+ne) of column k against B row k's
// positions [j0, j0+nj).  G lanes per entry; every thread of the workgroup
// calls it with the same arguments.
__device__ void expand_work(const ExpandArgs& g, ExpandLds& L, int k, long long e_begin,
                            long long ne, int j0, int nj, int lg) {
    const int G = 1 << lg;
    const int NE = kBlock >> lg;  // entries written at a time (G lanes each)
    const int T = g.T;
    const int b0 = g.b_ptr[k];
    const int tfirst = g.kwin ? (g.kwin[k].x >> g.tile_log2) : 0;  // windows are W-aligned
    // entries whose slots are reserved per round: as many as the LDS tables
    // hold, so one round of atomics serves several rounds of writes
    const int R = T > 1 ? min(kBlock, kTBCap / T) : kBlock;
    __syncthreads();  // LDS of the previous item is free
    if (T > 1)
        for (int t = threadIdx.x; t <= T; t += kBlock) L.sbs[t] = g.bsplit[(long long)k * (T + 1) + t];
    const int jlo = b0 + j0, jhi = b0 + j0 + nj;
    // Every A entry of the item is scaled against the same piece of B row k,
    // and a lane serves the same positions of it for every entry (lane l of
    // a G-lane group: l, l+G, l+2G, l+3G), so pieces of at most 4G columns
    // are loaded into registers once per item instead of once per entry.
    const int lj = threadIdx.x & (G - 1);
    const bool breg = nj <= 4 * G;
    int bc[4] = {0, 0, 0, 0};
    float bx[4] = {0.f, 0.f, 0.f, 0.f};
    if (breg) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u * G < nj) {  // uniform; nj >= 1, positions past it clamped
                const int q = min(jlo + lj + u * G, jhi - 1);
                bc[u] = g.b_idx[q];
                bx[u] = g.b_val[q];
            }
        }
    }
    for (long long eb = 0; eb < ne; eb += R) {
        int cnt = (int)min((long long)R, ne - eb);
        if (eb > 0) __syncthreads();
        // the round's entries (those of rows expanded in chunks left out,
        // the others packed to the front), each with its slot reservation
        int i = 0, q = threadIdx.x;
        float av = 0.f;
        bool keep = threadIdx.x < cnt;
        if (keep) {
            const long long e = e_begin + eb + threadIdx.x;
            i = g.a_idx[e];
            av = g.a_val[e];
        }
        if (g.rcskip) {  // uniform
            if (keep) keep = g.rcskip[i] < 0;
            q = block_exscan<int>(keep ? 1 : 0, &cnt, L.sh);
        }
        if (keep) {
            L.sa[q] = av;
            const int vs = g.vstart[i];
            if (g.vstart[i + 1] - vs == 1) {
                const int off = atomicAdd(&g.cursor[vs], nj);
                L.sbase[q] = g.vbase[vs] + off;
                L.sflag[q] = g.vwid[vs] <= (1 << g.tile_log2) ? 2 : 0;  // 2: narrow
            } else {
                L.sbase[q] = vs;
                L.sflag[q] = 1;
            }
        }
        __syncthreads();
        if (T > 1) {
            for (int pr = threadIdx.x; pr < cnt * T; pr += kBlock) {
                const int q = pr / T, t = pr - q * T;
                if (L.sflag[q] != 1) continue;
                const int s = max(L.sbs[t], jlo), e2 = min(L.sbs[t + 1], jhi);
                if (e2 > s) {
                    const int v = (int)L.sbase[q] + t;
                    const int off = atomicAdd(&g.cursor[v], e2 - s);
                    L.tbase[q * T + t] = g.vbase[v] + off - s;
                }
            }
            __syncthreads();
        }
        // every load before this point is complete here (the slot
        // reservation already waited for its atomics, which are issued after
        // them), said explicitly so that the stores below do not each wait
        // for the stores before them
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        for (int sb = 0; sb < cnt; sb += NE) {  // writes, NE entries at a time
            const int q = sb + (threadIdx.x >> lg);
            if (q >= cnt) continue;
            const float a = L.sa[q];
            const int fl = L.sflag[q];
            const long long base = L.sbase[q];
            const int* __restrict__ bidx = g.b_idx;
            const float* __restrict__ bval = g.b_val;
            int* __restrict__ icol = g.icol;
            uint16_t* __restrict__ icol16 = g.icol16;
            float* __restrict__ ival = g.ival;
            // fl: 0 an untiled row, 1 a tile of a tiled row, 2 an untiled
            // narrow row; tiles and narrow rows take 16-bit columns
            if (breg) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = lj + u * G;
                    if (jj < nj) {
                        long long slot;
                        if (fl != 1) slot = base + jj;
                        else slot = L.tbase[q * T + (bc[u] >> g.tile_log2) - tfirst] + jlo + jj;
                        const long long w = islot(slot);
                        if (fl) icol16[slot] = (uint16_t)bc[u];
                        else icol[w] = bc[u];
                        ival[w] = __fmul_rn(a, bx[u]);
                    }
                }
                continue;
            }
            for (int jj0 = lj; jj0 < nj; jj0 += 4 * G) {
                int col[4];
                float bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = jj0 + u * G;
                    col[u] = jj < nj ? bidx[jlo + jj] : 0;
                    bv[u] = jj < nj ? bval[jlo + jj] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = jj0 + u * G;
                    if (jj < nj) {
                        long long slot;
                        if (fl != 1) slot = base + jj;
                        else slot = L.tbase[q * T + (col[u] >> g.tile_log2) - tfirst] + jlo + jj;
                        const long long w = islot(slot);
                        if (fl) icol16[slot] = (uint16_t)col[u];
                        else icol[w] = col[u];
                        ival[w] = __fmul_rn(a, bv[u]);
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ int ceil_log2_6(int nj) {
    return min(6, 32 - __clz(max(nj, 1) - 1));
}

// Lanes per A entry (log2): the next power of two >= the B piece, at most a
// wave; pieces longer than a wave get more lanes (up to the workgroup) when
// the item has too few entries to give every wave one, so that no wave of
// the workgroup idles while one writes a long piece.
__device__ __forceinline__ int lanes_log2(int nj, long long ne) {
    int lg = ceil_log2_6(nj);
    while (lg < 8 && (long long)(kBlock >> lg) > ne && (1 << lg) < nj) ++lg;
    return lg;
}

// Binned expansion: persistent workgroups take chunks of kItemChunk
// consecutive work items (the k-bins are contiguous in item order, heavy
// bin first).  With the XCD mapping on, chunk c is served by the XCD c % 8
// so the items of one chunk (mostly one k) share that XCD's L2 copy of B
// row k.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kExpandWaves))) void k_expand(ExpandArgs g,
                                                    unsigned long long* __restrict__ ctr) {
    __shared__ ExpandLds L;
    __shared__ long long sj;
    const int parts = g.swizzle ? 8 : 1;
    const int part = g.swizzle ? (int)(blockIdx.x & 7) : 0;
    const long long nchunks = (g.nitems + kItemChunk - 1) / kItemChunk;
    for (;;) {
        const long long r = next_item(ctr + part, &sj);
        const long long chunk = r * parts + part;
        if (chunk >= nchunks) break;
        const long long first = chunk * kItemChunk;
        const long long last = min(first + kItemChunk, g.nitems);
        long long lo = 0, hi = g.np;  // largest p with kitem_off[p] <= first
        while (hi - lo > 1) {
            const long long mid = (lo + hi) >> 1;
            if (g.kitem_off[mid] <= first) lo = mid;
            else hi = mid;
        }
        long long p = lo;
        for (long long item = first; item < last; ++item) {
            while (g.kitem_off[p + 1] <= item) ++p;
            const int k = g.kperm[p];
            const long long local = item - g.kitem_off[p];
            const int a0 = g.a_ptr[k];
            const int b0 = g.b_ptr[k];
            const long long alen = g.a_ptr[k + 1] - a0;
            const int blen = g.b_ptr[k + 1] - b0;
            long long e_begin, ne;
            int j0, nj;
            if (blen <= g.P) {
                long long epi = g.P / blen;
                if (epi < 1) epi = 1;
                e_begin = a0 + local * epi;
                ne = min(epi, alen - local * epi);
                j0 = 0;
                nj = blen;
            } else {
                const long long jch = (blen + g.P - 1) / g.P;
                e_begin = a0 + local / jch;
                ne = 1;
                j0 = (int)(local % jch) * g.P;
                nj = min(g.P, blen - j0);
            }
            expand_work(g, L, k, e_begin, ne, j0, nj, lanes_log2(nj, ne));
        }
    }
}

// Plain expansion (simpleGEMM): one workgroup per k, 64 lanes per entry.
__global__ __launch_bounds__(kBlock) void k_expand_simple(ExpandArgs g) {
    __shared__ ExpandLds L;
    const int k = blockIdx.x;
    const int a0 = g.a_ptr[k], b0 = g.b_ptr[k];
    const long long ne = g.a_ptr[k + 1] - a0;
    const int nj = g.b_ptr[k + 1] - b0;
    if (ne == 0 || nj == 0) return;
    expand_work(g, L, k, a0, ne, 0, nj, lanes_log2(max(nj, 64), ne));
}

static ExpandArgs expand_args(spg_context* c) {
    ExpandArgs g{};
    g.a_ptr = c->A.ptr;
    g.a_idx = c->A.idx;
    g.a_val = c->A.val;
    g.b_ptr = c->B.ptr;
    g.b_idx = c->B.idx;
    g.b_val = c->B.val;
    g.kperm = (const int*)c->kperm.p;
    g.kitem_off = (const long long*)c->kitem_off.p;
    g.P = c->cfg.item_products;
    g.vstart = (const int*)c->vstart.p;
    g.vbase = (const long long*)c->vbase.p;
    g.vwid = (const int*)c->vwid.p;
    g.cursor = (int*)c->cursor.p;
    g.bsplit = (const int*)c->bsplit.p;
    g.kwin = (const int2*)c->kwin;
    g.T = c->have_long ? c->T : 1;
    g.tile_log2 = c->cfg.tile_log2;
    g.swizzle = c->cfg.xcd_swizzle;
    g.icol = inter_cols(c);
    g.icol16 = inter_cols16(c);
    g.ival = inter_vals(c);
    return g;
}

int launch_expand_simple(spg_context* c) {
    ExpandArgs g = expand_args(c);
    if (c->K > 0) {
        hipLaunchKernelGGL(k_expand_simple, dim3((unsigned)c->K), dim3(kBlock), 0, c->stream, g);
        SPG_HIP(hipGetLastError());
    }
    return SPG_OK;
}

// Sets bit c (c < 0: no bit) of an LDS bitmap.  Products of one k arrive in
// ascending column order, so neighbouring lanes usually hit the same 32-bit
// word; lanes first OR their bits along runs of equal words (segmented
// scan over the wave) and only the last lane of a run issues the LDS
// atomic, instead of up to 64 serialized atomics on one address.
__device__ __forceinline__ void wave_bitmap_or(unsigned* bits, int c) {
    const int lane = lane_id();
    const int w = c >= 0 ? (c >> 5) : -1 - lane;  // unique dummy word for idle lanes
    unsigned m = c >= 0 ? (1u << (c & 31)) : 0u;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int wp = __shfl_up(w, off, 64);
        const unsigned mp = __shfl_up(m, off, 64);
        if (lane >= off && wp == w) m |= mp;
    }
    const int wn = __shfl_down(w, 1, 64);
    if (c >= 0 && (lane == 63 || wn != w)) atomicOr(&bits[w], m);
}

// ---------------------------------------------- row-grouped expansion
// The same (i,k) products as the column-grouped kernels, grouped by C row:
// the work unit is one virtual row, whose slot range is filled front to
// back.  For every k of row i (A transposed to rows) the piece of B row k
// inside the virtual row's column range is scaled by a_ik; piece offsets
// are an exclusive scan of the piece lengths, so no slot is reserved with
// atomics and every store of a wave lands in one contiguous run (HBM write
// efficiency depends on run length; scattered 128-byte runs reach only a
// fraction of the sequential write bandwidth).  A lane finds the piece of
// its output position by a binary search over the 64 piece offsets.
constexpr long long kRowsWaveMax = 16384;  // larger virtual rows: workgroup kernel

struct RowsArgs {
    const int* rptr;
    const int* rk;
    const float* rv;
    const int* b_ptr;
    const int2* bpair;  // B's (column, value bits) pairs
    const int* bsplit;
    const int* vrow_row;
    const int* vstart;
    const long long* vbase;
    const int* vwid;
    int T, W;
    int* icol;
    uint16_t* icol16;  // columns of narrow rows (width <= W)
    float* ival;
};

// Piece [ps, ps+len) of B row k inside tile t (whole row when untiled).
__device__ __forceinline__ void piece_of(const RowsArgs& g, int k, int t, bool tiled, int& ps,
                                         int& len) {
    if (tiled) {
        const int* bs = g.bsplit + (long long)k * (g.T + 1);
        ps = bs[t];
        len = bs[t + 1] - ps;
    } else {
        ps = g.b_ptr[k];
        len = g.b_ptr[k + 1] - ps;
    }
}

__global__ __launch_bounds__(kPT) void k_expand_rows_wave(RowsArgs g, const int* __restrict__ vlist,
                                                           long long nlist) {
    const int wv = threadIdx.x >> 6, lane = lane_id();
    constexpr int NWV = kPT / 64;
    const long long stride = (long long)gridDim.x * NWV;
    const int2* __restrict__ bpair = g.bpair;
    int* __restrict__ icol = g.icol;
    uint16_t* __restrict__ icol16 = g.icol16;
    float* __restrict__ ival = g.ival;
    for (long long j = (long long)blockIdx.x * NWV + wv; j < nlist; j += stride) {
        const int v = vlist[j];
        const int i = g.vrow_row[v];
        const int vs = g.vstart[i];
        const bool tiled = g.vstart[i + 1] - vs > 1;
        const bool nar = g.vwid[v] <= g.W;
        const int t = v - vs;
        const int e1 = g.rptr[i + 1];
        long long dst = g.vbase[v];
        for (int c0 = g.rptr[i]; c0 < e1; c0 += 64) {
            const int e = c0 + lane;
            int ps = 0, len = 0;
            float a = 0.f;
            if (e < e1) {
                piece_of(g, g.rk[e], t, tiled, ps, len);
                a = g.rv[e];
            }
            int inc = len;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(inc, o, 64);
                if (lane >= o) inc += y;
            }
            const int off = inc - len;
            const int total = __shfl(inc, 63, 64);
            // four independent output groups per iteration keep four gathers
            // in flight per lane
            for (int pb = 0; pb < total; pb += 4 * 64) {
                int p[4], lo[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    p[u] = pb + u * 64 + lane;
                    lo[u] = 0;  // largest lane whose piece starts at or before p
                }
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int cand = lo[u] + step;
                        const int oc = __shfl(off, cand, 64);
                        if (oc <= p[u]) lo[u] = cand;
                    }
                }
                int src[4];
                float sa[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    src[u] = __shfl(ps, lo[u], 64) + (p[u] - __shfl(off, lo[u], 64));
                    sa[u] = __shfl(a, lo[u], 64);
                }
                int col[4];
                float bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int2 pr = p[u] < total ? bpair[src[u]] : make_int2(0, 0);
                    col[u] = pr.x;
                    bv[u] = __int_as_float(pr.y);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (p[u] < total) {
                        const long long w = islot(dst + p[u]);
                        if (nar) icol16[dst + p[u]] = (uint16_t)col[u];
                        else icol[w] = col[u];
                        ival[w] = __fmul_rn(sa[u], bv[u]);
                    }
                }
            }
            dst += total;
        }
    }
}

// Virtual rows with more than kRowsWaveMax products: one workgroup each,
// k taken kBlock at a time, offsets scanned across the workgroup and the
// piece of each output found by a binary search in LDS.
__global__ __launch_bounds__(kBlock) void k_expand_rows_wg(RowsArgs g, const int* __restrict__ vlist,
                                                            long long nlist) {
    __shared__ int soff[kBlock + 1];
    __shared__ int sps[kBlock];
    __shared__ float sa[kBlock];
    __shared__ int sh[kBlock / 64];
    for (long long j = blockIdx.x; j < nlist; j += gridDim.x) {
        const int v = vlist[j];
        const int i = g.vrow_row[v];
        const int vs = g.vstart[i];
        const bool tiled = g.vstart[i + 1] - vs > 1;
        const bool nar = g.vwid[v] <= g.W;
        const int t = v - vs;
        const int e1 = g.rptr[i + 1];
        long long dst = g.vbase[v];
        for (int c0 = g.rptr[i]; c0 < e1; c0 += kBlock) {
            const int e = c0 + threadIdx.x;
            int ps = 0, len = 0;
            float a = 0.f;
            if (e < e1) {
                piece_of(g, g.rk[e], t, tiled, ps, len);
                a = g.rv[e];
            }
            int total;
            const int off = block_exscan<int>(len, &total, sh);
            soff[threadIdx.x] = off;
            sps[threadIdx.x] = ps;
            sa[threadIdx.x] = a;
            if (threadIdx.x == 0) soff[kBlock] = total;
            __syncthreads();
            for (int pb = 0; pb < total; pb += 4 * kBlock) {
                int p[4], lo[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    p[u] = pb + u * kBlock + threadIdx.x;
                    lo[u] = 0;
                }
#pragma unroll
                for (int step = kBlock / 2; step > 0; step >>= 1) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (soff[lo[u] + step] <= p[u]) lo[u] += step;
                }
                int col[4];
                float bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int src = sps[lo[u]] + (p[u] - soff[lo[u]]);
                    const int2 pr = p[u] < total ? g.bpair[src] : make_int2(0, 0);
                    col[u] = pr.x;
                    bv[u] = __int_as_float(pr.y);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (p[u] < total) {
                        const long long w = islot(dst + p[u]);
                        if (nar) g.icol16[dst + p[u]] = (uint16_t)col[u];
                        else g.icol[w] = col[u];
                        g.ival[w] = __fmul_rn(sa[lo[u]], bv[u]);
                    }
                }
            }
            __syncthreads();
            dst += total;
        }
    }
}

// Tiled rows with at most kRCMaxTiles tiles, in chunks of E consecutive
// entries (E = min(64, kRCTab / tiles)); a workgroup takes one chunk at a
// time from a queue (the heaviest rows' chunks first; the next chunk is
// fetched while the current one is expanded).  For every chunk and tile,
// the pieces of the chunk's entries inside the tile (ranges of B's split
// points) are laid end to end into one run, reserved with one atomic on the
// tile's cursor, and each wave writes a contiguous share of the chunk's
// (tile, 64-slot block) pairs: the 64 lanes of one store cover 64
// consecutive slots of one run, each lane finding its entry by a binary
// search over the run offsets and gathering its B column and value.  The
// runs of a row's chunks are reserved at nearly the same time, so its lines
// are completed within microseconds, where the binned expansion's short
// pieces leave partial lines written far apart in time (L2 write-back for
// the 23.5 GB of intermediates of R-MAT scale 18: 24.2 GB here, 29.0 GB
// binned).  The price is that B is read once per product (8-byte pairs,
// from L2 or the last-level cache: 25 GB of L2 misses there) instead of
// once per work item: 9.8 ms on that product, against 12.6 ms for the
// binned kernel on 64-product groups.  Lanes past a run's end repeat its
// last product (same slot, same data), so no store is under a lane
// condition.
constexpr int kRCTab = 2048;      // (entry, tile) slot bases of one chunk
constexpr int kRCMaxTiles = 256;  // rows with more tiles: per-vrow kernels
constexpr int kRCMaxE = 64;       // entries per chunk
constexpr int kRCU = 8;           // (tile, block) pairs in flight per wave

// Entries per chunk: the largest power of two <= min(64, kRCTab / Ti), so
// that a tile's row of the chunk table is searched by binary lifting.
__device__ __forceinline__ int rc_chunk_entries(int Ti) {
    const int e = min(kRCMaxE, kRCTab / Ti);
    return 1 << (31 - __clz(e));
}

struct RCArgs {
    const int* rk;
    const float* rv;
    const int2* bpair;  // B's (column, value bits) pairs
    const int* bsplit;
    const long long* vbase;
    int* cursor;
    int T;
    uint16_t* icol16;  // tiles are narrow rows: 16-bit columns
    float* ival;
};

// Chunks of every listed row.
__global__ void k_rc_chunks(const int* __restrict__ rows, long long nrows,
                            const int* __restrict__ rptr, const int* __restrict__ vstart,
                            int* __restrict__ nch) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int i = rows[r];
    const int E = rc_chunk_entries(vstart[i + 1] - vstart[i]);
    nch[r] = (rptr[i + 1] - rptr[i] + E - 1) / E;
}

// Every chunk's (first vrow, tiles, first entry, entries).
__global__ void k_rc_units(const int* __restrict__ rows, long long nrows,
                           const int* __restrict__ rptr, const int* __restrict__ vstart,
                           const int* __restrict__ choff, int4* __restrict__ units) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int i = rows[r];
    const int vs = vstart[i], Ti = vstart[i + 1] - vs;
    const int E = rc_chunk_entries(Ti);
    const int e0 = rptr[i], e1 = rptr[i + 1];
    int u = choff[r];
    for (int c0 = e0; c0 < e1; c0 += E) units[u++] = make_int4(vs, Ti, c0, min(E, e1 - c0));
}

template <bool B32>
__global__ __launch_bounds__(kBlock) void k_expand_rowchunks(RCArgs g, const int4* __restrict__ units,
                                                              int nunits,
                                                              unsigned long long* __restrict__ ctr) {
    // [t][e], row stride E + 1: tile t's run offset of entry e (INT_MAX for
    // e >= ne, so a row is searched with no bound test) and the B position
    // of the entry's piece in the tile
    __shared__ int2 tab[kRCTab + kRCMaxTiles];
    __shared__ long long sbase[kRCMaxTiles];  // first slot of the chunk's run in tile t
    __shared__ int rlen[kRCMaxTiles];         // length of that run
    __shared__ int gsum[kBlock];
    __shared__ int tpb[kRCMaxTiles + 1];      // 64-slot blocks of the tiles, scanned
    __shared__ float sa[kRCMaxE];
    __shared__ int sh[kBlock / 64];
    __shared__ int snext;
    constexpr int NW = kBlock / 64;
    const int wv = uni(threadIdx.x >> 6), lane = lane_id();
    const char* __restrict__ bpair = (const char*)g.bpair;
    // the next unit is taken from the queue (and its description loaded)
    // while the current one is expanded
    if (threadIdx.x == 0) snext = (int)atomicAdd(ctr, 1ull);
    __syncthreads();
    int u = uni(snext);
    int4 inf = u < nunits ? units[u] : make_int4(0, 1, 0, 0);
    while (u < nunits) {
        unsigned long long nx = 0;
        if (threadIdx.x == 0) nx = atomicAdd(ctr, 1ull);
        // Ti in 2 .. kRCMaxTiles
        const int vs = uni(inf.x), Ti = uni(inf.y), c0 = uni(inf.z), ne = uni(inf.w);
        const int E = rc_chunk_entries(Ti), RS = E + 1;
        const int NG = kBlock / Ti;         // entry groups of the prefix pass
        const int EPG = (E + NG - 1) / NG;  // entries per group: <= 8, as E * Ti <= kRCTab
        const int t = threadIdx.x % Ti, eg = threadIdx.x / Ti;
        const int ea = eg * EPG;
        if (threadIdx.x < ne) sa[threadIdx.x] = g.rv[c0 + threadIdx.x];
        if (threadIdx.x == 0) snext = (int)nx;
        // tile t's piece of every entry in [ea, ea + EPG): B start and length
        int xv[8], lv[8], run = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = ea + j;
            xv[j] = 0;
            lv[j] = 0;
            if (eg < NG && j < EPG && e < ne) {
                const int* bs = g.bsplit + (long long)g.rk[c0 + e] * (g.T + 1) + t;
                xv[j] = bs[0];
                lv[j] = bs[1] - xv[j];
            }
            run += lv[j];
        }
        if (eg < NG) gsum[eg * Ti + t] = run;
        __syncthreads();
        const int un = uni(snext);
        const int4 infn = un < nunits ? units[un] : make_int4(0, 1, 0, 0);
        int nbt = 0;
        if (threadIdx.x < Ti) {  // the chunk's run of every tile, reserved
            int sum = 0;
            for (int q = 0; q < NG; ++q) sum += gsum[q * Ti + threadIdx.x];
            long long base = 0;
            if (sum > 0) base = g.vbase[vs + threadIdx.x] + atomicAdd(&g.cursor[vs + threadIdx.x], sum);
            sbase[threadIdx.x] = base;
            rlen[threadIdx.x] = sum;
            nbt = (sum + 63) >> 6;
        }
        if (eg < NG) {  // the entries' run offsets, INT_MAX past the chunk's last entry
            int pre = 0;
            for (int q = 0; q < eg; ++q) pre += gsum[q * Ti + t];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int e = ea + j;
                if (j < EPG && e < E) tab[t * RS + e] = make_int2(e < ne ? pre : INT_MAX, xv[j]);
                pre += lv[j];
            }
        }
        int nbk;
        const int tb = block_exscan<int>(nbt, &nbk, sh);
        if (threadIdx.x < Ti) tpb[threadIdx.x] = tb;
        if (threadIdx.x == 0) tpb[Ti] = nbk;
        __syncthreads();
        // (tile, 64-slot block) pairs, a contiguous share per wave: the lanes
        // of one store cover 64 consecutive slots of one tile's run.  A
        // block's tile, run and first slot are wave-uniform (scalar); each
        // lane finds its entry by binary lifting over the tile's row.
        const int nb = uni(nbk);
        const int per = (nb + NW - 1) / NW;
        const int pb = min(nb, wv * per), pe = min(nb, pb + per);
        if (pb < pe) {
            int cnt = 0;  // the tile of block pb: the last t with tpb[t] <= pb
            for (int t0 = 0; t0 < Ti; t0 += 64) {
                const int tt = t0 + lane;
                cnt += __popcll(__ballot(tt < Ti && tpb[tt] <= pb));
            }
            int tc = cnt - 1;
            int t_beg = uni(tpb[tc]), t_end = uni(tpb[tc + 1]), rl = uni(rlen[tc]);
            long long sb = uni64(sbase[tc]);
            for (int p0 = pb; p0 < pe; p0 += kRCU) {
                int row[kRCU], boff[kRCU], pos[kRCU];
                unsigned ra[kRCU];  // byte offset in tab of the entry found so far
                long long s0[kRCU];
#pragma unroll
                for (int v = 0; v < kRCU; ++v) {
                    const int P = min(p0 + v, pe - 1);
                    while (P >= t_end) {
                        ++tc;
                        t_beg = t_end;
                        t_end = uni(tpb[tc + 1]);
                        rl = uni(rlen[tc]);
                        sb = uni64(sbase[tc]);
                    }
                    row[v] = tc * RS;
                    boff[v] = (P - t_beg) << 6;
                    s0[v] = sb + boff[v];
                    pos[v] = min(boff[v] + lane, rl - 1);
                    ra[v] = (unsigned)row[v] * sizeof(int2);  // entry 0: run offset 0 <= pos
                }
                // the largest entry whose run offset is <= pos
#pragma unroll
                for (int step = kRCMaxE / 2; step > 0; step >>= 1) {
                    if (step < E) {
#pragma unroll
                        for (int v = 0; v < kRCU; ++v) {
                            const unsigned cand = ra[v] + step * (unsigned)sizeof(int2);
                            const int x = *(const int*)((const char*)tab + cand);
                            ra[v] = x <= pos[v] ? cand : ra[v];
                        }
                    }
                }
                int col[kRCU];
                float val[kRCU];
#pragma unroll
                for (int v = 0; v < kRCU; ++v) {
                    const int2 pr = *(const int2*)((const char*)tab + ra[v]);
                    const int ev = (int)(ra[v] / sizeof(int2)) - row[v];
                    const unsigned q = (unsigned)(pr.y + pos[v] - pr.x);
                    const int2 b = B32 ? *(const int2*)(bpair + (q << 3)) : ((const int2*)bpair)[q];
                    col[v] = b.x;
                    val[v] = __fmul_rn(sa[ev], __int_as_float(b.y));
                }
#pragma unroll
                for (int v = 0; v < kRCU; ++v) {
                    // slot s0 + d: its value at word (s0 >> 6) * 128 + o + (o & 64) of
                    // the value array, o = (s0 & 63) + d; its 16-bit column at s0 + d
                    const long long w0 = (s0[v] >> kIGroupLog2) << (kIGroupLog2 + 1);
                    const int d = pos[v] - boff[v];
                    const int o = (int)(s0[v] & (kIGroup - 1)) + d;
                    const unsigned off = (unsigned)(o + (o & kIGroup)) << 2;
                    *(uint16_t*)((char*)(g.icol16 + s0[v]) + ((unsigned)d << 1)) = (uint16_t)col[v];
                    *(float*)((char*)(g.ival + w0) + off) = val[v];
                }
            }
        }
        __syncthreads();  // the unit's LDS tables are free
        u = un;
        inf = infn;
    }
}

__global__ void k_interleave(const int* __restrict__ idx, const float* __restrict__ val, long long n,
                             int2* __restrict__ out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = make_int2(idx[q], __float_as_int(val[q]));
}

// Untiled rows, by chunks of kBlock consecutive A entries (in row order):
// the slot of each entry's first product is known from its row's base and
// the lengths of the row's earlier B rows, so a chunk's products form runs
// of consecutive slots (consecutive untiled rows have adjacent slot ranges)
// that each wave writes 64 slots per store, lanes finding their entry by a
// binary search over the chunk's product offsets, as the row-chunk kernel
// does per tile.  One workgroup per chunk amortizes the per-row chain of
// dependent loads that a wave per row pays.
// An entry's base carries kNarrowBit when its row is narrow (width <= W:
// 16-bit columns).
constexpr long long kNarrowBit = 1LL << 62;

__global__ void k_expand_entry_base(int m, const int* __restrict__ rptr, const int* __restrict__ rk,
                                    const int* __restrict__ b_ptr, const int* __restrict__ vstart,
                                    const long long* __restrict__ vbase,
                                    const int* __restrict__ vwid, int W,
                                    long long* __restrict__ ebase) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int vs = vstart[i];
    if (vstart[i + 1] - vs != 1) return;  // a tiled row: its entries keep -1
    // slots stay below 2^62, so the flag does not touch them
    long long run = vbase[vs] | (vwid[vs] <= W ? kNarrowBit : 0LL);
    const int e1 = rptr[i + 1];
    int e = rptr[i];
    for (; e + 4 <= e1; e += 4) {  // four entries' loads in flight
        int k[4], l[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) k[u] = rk[e + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) l[u] = b_ptr[k[u] + 1] - b_ptr[k[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ebase[e + u] = run;
            run += l[u];
        }
    }
    for (; e < e1; ++e) {
        const int k = rk[e];
        ebase[e] = run;
        run += b_ptr[k + 1] - b_ptr[k];
    }
}

struct EntryArgs {
    const int* rk;
    const float* rv;
    const int* b_ptr;
    const int2* bpair;
    const long long* ebase;
    int* icol;
    uint16_t* icol16;
    float* ival;
};

__global__ __launch_bounds__(kBlock) void k_expand_entrychunks(EntryArgs g, long long nent) {
    __shared__ int xs[kBlock];      // B position of each entry's row
    __shared__ int pref[kBlock];    // each entry's offset in the chunk's products
    __shared__ long long sbo[kBlock];  // slot of product offset 0 for the entry
    __shared__ float sa[kBlock];
    __shared__ int snar[kBlock];       // the entry's row is narrow (16-bit columns)
    __shared__ int sh[kBlock / 64];
    constexpr int NW = kBlock / 64;
    const int wv = threadIdx.x >> 6, lane = lane_id();
    const int2* __restrict__ bpair = g.bpair;
    int* __restrict__ icol = g.icol;
    uint16_t* __restrict__ icol16 = g.icol16;
    float* __restrict__ ival = g.ival;
    const long long nunits = (nent + kBlock - 1) / kBlock;
    for (long long u = blockIdx.x; u < nunits; u += gridDim.x) {
        const long long c0 = u * kBlock;
        const int ne = (int)min((long long)kBlock, nent - c0);
        int len = 0, x = 0, nar = 0;
        long long base = -1;
        float a = 0.f;
        if (threadIdx.x < ne) {
            const long long e = c0 + threadIdx.x;
            base = g.ebase[e];
            if (base >= 0) {  // an entry of an untiled row
                nar = (base & kNarrowBit) != 0;
                base &= kNarrowBit - 1;
                const int k = g.rk[e];
                x = g.b_ptr[k];
                len = g.b_ptr[k + 1] - x;
                a = g.rv[e];
            }
        }
        int sum;
        const int off = block_exscan<int>(len, &sum, sh);  // also: the previous chunk is done
        if (threadIdx.x < ne) {
            xs[threadIdx.x] = x;
            pref[threadIdx.x] = off;
            sbo[threadIdx.x] = base - off;
            sa[threadIdx.x] = a;
            snar[threadIdx.x] = nar;
        }
        __syncthreads();
        const int nbk = (sum + 63) >> 6;
        const int per = (nbk + NW - 1) / NW;
        const int pb = min(nbk, wv * per), pe = min(nbk, pb + per);
        for (int p0 = pb; p0 < pe; p0 += 4) {
            int pos[4], ev[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                pos[v] = min((min(p0 + v, pe - 1) << 6) + lane, sum - 1);
                ev[v] = 0;  // largest entry whose offset is <= pos
            }
#pragma unroll
            for (int step = kBlock / 2; step > 0; step >>= 1) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int cand = ev[v] + step;
                    if (cand < ne && pref[cand] <= pos[v]) ev[v] = cand;
                }
            }
            int col[4];
            float bv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int2 pr = bpair[xs[ev[v]] + pos[v] - pref[ev[v]]];
                col[v] = pr.x;
                bv[v] = __int_as_float(pr.y);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long sl = sbo[ev[v]] + pos[v];
                const long long w = islot(sl);
                if (snar[ev[v]]) icol16[sl] = (uint16_t)col[v];
                else icol[w] = col[v];
                ival[w] = __fmul_rn(sa[ev[v]], bv[v]);
            }
        }
    }
}

// Row-chunk classes: tiled rows with at most kRCMaxTiles tiles, by bound
// (>= 2^20, >= 2^16, the rest), so that the heaviest rows start first.
__global__ void k_rc_key(int m, const int* __restrict__ vstart, const long long* __restrict__ ub,
                         long long max_ub, int* __restrict__ key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int Ti = vstart[i + 1] - vstart[i];
    const long long u = ub[i];
    key[i] = (Ti > 1 && Ti <= kRCMaxTiles && u < max_ub)
                 ? (u >= (1LL << 20) ? 0 : (u >= (1LL << 16) ? 1 : 2))
                 : -1;
}

// The row-chunk work of an expansion: the rows' keys (>= 0: the row is
// expanded in chunks; the binned kernel skips its entries), the chunk table
// and its size.
struct RCPlan {
    const int* rkey = nullptr;
    const int4* units = nullptr;
    int nunits = 0;
};

static int plan_rowchunks(spg_context* c, long long max_ub, RCPlan& pl) {
    const int m = c->m;
    pl = RCPlan{};
    if (!(c->have_long && c->T > 1 && m > 0 && max_ub > 0)) return SPG_OK;
    int* rkey = (int*)scratch_alloc(c, sizeof(int) * (size_t)m);
    int* rlist = (int*)scratch_alloc(c, sizeof(int) * (size_t)m);
    int64_t* roff = (int64_t*)scratch_alloc(c, sizeof(int64_t) * 4);
    if (!rkey || !rlist || !roff) return fail(c, SPG_ENOMEM, "row-chunk scratch");
    hipLaunchKernelGGL(k_rc_key, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, m,
                       (const int*)c->vstart.p, (const long long*)c->ub.p, max_ub, rkey);
    SPG_HIP(hipGetLastError());
    pl.rkey = rkey;
    // heaviest rows first
    SPG_TRY(stable_partition(c, rkey, m, 3, rlist, roff));
    long long roff_h[4];
    SPG_TRY(readback(c, roff, sizeof(roff_h), roff_h));
    const long long nrc = roff_h[3];
    if (nrc <= 0) return SPG_OK;
    int* nch = (int*)scratch_alloc(c, sizeof(int) * (size_t)nrc);
    int* choff = (int*)scratch_alloc(c, sizeof(int) * (size_t)(nrc + 1));
    if (!nch || !choff) return fail(c, SPG_ENOMEM, "row-chunk scratch");
    const unsigned gr = (unsigned)((nrc + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_rc_chunks, dim3(gr), dim3(kBlock), 0, c->stream, (const int*)rlist, nrc,
                       (const int*)c->arow_ptr.p, (const int*)c->vstart.p, nch);
    SPG_HIP(hipGetLastError());
    SPG_TRY(scan_i32(c, nch, choff, nrc));
    int nunits = 0;
    SPG_TRY(readback(c, choff + nrc, sizeof(int), &nunits));
    if (nunits <= 0) return SPG_OK;
    int4* units = (int4*)scratch_alloc(c, sizeof(int4) * (size_t)nunits);
    if (!units) return fail(c, SPG_ENOMEM, "row-chunk scratch");
    hipLaunchKernelGGL(k_rc_units, dim3(gr), dim3(kBlock), 0, c->stream, (const int*)rlist, nrc,
                       (const int*)c->arow_ptr.p, (const int*)c->vstart.p, (const int*)choff, units);
    SPG_HIP(hipGetLastError());
    pl.units = units;
    pl.nunits = nunits;
    return fill<int>(c, (int*)c->cursor.p, 0, c->V);
}

// B's (column, value) pairs interleaved into c->bpair.
static int make_bpair(spg_context* c) {
    const long long nb = c->B.nnz;
    SPG_TRY(ensure(c, c->bpair, sizeof(int2) * (size_t)std::max<long long>(nb, 1)));
    if (nb > 0) {
        hipLaunchKernelGGL(k_interleave, dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, c->B.idx, c->B.val, nb, (int2*)c->bpair.p);
        SPG_HIP(hipGetLastError());
    }
    return SPG_OK;
}

static int launch_rowchunks(spg_context* c, const RCPlan& pl) {
    if (pl.nunits <= 0) return SPG_OK;
    unsigned long long* ctr = (unsigned long long*)c->counters.p + 32;
    SPG_TRY(fill<unsigned long long>(c, ctr, 0ull, 1));
    RCArgs h{};
    h.rk = (const int*)c->arow_k.p;
    h.rv = (const float*)c->arow_v.p;
    h.bpair = (const int2*)c->bpair.p;
    h.bsplit = (const int*)c->bsplit.p;
    h.vbase = (const long long*)c->vbase.p;
    h.cursor = (int*)c->cursor.p;
    h.T = c->T;
    h.icol16 = inter_cols16(c);
    h.ival = inter_vals(c);
    // B's pairs addressed by 32-bit byte offsets while they fit in 4 GiB
    const dim3 grid(persistent_grid(c, pl.nunits, 6));
    if (c->B.nnz <= (1LL << 29))
        hipLaunchKernelGGL(k_expand_rowchunks<true>, grid, dim3(kBlock), 0, c->stream, h, pl.units,
                           pl.nunits, ctr);
    else
        hipLaunchKernelGGL(k_expand_rowchunks<false>, grid, dim3(kBlock), 0, c->stream, h, pl.units,
                           pl.nunits, ctr);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

int launch_expand_binned(spg_context* c) {
    ExpandArgs g = expand_args(c);
    g.np = c->kbin_off_h[SPG_NUM_KBINS];
    g.nitems = c->kitems_bin_h[SPG_NUM_KBINS];
    if (g.nitems <= 0) return SPG_OK;
    // tiled rows with a bound below rowchunk_ub: expanded in chunks first
    RCPlan pl;
    SPG_TRY(plan_rowchunks(c, c->cfg.rowchunk_ub, pl));
    if (pl.nunits > 0) SPG_TRY(make_bpair(c));
    SPG_TRY(launch_rowchunks(c, pl));
    g.rcskip = pl.nunits > 0 ? pl.rkey : nullptr;
    unsigned long long* ctr = (unsigned long long*)c->counters.p + 8;
    SPG_TRY(fill<unsigned long long>(c, ctr, 0ull, 8));
    const long long nchunks = (g.nitems + kItemChunk - 1) / kItemChunk;
    unsigned grid = persistent_grid(c, nchunks, kExpandWaves);
    if (g.swizzle) grid = (grid + 7) & ~7u;  // every XCD part gets workgroups
    hipLaunchKernelGGL(k_expand, dim3(grid), dim3(kBlock), 0, c->stream, g, ctr);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

// Per-vrow kernels: 0 wave, 1 workgroup; vrows of row-chunk rows and empty
// vrows are dropped.
__global__ void k_vclass_expand(long long V, const long long* __restrict__ vub,
                                const int* __restrict__ vrow_row, const int* __restrict__ vstart,
                                int rc, int* __restrict__ cls) {
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const long long u = vub[v];
    const int i = vrow_row[v];
    const int Ti = vstart[i + 1] - vstart[i];
    if (Ti == 1 || (rc && Ti <= kRCMaxTiles)) {  // entry chunks / row chunks
        cls[v] = -1;
        return;
    }
    cls[v] = u == 0 ? -1 : (u <= kRowsWaveMax ? 0 : 1);
}

int launch_expand_rows(spg_context* c) {
    const long long V = c->V;
    if (V <= 0) return SPG_OK;
    // every tiled row of at most kRCMaxTiles tiles is expanded in chunks
    RCPlan pl;
    SPG_TRY(plan_rowchunks(c, LLONG_MAX, pl));
    const int rc = pl.rkey != nullptr;
    SPG_TRY(make_bpair(c));
    hipLaunchKernelGGL(k_vclass_expand, dim3((unsigned)((V + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, V, (const long long*)c->vub.p, (const int*)c->vrow_row.p,
                       (const int*)c->vstart.p, rc, (int*)c->vclass.p);
    SPG_HIP(hipGetLastError());
    SPG_TRY(stable_partition(c, (const int32_t*)c->vclass.p, V, 2, (int32_t*)c->vlist.p,
                             (int64_t*)c->vclass_off.p));
    long long off_h[3];
    SPG_TRY(readback(c, c->vclass_off.p, sizeof(off_h), off_h));
    const long long n0 = off_h[1] - off_h[0], n1 = off_h[2] - off_h[1];
    RowsArgs g{};
    g.rptr = (const int*)c->arow_ptr.p;
    g.rk = (const int*)c->arow_k.p;
    g.rv = (const float*)c->arow_v.p;
    g.b_ptr = c->B.ptr;
    g.bpair = (const int2*)c->bpair.p;
    g.bsplit = (const int*)c->bsplit.p;
    g.vrow_row = (const int*)c->vrow_row.p;
    g.vstart = (const int*)c->vstart.p;
    g.vbase = (const long long*)c->vbase.p;
    g.vwid = (const int*)c->vwid.p;
    g.T = c->T;
    g.W = c->W;
    g.icol = inter_cols(c);
    g.icol16 = inter_cols16(c);
    g.ival = inter_vals(c);
    const int* vl = (const int*)c->vlist.p;
    // untiled rows: the slot of every entry's first product
    const long long nent = c->A.nnz;
    SPG_TRY(ensure(c, c->ebase, sizeof(long long) * (size_t)std::max<long long>(nent, 1)));
    tmark(c, SPG_PHASE_EXPAND, 0);
    if (c->m > 0 && nent > 0) {
        SPG_TRY(fill<long long>(c, (long long*)c->ebase.p, -1LL, nent));
        hipLaunchKernelGGL(k_expand_entry_base, dim3((c->m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, c->m, g.rptr, g.rk, g.b_ptr, g.vstart, g.vbase,
                           g.vwid, g.W, (long long*)c->ebase.p);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(launch_rowchunks(c, pl));
    if (nent > 0) {
        EntryArgs h{};
        h.rk = g.rk;
        h.rv = g.rv;
        h.b_ptr = g.b_ptr;
        h.bpair = g.bpair;
        h.ebase = (const long long*)c->ebase.p;
        h.icol = g.icol;
        h.icol16 = g.icol16;
        h.ival = g.ival;
        const long long nunits = (nent + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_expand_entrychunks, dim3(persistent_grid(c, nunits, 8)), dim3(kBlock), 0,
                           c->stream, h, nent);
        SPG_HIP(hipGetLastError());
    }
    if (n1 > 0) {  // largest rows first
        hipLaunchKernelGGL(k_expand_rows_wg, dim3(persistent_grid(c, n1, 8)), dim3(kBlock), 0,
                           c->stream, g, vl + off_h[1], n1);
        SPG_HIP(hipGetLastError());
    }
    if (n0 > 0) {
        hipLaunchKernelGGL(k_expand_rows_wave,
                           dim3(persistent_grid(c, (n0 + kPT / 64 - 1) / (kPT / 64), 4)), dim3(kPT),
                           0, c->stream, g, vl + off_h[0], n0);
        SPG_HIP(hipGetLastError());
    }
    tmark(c, SPG_PHASE_EXPAND, 1);
    return SPG_OK;
}


__device__ __forceinline__ unsigned hash_slot(int key, int log2h) {
    return ((unsigned)key * 2654435761u) >> (32 - log2h);
}

// Insert key into an LDS open-addressing table; returns the slot and sets
// fresh when the key was not present, or returns -1 when all H slots hold
// other keys (possible only when the table was sized from a prediction; a
// table sized from an exact count or an upper bound is at most half full).
template <int LOG2H>
__device__ __forceinline__ int hash_insert(int* keys, int key, bool& fresh) {
    constexpr int H = 1 << LOG2H;
    unsigned h = hash_slot(key, LOG2H);
    fresh = false;
    for (int probe = 0; probe < H; ++probe) {
        int old = atomicCAS(&keys[h], -1, key);
        if (old == -1) {
            fresh = true;
            return (int)h;
        }
        if (old == key) return (int)h;
        h = (h + 1) & (H - 1);
    }
    return -1;
}

// Completion times of batched products (spgBatchMultiply with
// SPG_BATCH_LATENCY): the kernel that writes a virtual row's final C
// entries raises its product's time to the wall clock once the row is
// written.  t == null: nothing is recorded.
struct Done {
    unsigned long long* t;  // per product, wall-clock ticks
    const int* vprod;       // product of every virtual row
};
__device__ __forceinline__ void mark_done(const Done& d, int v) {
    if (d.t) atomicMax(&d.t[d.vprod[v]], (unsigned long long)wall_clock64());
}

// --------------------------------------------------------------- symbolic
// Classes of a virtual row.  0..7: LDS hash with 2^(6+class) slots; TILE:
// the row's column range fits one tile (bitmap / dense FP64 accumulator);
// RANK: a row over the full column range (n <= 2^18) merged through a
// bitmap of all columns whose prefix popcounts give each column its output
// position (no sort).
constexpr int kClassTile = kNumHashClasses;       // 8: large tile rows (workgroup)
constexpr int kClassRank = kNumHashClasses + 1;   // 9
constexpr int kClassTileS = kNumHashClasses + 2;  // 10: tile rows, count <= 256 (wave)
constexpr int kTileSMax = 256;                    // largest count of the small tile class
constexpr int kClassTileM = kNumHashClasses + 3;  // 11: tile rows, count <= 1024 (wave)
constexpr int kClassRankW = kNumHashClasses + 4;  // 12: rank rows of a narrow window (wave)
constexpr int kTileWords = (1 << kTileMaxLog2) / 32;  // bitmap words of one tile
constexpr int kRankMaxWords = 8192;              // bitmap of 2^18 columns
constexpr int kRankWaveWords = 2048;             // wave rank kernel: windows <= 2^16 columns
constexpr int kRankWaveCap = 1024;               // ... and counts <= 1024
// Rows over a full-width window with more than cfg.rank_min_count (default
// 128) distinct columns take the rank kernel, smaller ones a hash.

static_assert(kNumVClasses >= kClassRankW + 1, "class table too small");

__global__ void k_vclass_symbolic(long long V, const long long* __restrict__ vub,
                                  const int* __restrict__ vwid, int W,
                                  int* __restrict__ cls) {
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const long long u = vub[v];
    if (u == 0) {
        cls[v] = -1;
        return;
    }
    const int width = vwid[v];
    if (width <= W) {
        cls[v] = kClassTile;
        return;
    }
    long long need = 2 * min(u, (long long)width);
    int lg = kHashMinLog2;
    while ((1LL << lg) < need) ++lg;
    // a row of a narrow window (at most 2^16 columns) with more than 256
    // products: counted through a bitmap of its window by one wave
    if (width <= kRankWaveWords * 32 && lg >= 10) {
        cls[v] = kClassRankW;
        return;
    }
    cls[v] = lg - kHashMinLog2;  // lg <= kHashMaxLog2 since u <= long_row_ub
}

// Symbolic count of rows of a window of at most NW * 32 columns (the
// batch's untiled rows of more than 256 products): one wave per row with a
// private bitmap of the window (8 KB), an LDS OR per product and a
// popcount over the window's words, which are cleared as they are counted.
// A workgroup hash per row (2048 slots for up to 1024 products) zeroed its
// table, probed with compare-and-swap and synchronized the workgroup twice
// per row.
template <int NW, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_symbolic_rank_wave(
    const int* __restrict__ vlist, long long nlist, const long long* __restrict__ vbase,
    const int* __restrict__ vlo, const int* __restrict__ vwid, const int* __restrict__ icol,
    long long* __restrict__ vcnt) {
    static_assert(NW % 256 == 0, "four words per lane and pass");
    __shared__ __align__(16) unsigned bitsall[WPB][NW];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    unsigned* bits = bitsall[wv];
    for (int w = lane * 4; w < NW; w += 256) *(uint4*)(bits + w) = make_uint4(0u, 0u, 0u, 0u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const long long stride = (long long)gridDim.x * WPB;
    for (long long j = (long long)blockIdx.x * WPB + wv; j < nlist; j += stride) {
        const int v = vlist[j];
        const long long s0 = vbase[v], s1 = vbase[v + 1];
        const int lo = vlo[v], nw = (vwid[v] + 31) >> 5;  // <= NW (class condition)
        // four products per lane and pass, their loads issued together (a
        // position past the row reads its last product and is masked)
        for (long long sb = s0; sb < s1; sb += 4 * 64) {
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = icol[islot(min(sb + u * 64 + lane, s1 - 1))];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (sb + u * 64 + lane < s1) {
                    const int cc = c[u] - lo;
                    atomicOr(&bits[cc >> 5], 1u << (cc & 31));
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        int mine = 0;
        for (int w = lane * 4; w < nw; w += 256) {  // words past nw are zero
            const uint4 q = *(const uint4*)(bits + w);
            mine += __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
            *(uint4*)(bits + w) = make_uint4(0u, 0u, 0u, 0u);
        }
        mine = wave_sum(mine);
        if (lane == 0) vcnt[v] = mine;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

template <int LOG2H>
__global__ __launch_bounds__(kBlock) void k_symbolic_hash(const int* __restrict__ vlist,
                                                           const long long* __restrict__ vbase,
                                                           const int* __restrict__ icol,
                                                           long long* __restrict__ vcnt) {
    constexpr int H = 1 << LOG2H;
    __shared__ int keys[H];
    __shared__ int cnt;
    const int v = vlist[blockIdx.x];
    for (int h = threadIdx.x; h < H; h += kBlock) keys[h] = -1;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const long long s1 = vbase[v + 1];
    int mine = 0;
    for (long long s = vbase[v] + threadIdx.x; s < s1; s += kBlock) {
        bool fresh;
        hash_insert<LOG2H>(keys, icol[islot(s)], fresh);
        mine += fresh;
    }
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) vcnt[v] = cnt;
}

// Exclusive scan over an NT-thread workgroup (sh: NT/64 slots).
template <typename T, int NT>
__device__ __forceinline__ T block_exscan_n(T x, T* total, T* sh) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    T inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    T wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        T s = sh[i];
        if (i < w) wpre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wpre + inc - x;
}

// Symbolic count of tile-width virtual rows: one wave per row with a
// wave-private bitmap of the tile's columns (no workgroup barriers, so the
// waves of a CU overlap their load latencies).  Products are read four
// 16-bit columns (8 bytes) per lane and load, the next row's
// descriptor is loaded while the current one is counted, and the finished
// bitmap is stored for the merge pass when the row is larger than the small
// tile class.
constexpr int kSymWaves = kPT / 64;

__global__ __launch_bounds__(kPT) void k_symbolic_tile(
    const int* __restrict__ vlist, long long nlist, const long long* __restrict__ vbase,
    const int* __restrict__ vlo, const uint16_t* __restrict__ icol16, long long* __restrict__ vcnt,
    unsigned* __restrict__ tilebits) {
    __shared__ unsigned bitsall[kSymWaves][kTileWords];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    unsigned* bits = bitsall[wv];
    for (int w = lane; w < kTileWords; w += 64) bits[w] = 0u;
    const long long stride = (long long)gridDim.x * kSymWaves;
    long long j = (long long)blockIdx.x * kSymWaves + wv;
    if (j >= nlist) return;
    int v = vlist[j];
    long long s0 = vbase[v], s1 = vbase[v + 1];
    int lo = vlo[v];
    // 8-byte groups of four 16-bit columns past the row's end are clamped to
    // the group holding its last product (rows of the class are non-empty;
    // the column array is padded) and masked where used, so no load sits
    // under a lane-divergent branch; the next groups are in flight while
    // the current ones are applied
    uint2 q[2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
        q[u] = *(const uint2*)(icol16 + min((s0 & ~3LL) + (long long)(lane + u * 64) * 4,
                                            (s1 - 1) & ~3LL));
    for (;;) {
        const long long jn = j + stride;
        const int vn = jn < nlist ? vlist[jn] : -1;
        for (long long sb = s0 & ~3LL;;) {
            const long long nsb = sb + 2 * 4 * 64;
            uint2 qn[2] = {q[0], q[1]};
            if (nsb < s1) {  // uniform
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    qn[u] = *(const uint2*)(icol16 + min(nsb + (long long)(lane + u * 64) * 4,
                                                         (s1 - 1) & ~3LL));
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const long long base = sb + (long long)(lane + u * 64) * 4;
                int cc[4];
                rel4(q[u], lo, cc);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (base + e >= s0 && base + e < s1) {
                        const int c = cc[e];
                        atomicOr(&bits[c >> 5], 1u << (c & 31));
                    }
                }
            }
            q[0] = qn[0];
            q[1] = qn[1];
            sb = nsb;
            if (sb >= s1) break;
        }
        long long s0n = 0, s1n = 0;
        int lon = 0;
        if (vn >= 0) {
            s0n = vbase[vn];
            s1n = vbase[vn + 1];
            lon = vlo[vn];
#pragma unroll
            for (int u = 0; u < 2; ++u)
                q[u] = *(const uint2*)(icol16 + min((s0n & ~3LL) + (long long)(lane + u * 64) * 4,
                                                    (s1n - 1) & ~3LL));
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        static_assert(kTileWords == 4 * 64, "four bitmap words per lane");
        uint4 wq = *(const uint4*)(bits + lane * 4);
        int mine = __popc(wq.x) + __popc(wq.y) + __popc(wq.z) + __popc(wq.w);
        *(uint4*)(bits + lane * 4) = make_uint4(0u, 0u, 0u, 0u);
        mine = wave_sum(mine);
        // rows of the small tile class rebuild their bitmap in the merge
        // from their few products (cheaper than 1 KB stored and loaded)
        if (mine > kTileSMax) *(uint4*)(tilebits + (long long)v * kTileWords + lane * 4) = wq;
        if (lane == 0) vcnt[v] = mine;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (vn < 0) break;
        j = jn;
        v = vn;
        s0 = s0n;
        s1 = s1n;
        lo = lon;
    }
}

__global__ void k_rowptr(const int* __restrict__ vstart, const long long* __restrict__ vout,
                         int m, long long* __restrict__ crow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= m) crow[i] = vout[vstart[i]];
}

// Small wide rows (at most 2^(LOG2H-1) distinct columns, LOG2H <= 9): one
// wave per row and a wave-private LDS hash, so many rows share a CU without
// workgroup barriers.
constexpr int kHashWaveWPB = 8;   // waves per workgroup (4 for 512 slots)

template <int LOG2H, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_symbolic_hash_wave(
    const int* __restrict__ vlist, long long nlist, const long long* __restrict__ vbase,
    const int* __restrict__ icol, long long* __restrict__ vcnt) {
    constexpr int H = 1 << LOG2H;
    __shared__ int keysall[WPB][H];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    int* keys = keysall[wv];
    const long long stride = (long long)gridDim.x * WPB;
    for (long long j = (long long)blockIdx.x * WPB + wv; j < nlist; j += stride) {
        const int v = vlist[j];
        const long long s1 = vbase[v + 1];
        for (int q = lane; q < H; q += 64) keys[q] = -1;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        int mine = 0;
        for (long long s = vbase[v] + lane; s < s1; s += 64) {
            bool fresh;
            hash_insert<LOG2H>(keys, icol[islot(s)], fresh);
            mine += fresh;
        }
        mine = wave_sum(mine);
        if (lane == 0) vcnt[v] = mine;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// Merge of the same rows: keys and FP64 sums in the wave's hash, the occupied
// slots compacted, and each key's output position found by counting the
// smaller keys (at most 256 x 256 comparisons of broadcast LDS reads).
// cnt_out as in k_merge_hash (-1: the predicted table size was too small).
template <int LOG2H, int WPB, bool STAMP>
__global__ __launch_bounds__(64 * WPB) void k_merge_hash_wave(
    const int* __restrict__ vlist, long long nlist, const long long* __restrict__ vbase,
    const int* __restrict__ icol, const float* __restrict__ ival,
    const long long* __restrict__ vout, int* __restrict__ ccol, float* __restrict__ cval,
    long long* __restrict__ cnt_out, const Done dn) {
    constexpr int H = 1 << LOG2H, P = H / 2, SPL = H / 64;  // slots per lane
    static_assert(H >= 64 && H <= 512, "wave hash sizes");
    __shared__ int keysall[WPB][H];
    __shared__ double valsall[WPB][H];
    __shared__ int skall[WPB][P];
    __shared__ double svall[WPB][P];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    int* keys = keysall[wv];
    double* vals = valsall[wv];
    int* sk = skall[wv];
    double* sv = svall[wv];
    const long long stride = (long long)gridDim.x * WPB;
    for (long long j = (long long)blockIdx.x * WPB + wv; j < nlist; j += stride) {
        const int v = vlist[j];
        const long long s1 = vbase[v + 1];
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
            keys[lane + 64 * q] = -1;
            vals[lane + 64 * q] = 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool full = false;
        for (long long s = vbase[v] + lane; s < s1; s += 64) {
            bool fresh;
            const int h = hash_insert<LOG2H>(keys, icol[islot(s)], fresh);
            if (h >= 0)
                unsafeAtomicAdd(&vals[h], (double)ival[islot(s)]);
            else
                full = true;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool occ[SPL];
        int mine = 0;
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
            occ[q] = keys[lane + 64 * q] != -1;
            mine += occ[q];
        }
        int inc = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        const int total = __shfl(inc, 63, 64);
        if (__ballot(full) != 0ull || total > P) {  // only with a predicted table size
            if (lane == 0 && cnt_out) cnt_out[v] = -1;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        int pos = inc - mine;
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
            if (occ[q]) {
                sk[pos] = keys[lane + 64 * q];
                sv[pos] = vals[lane + 64 * q];
                ++pos;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const long long o = vout[v];
        const int lim = (int)min((long long)total, vout[v + 1] - o);
        for (int a = lane; a < total; a += 64) {
            const int key = sk[a];
            int r = 0;
            for (int b = 0; b < total; ++b) r += sk[b] < key;
            if (r < lim) {
                ccol[o + r] = key;
                cval[o + r] = (float)sv[a];
            }
        }
        if (lane == 0 && cnt_out) cnt_out[v] = total;
        if (STAMP && lane == 0) mark_done(dn, v);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

template <int LOG2H>
static void launch_sym_hash(spg_context* c, const int* list, long long cnt) {
    hipLaunchKernelGGL(k_symbolic_hash<LOG2H>, dim3((unsigned)cnt), dim3(kBlock), 0, c->stream,
                       list, (const long long*)c->vbase.p, (const int*)inter_cols(c),
                       (long long*)c->vcnt.p);
}

static int class_offsets(spg_context* c, long long* off) {
    return readback(c, c->vclass_off.p, sizeof(long long) * (kNumVClasses + 1), off);
}

static bool aux_streams(spg_context* c);
static int fork_aux(spg_context* c);
static int join_aux(spg_context* c);
constexpr long long kForkMinProducts = 1LL << 26;

int launch_symbolic(spg_context* c) {
    const long long V = c->V;
    SPG_TRY(fill<long long>(c, (long long*)c->vcnt.p, 0, V));
    SPG_TRY(fill<unsigned long long>(c, (unsigned long long*)c->counters.p, 0ull, 64));
    if (V > 0) {
        hipLaunchKernelGGL(k_vclass_symbolic, dim3((unsigned)((V + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, c->stream, V, (const long long*)c->vub.p,
                           (const int*)c->vwid.p, c->W, (int*)c->vclass.p);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(stable_partition(c, (const int32_t*)c->vclass.p, V, kNumVClasses,
                             (int32_t*)c->vlist.p, (int64_t*)c->vclass_off.p));
    long long off[kNumVClasses + 1];
    SPG_TRY(class_offsets(c, off));
    const int* vl = (const int*)c->vlist.p;
    tmark(c, SPG_PHASE_SYMBOLIC, 0);
    // the tile class first, on the context's stream; the wave bitmap class
    // and the hash classes beside it on the two merge streams (they hold
    // little LDS), the wave bitmap class first (the largest of them)
    const bool par = c->nprod >= kForkMinProducts && aux_streams(c);
    if (par) SPG_TRY(fork_aux(c));
    const hipStream_t main = c->stream;
    int nside = 0;
    const int order[] = {kClassTile, kClassRankW, 0, 1, 2, 3, 4, 5, 6, 7, kClassRank, kClassTileS,
                         kClassTileM};
    static_assert(sizeof(order) / sizeof(order[0]) == kNumVClasses, "every class once");
    static_assert(kNumHashClasses == 8, "hash classes 0..7");
    for (const int b : order) {
        const long long cnt = off[b + 1] - off[b];
        if (cnt <= 0) continue;
        const int* list = vl + off[b];
        if (par && b != kClassTile) c->stream = c->aux[nside++ & 1];
        if (b == kClassRankW) {
            constexpr int WPB = 4;  // 32 KB of bitmaps per workgroup
            hipLaunchKernelGGL((k_symbolic_rank_wave<kRankWaveWords, WPB>),
                               dim3(persistent_grid(c, (cnt + WPB - 1) / WPB, 5)), dim3(64 * WPB), 0,
                               c->stream, list, cnt, (const long long*)c->vbase.p, (const int*)c->vlo.p,
                               (const int*)c->vwid.p, (const int*)inter_cols(c), (long long*)c->vcnt.p);
        } else if (b == kClassTile) {
            // three workgroups per CU (24 waves) leave room for the classes
            // beside it (four: batch symbolic 3.78 instead of 3.53 ms)
            hipLaunchKernelGGL(k_symbolic_tile,
                               dim3(persistent_grid(c, (cnt + kSymWaves - 1) / kSymWaves, 3)),
                               dim3(kPT), 0, c->stream, list, cnt, (const long long*)c->vbase.p,
                               (const int*)c->vlo.p,
                               (const uint16_t*)inter_cols16(c), (long long*)c->vcnt.p,
                               (unsigned*)c->tilebits.p);
        } else {
            const dim3 gw(persistent_grid(c, (cnt + kHashWaveWPB - 1) / kHashWaveWPB, 4));
            const dim3 bw(64 * kHashWaveWPB);
            const long long* vb = (const long long*)c->vbase.p;
            const int* ic = (const int*)inter_cols(c);
            long long* vc = (long long*)c->vcnt.p;
            switch (b + kHashMinLog2) {
                case 6:
                    hipLaunchKernelGGL((k_symbolic_hash_wave<6, kHashWaveWPB>), gw, bw, 0,
                                       c->stream, list, cnt, vb, ic, vc);
                    break;
                case 7:
                    hipLaunchKernelGGL((k_symbolic_hash_wave<7, kHashWaveWPB>), gw, bw, 0,
                                       c->stream, list, cnt, vb, ic, vc);
                    break;
                case 8:
                    hipLaunchKernelGGL((k_symbolic_hash_wave<8, kHashWaveWPB>), gw, bw, 0,
                                       c->stream, list, cnt, vb, ic, vc);
                    break;
                case 9:
                    hipLaunchKernelGGL((k_symbolic_hash_wave<9, kHashWaveWPB>), gw, bw, 0,
                                       c->stream, list, cnt, vb, ic, vc);
                    break;
                case 10: launch_sym_hash<10>(c, list, cnt); break;
                case 11: launch_sym_hash<11>(c, list, cnt); break;
                case 12: launch_sym_hash<12>(c, list, cnt); break;
                case 13: launch_sym_hash<13>(c, list, cnt); break;
                default: c->stream = main; return fail(c, SPG_EINVAL, "bad symbolic class");
            }
        }
        c->stream = main;
        SPG_HIP(hipGetLastError());
    }
    if (par) SPG_TRY(join_aux(c));
    tmark(c, SPG_PHASE_SYMBOLIC, 1);
    SPG_TRY(scan_i64(c, (const int64_t*)c->vcnt.p, (int64_t*)c->vout.p, V));
    return SPG_OK;
}

int launch_rowptr(spg_context* c) {
    hipLaunchKernelGGL(k_rowptr, dim3((c->m + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       c->stream, (const int*)c->vstart.p, (const long long*)c->vout.p, c->m,
                       (long long*)c->crow.p);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

// ------------------------------------------------------------------ merge
// src/mm.cu:366-404 with D1-D4 fixed.  FP64 accumulators in LDS keyed by
// column, first touch by occupancy (bitmap or hash key, never by value), and
// columns emitted in ascending order directly at their final CSR position
// (the symbolic pass fixed every row's offset).
__global__ void k_vclass_merge(long long V, const long long* __restrict__ vcnt,
                               const int* __restrict__ vwid, int W, int rank_min,
                               int* __restrict__ cls) {
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const long long cnt = vcnt[v];
    if (cnt == 0) {
        cls[v] = -1;
        return;
    }
    const int width = vwid[v];
    if (width <= W) {
        cls[v] = cnt <= kTileSMax ? kClassTileS : (cnt <= 1024 ? kClassTileM : kClassTile);
        return;
    }
    if (cnt > rank_min && width <= kRankWaveWords * 32 && cnt <= kRankWaveCap) {
        cls[v] = kClassRankW;
        return;
    }
    if (cnt > rank_min && width <= kRankMaxWords * 32 && cnt <= kHashMaxCount) {
        cls[v] = kClassRank;
        return;
    }
    int lg = kHashMinLog2;
    while ((1LL << lg) < 2 * cnt) ++lg;
    cls[v] = lg - kHashMinLog2;
}

// cnt_out (may be null) receives the row's distinct-column count, or -1
// when the table (sized from a predicted count) could not hold the row; the
// row's output is then incomplete and the caller merges it again.
template <int LOG2H, bool STAMP>
__global__ __launch_bounds__(kBlock) void k_merge_hash(
    const int* __restrict__ vlist, const long long* __restrict__ vbase,
    const int* __restrict__ icol, const float* __restrict__ ival,
    const long long* __restrict__ vout, int* __restrict__ ccol, float* __restrict__ cval,
    long long* __restrict__ cnt_out, const Done dn) {
    constexpr int H = 1 << LOG2H;
    constexpr int P = H / 2;  // at most H/2 distinct keys are sorted
    __shared__ int keys[H];
    __shared__ double vals[H];
    __shared__ int skey[P];
    __shared__ double sval[P];
    __shared__ int cnt;
    __shared__ int full;
    const int v = vlist[blockIdx.x];
    for (int h = threadIdx.x; h < H; h += kBlock) {
        keys[h] = -1;
        vals[h] = 0.0;
    }
    if (threadIdx.x == 0) {
        cnt = 0;
        full = 0;
    }
    __syncthreads();
    const long long s1 = vbase[v + 1];
    for (long long s = vbase[v] + threadIdx.x; s < s1; s += kBlock) {
        bool fresh;
        const int h = hash_insert<LOG2H>(keys, icol[islot(s)], fresh);
        if (h >= 0)
            unsafeAtomicAdd(&vals[h], (double)ival[islot(s)]);
        else
            full = 1;
    }
    __syncthreads();
    for (int h = threadIdx.x; h < H; h += kBlock) {
        if (keys[h] != -1) {
            const int q = atomicAdd(&cnt, 1);
            if (q < P) {
                skey[q] = keys[h];
                sval[q] = vals[h];
            }
        }
    }
    __syncthreads();
    if (full || cnt > P) {  // only with a predicted table size
        if (threadIdx.x == 0 && cnt_out) cnt_out[v] = -1;
        return;
    }
    const int c = cnt;
    if (threadIdx.x == 0 && cnt_out) cnt_out[v] = c;
    int np = 1;
    while (np < c) np <<= 1;
    for (int q = c + threadIdx.x; q < np; q += kBlock) skey[q] = 0x7fffffff;
    __syncthreads();
    // bitonic sort of (skey, sval), np <= P
    for (int k = 2; k <= np; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np; i += kBlock) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const int ki = skey[i], kj = skey[ixj];
                    const bool up = (i & k) == 0;
                    if ((ki > kj) == up) {
                        skey[i] = kj;
                        skey[ixj] = ki;
                        const double t = sval[i];
                        sval[i] = sval[ixj];
                        sval[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    // never write past the row's slot range fixed by the symbolic pass
    const long long o = vout[v];
    const int lim = (int)min((long long)c, vout[v + 1] - o);
    for (int q = threadIdx.x; q < lim; q += kBlock) {
        ccol[o + q] = skey[q];
        cval[o + q] = (float)sval[q];
    }
    if (STAMP && threadIdx.x == 0) mark_done(dn, v);
}

// Tile-width virtual rows: dense FP64 accumulator over the tile's columns
// plus an occupancy bitmap (stored by the symbolic pass, or built here with
// BUILD when none ran).  Persistent workgroups walk the list with a fixed
// stride; each thread loads two groups of four consecutive products per
// pass (16-byte loads) and the next pass's groups are in flight while the
// current ones are accumulated.  A row's descriptor is loaded one row ahead
// and its first products while the row before it is emitted.  The emit
// gives each wave 32 consecutive bitmap words (1024 columns) held in its
// lanes: 64 columns at a time, the two words form a scalar mask, a column's
// rank is the wave's base plus the mask bits below its lane, and the set
// columns are written to consecutive output positions.  Only touched
// accumulators are cleared (while being emitted).
template <bool BUILD, bool STAMP>
__global__ __launch_bounds__(kPT) void k_merge_tile(
    const int* __restrict__ vlist, long long nlist, const long long* __restrict__ vbase,
    const int* __restrict__ vlo, const int* __restrict__ vwid,
    const unsigned* __restrict__ tilebits, const uint16_t* __restrict__ icol16,
    const float* __restrict__ ival, const long long* __restrict__ vout, int* __restrict__ ccol,
    float* __restrict__ cval, long long* __restrict__ cnt_out, const Done dn) {
    constexpr int WMAX = 1 << kTileMaxLog2;
    constexpr int NW = WMAX / 32;
    constexpr int NWV = kPT / 64;
    static_assert(NW <= kPT, "one bitmap word per thread");
    static_assert(NW == 32 * NWV, "32 bitmap words per wave");
    __shared__ double acc[WMAX];
    __shared__ unsigned bits[NW];
    __shared__ int wtot[NWV];  // set columns of each wave's words
    for (int x = threadIdx.x; x < WMAX; x += kPT) acc[x] = 0.0;
    for (int w = threadIdx.x; w < NW; w += kPT) bits[w] = 0u;
    __syncthreads();
    long long j = blockIdx.x;
    if (j >= nlist) return;
    const long long G = gridDim.x;
    const int wv = threadIdx.x >> 6, lane = lane_id();
    int v = vlist[j];
    long long s0 = vbase[v], s1 = vbase[v + 1], o0 = vout[v], lim = vout[v + 1];
    int lo = vlo[v], nw = (vwid[v] + 31) >> 5;
    // the next row's descriptor, one row ahead
    int vn = j + G < nlist ? vlist[j + G] : -1;
    long long s0n = 0, s1n = 0, o0n = 0, limn = 0;
    int lon = 0, nwn = 0;
    if (vn >= 0) {
        s0n = vbase[vn];
        s1n = vbase[vn + 1];
        o0n = vout[vn];
        limn = vout[vn + 1];
        lon = vlo[vn];
        nwn = (vwid[vn] + 31) >> 5;
    }
    // rows of the class are non-empty; a group of four products past the
    // row's end is clamped to the group holding its last product (the
    // intermediates are padded to whole groups), and its lanes are masked
    // where used.  D groups of four products per thread and pass (2 * 12 KB
    // in flight per workgroup while the previous pass is applied)
    constexpr int D = 2;
    uint2 qc[D];
    float4 qv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const long long b0 =
            min((s0 & ~3LL) + (long long)(d * kPT + threadIdx.x) * 4, (s1 - 1) & ~3LL);
        qc[d] = *(const uint2*)(icol16 + b0);
        qv[d] = *(const float4*)(ival + islot(b0));
    }
    for (;;) {
        const long long jnn = j + 2 * G;
        const int vnn = vn >= 0 && jnn < nlist ? vlist[jnn] : -1;
        unsigned wb = 0u;  // this row's bitmap word (exact path)
        if (!BUILD && threadIdx.x < nw) wb = tilebits[(long long)v * kTileWords + threadIdx.x];
        for (long long sb = s0 & ~3LL;;) {
            const long long nsb = sb + D * 4 * kPT;
            uint2 nc[D];
            float4 nv[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                nc[d] = qc[d];
                nv[d] = qv[d];
            }
            if (nsb < s1) {  // uniform: the next groups of the row, in flight
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const long long nb =
                        min(nsb + (long long)(d * kPT + threadIdx.x) * 4, (s1 - 1) & ~3LL);
                    nc[d] = *(const uint2*)(icol16 + nb);
                    nv[d] = *(const float4*)(ival + islot(nb));
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const long long bb = sb + (long long)(d * kPT + threadIdx.x) * 4;
                int cc[4];
                rel4(qc[d], lo, cc);
                const float xx[4] = {qv[d].x, qv[d].y, qv[d].z, qv[d].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (bb + e >= s0 && bb + e < s1) {
                        const int c = cc[e];
                        if constexpr (BUILD) atomicOr(&bits[c >> 5], 1u << (c & 31));
                        unsafeAtomicAdd(&acc[c], (double)xx[e]);
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                qc[d] = nc[d];
                qv[d] = nv[d];
            }
            sb = nsb;
            if (sb >= s1) break;
        }
        if constexpr (!BUILD) {
            if (threadIdx.x < NW) bits[threadIdx.x] = wb;  // zero past the window
        }
        __syncthreads();  // the row's accumulators and bitmap are complete
        if (vn >= 0) {  // the next row's first products, in flight during the emit
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const long long nb = min((s0n & ~3LL) + (long long)(d * kPT + threadIdx.x) * 4,
                                         (s1n - 1) & ~3LL);
                qc[d] = *(const uint2*)(icol16 + nb);
                qv[d] = *(const float4*)(ival + islot(nb));
            }
        }
        // the wave's 32 words, one per lane (lanes 32..63 hold zero)
        const unsigned bw = lane < 32 ? bits[wv * 32 + lane] : 0u;
        if constexpr (BUILD) {
            if (lane < 32) bits[wv * 32 + lane] = 0u;  // read only by this wave; the next row ORs into it
        }
        int wc = __popc(bw);
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) wc += __shfl_xor(wc, o, 64);
        if (lane == 0) wtot[wv] = wc;
        __syncthreads();
        long long ob = o0;
        int tot = 0;
#pragma unroll
        for (int w = 0; w < NWV; ++w) {
            const int x = wtot[w];
            if (w < wv) ob += x;
            tot += x;
        }
        if (cnt_out && threadIdx.x == 0) cnt_out[v] = tot;
        ob = uni64(ob);
        // the wave's outputs from a scalar base, 32-bit ranks per lane; after
        // a symbolic pass the row's range holds every column, a predicted
        // capacity (BUILD) may cut it short
        int* __restrict__ cw = ccol + ob;
        float* __restrict__ vw = cval + ob;
        const int room = BUILD ? (int)max(0LL, min((long long)INT_MAX, lim - ob)) : INT_MAX;
        const int ncol = uni(nw) * 32, wcol = wv * 1024;
        int rb = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (wcol + 64 * i >= ncol) break;  // past the row's window
            const unsigned long long m =
                ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)bw, 2 * i + 1) << 32) |
                (unsigned)__builtin_amdgcn_readlane((int)bw, 2 * i);
            if (m == 0ull) continue;
            if ((m >> lane) & 1ull) {
                const int col = wcol + 64 * i + lane;
                const int r = rb + (int)__popcll(m & lanemask_lt());
                const double a = acc[col];
                acc[col] = 0.0;
                if (!BUILD || r < room) {
                    cw[r] = lo + col;
                    vw[r] = (float)a;
                }
            }
            rb += (int)__popcll(m);
        }
        __syncthreads();  // emitted: accumulators cleared, bitmap and wtot free
        if (STAMP && threadIdx.x == 0) mark_done(dn, v);
        if (vn < 0) break;
        j += G;
        v = vn;
        s0 = s0n;
        s1 = s1n;
        o0 = o0n;
        lim = limn;
        lo = lon;
        nw = nwn;
        vn = vnn;
        if (vn >= 0) {
            s0n = vbase[vn];
            s1n = vbase[vn + 1];
            o0n = vout[vn];
            limn = vout[vn + 1];
            lon = vlo[vn];
            nwn = (vwid[vn] + 31) >> 5;
        }
    }
}

// Tile-width virtual rows with at most CAP columns: one wave per row, no
// workgroup barriers.  The row's column bitmap (stored by the symbolic
// pass, or built here from the products when BUILD) and a prefix popcount
// per word give every column its rank, i.e. its output position; products
// are accumulated in FP64 at acc[rank], the column's offset in the tile is
// recorded at col[rank] (16 bits), and the row is written out with fully
// coalesced stores.  With
// BUILD, cnt_out receives the row's distinct-column count (which may exceed
// its output capacity; only the capacity is written).  Rows are
// software-pipelined: the next row's descriptor is loaded while the current
// one is accumulated, and its first 256 products (and its stored bitmap)
// while the current one is emitted, so a row of the small class costs about
// one memory round trip; a row whose products fit one pass builds its
// bitmap from the products already in registers.
constexpr int kTileSWPB = 3;  // waves per workgroup of the small tile class

template <int CAP, int WPB, bool BUILD, bool STAMP>
__global__ __launch_bounds__(64 * WPB) void k_merge_tile_wave(
    const int* __restrict__ vlist, long long nlist, unsigned long long* __restrict__ ctr,
    const long long* __restrict__ vbase,
    const int* __restrict__ vlo, const unsigned* __restrict__ tilebits,
    const uint16_t* __restrict__ icol16, const float* __restrict__ ival,
    const long long* __restrict__ vout, int* __restrict__ ccol, float* __restrict__ cval,
    long long* __restrict__ cnt_out, const Done dn) {
    __shared__ double accall[WPB][CAP];
    __shared__ unsigned short colall[WPB][CAP];
    __shared__ unsigned bitsall[WPB][kTileWords];
    __shared__ unsigned short preall[WPB][kTileWords];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    double* acc = accall[wv];
    unsigned short* colr = colall[wv];
    unsigned* bits = bitsall[wv];
    unsigned short* pre = preall[wv];
    // Rows in chunks of R: a wave's first chunk by its index, the others
    // from the counter *ctr (zero at launch), so that waves that start late
    // (while the dense class holds the LDS) take less; the wave's next chunk
    // is claimed when it enters the current one.
    constexpr int R = CAP <= 256 ? 16 : 8;
    const long long nwave = (long long)gridDim.x * WPB;
    auto ticket = [&]() -> long long {
        unsigned t = 0;
        if (lane == 0) t = (unsigned)atomicAdd(ctr, 1ull);
        return (nwave + (long long)__builtin_amdgcn_readfirstlane(t)) * R;
    };
    // f: the last row handed out; fe: the end of its chunk; fnext: the next chunk
    long long f = ((long long)blockIdx.x * WPB + wv) * R;
    if (f >= nlist) return;
    long long fe = min(f + R, nlist), fnext = ticket();
    auto next_row = [&]() -> long long {  // the row after f, or -1
        if (f < 0) return -1;
        if (++f < fe) return f;
        f = fnext;
        if (f >= nlist) return f = -1;
        fe = min(f + R, nlist);
        fnext = ticket();
        return f;
    };
    long long j = f;
    // row j: descriptor, first products and (stored) bitmap.  Groups of four
    // products past the row's end are clamped to the group holding its last
    // product (rows of the class are non-empty; the intermediates are padded
    // to whole groups) and masked where used, so no load sits under a
    // lane-divergent branch.
    int v = vlist[j];
    long long s0 = vbase[v], s1 = vbase[v + 1], o0 = vout[v], o1 = vout[v + 1];
    int lo = vlo[v];
    uint2 qc;
    float4 qv;
    uint4 wq = make_uint4(0u, 0u, 0u, 0u);
    {
        const long long ga = min((s0 & ~3LL) + (long long)lane * 4, (s1 - 1) & ~3LL);
        qc = *(const uint2*)(icol16 + ga);
        qv = *(const float4*)(ival + islot(ga));
        if constexpr (!BUILD) wq = *(const uint4*)(tilebits + (long long)v * kTileWords + lane * 4);
    }
    long long jn = next_row();
    int vn = jn >= 0 ? vlist[jn] : -1;
    for (;;) {
        // the next row's descriptor, in flight during this row
        long long s0n = 0, s1n = 0, o0n = 0, o1n = 0;
        int lon = 0;
        if (vn >= 0) {
            s0n = vbase[vn];
            s1n = vbase[vn + 1];
            o0n = vout[vn];
            o1n = vout[vn + 1];
            lon = vlo[vn];
        }
        const long long jnn = vn >= 0 ? next_row() : -1;
        const int vnn = jnn >= 0 ? vlist[jnn] : -1;
        const long long sfirst = s0 & ~3LL, glast = (s1 - 1) & ~3LL;
        int cnt = (int)min((long long)CAP, o1 - o0);
        if constexpr (BUILD) {
            *(uint4*)(bits + lane * 4) = make_uint4(0u, 0u, 0u, 0u);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (long long sb = sfirst; sb < s1; sb += 4 * 64) {
                const long long base = sb + (long long)lane * 4;
                uint2 q = qc;  // the first pass: the products already loaded
                if (sb != sfirst) q = *(const uint2*)(icol16 + min(base, glast));  // uniform
                int cc[4];
                rel4(q, lo, cc);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (base + e >= s0 && base + e < s1) {
                        const int c = cc[e];
                        atomicOr(&bits[c >> 5], 1u << (c & 31));
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            wq = *(const uint4*)(bits + lane * 4);
        }
        const int p0 = __popc(wq.x), p1 = __popc(wq.y), p2 = __popc(wq.z), p3 = __popc(wq.w);
        const int mine = p0 + p1 + p2 + p3;
        int inc = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        const int ex = inc - mine;
        if constexpr (BUILD) {
            const int total = __shfl(inc, 63, 64);
            if (cnt_out && lane == 0) cnt_out[v] = total;
            cnt = min(cnt, total);  // the capacity may exceed the count
        } else {
            *(uint4*)(bits + lane * 4) = wq;
        }
        pre[lane * 4 + 0] = (unsigned short)ex;
        pre[lane * 4 + 1] = (unsigned short)(ex + p0);
        pre[lane * 4 + 2] = (unsigned short)(ex + p0 + p1);
        pre[lane * 4 + 3] = (unsigned short)(ex + p0 + p1 + p2);
        for (int q = lane; q < cnt; q += 64) acc[q] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (long long sb = sfirst;;) {
            const long long base = sb + (long long)lane * 4;
            const long long nsb = sb + 4 * 64;
            uint2 nc = qc;
            float4 nv = qv;
            if (nsb < s1) {  // uniform
                const long long ga = min(nsb + (long long)lane * 4, glast);
                nc = *(const uint2*)(icol16 + ga);
                nv = *(const float4*)(ival + islot(ga));
            }
            int cc[4];
            rel4(qc, lo, cc);
            const float xx[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (base + e >= s0 && base + e < s1) {
                    const int c = cc[e];
                    const int w = c >> 5;
                    const int r = pre[w] + __popc(bits[w] & ((1u << (c & 31)) - 1u));
                    if (r < cnt) {
                        unsafeAtomicAdd(&acc[r], (double)xx[e]);
                        colr[r] = (unsigned short)c;
                    }
                }
            }
            qc = nc;
            qv = nv;
            sb = nsb;
            if (sb >= s1) break;
        }
        // the next row's first products and stored bitmap, in flight during the emit
        if (vn >= 0) {
            const long long ga = min((s0n & ~3LL) + (long long)lane * 4, (s1n - 1) & ~3LL);
            qc = *(const uint2*)(icol16 + ga);
            qv = *(const float4*)(ival + islot(ga));
            if constexpr (!BUILD) wq = *(const uint4*)(tilebits + (long long)vn * kTileWords + lane * 4);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int q = lane; q < cnt; q += 64) {
            ccol[o0 + q] = lo + (int)colr[q];
            cval[o0 + q] = (float)acc[q];
        }
        if (STAMP && lane == 0) mark_done(dn, v);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (vn < 0) break;
        v = vn;
        s0 = s0n;
        s1 = s1n;
        o0 = o0n;
        o1 = o1n;
        lo = lon;
        jn = jnn;
        vn = vnn;
    }
}

// Rows over the full column range (n <= 2^18) with 128 < count <= 4096:
// pass 1 sets the row's column bitmap, a prefix popcount per word gives
// every column its rank, pass 2 accumulates into acc[rank] in FP64, and the
// columns are emitted from the bitmap in ascending order (no sort).
constexpr int kRankChunk = 8;  // rows taken per counter increment

template <bool STAMP>
__global__ __launch_bounds__(kPT) void k_merge_rank(
    const int* __restrict__ vlist, long long nlist, unsigned long long* __restrict__ ctr,
    const long long* __restrict__ vbase, const int* __restrict__ vlo,
    const int* __restrict__ icol, const float* __restrict__ ival,
    const long long* __restrict__ vout, int* __restrict__ ccol, float* __restrict__ cval,
    long long* __restrict__ cnt_out, const Done dn) {
    __shared__ unsigned bits[kRankMaxWords];
    __shared__ unsigned short pre[kRankMaxWords / 2];  // rank before each word pair
    __shared__ double acc[kHashMaxCount];
    __shared__ int sh[kPT / 64];
    __shared__ long long sj;
    __shared__ int smin, smax;  // the row's smallest and largest window column
    for (int w = threadIdx.x; w < kRankMaxWords; w += kPT) bits[w] = 0u;
    if (threadIdx.x == 0) {
        smin = INT_MAX;
        smax = -1;
    }
    // rows in chunks of kRankChunk from a global counter (few atomics, and
    // rows of uneven size still balance)
    for (;;) {
        const long long chunk = next_item(ctr, &sj);
        const long long j0 = chunk * kRankChunk;
        if (j0 >= nlist) break;
        const long long j1 = min(j0 + kRankChunk, nlist);
        for (long long j = j0; j < j1; ++j) {
            const int v = vlist[j];
            const long long s0 = vbase[v], s1 = vbase[v + 1];
            const long long o = vout[v];
            const int cnt = (int)min((long long)kHashMaxCount, vout[v + 1] - o);
            const int lo = vlo[v];                  // columns of the row: [lo, lo + width)
            int cmin = INT_MAX, cmax = -1;
            // trip count uniform across the workgroup (the OR below shuffles);
            // positions past the row read its last product (rows of the
            // class are non-empty) and are masked after the loads, so the
            // loads do not wait for each other
            for (long long sb = s0; sb < s1; sb += 4 * kPT) {
                int c[4];
    #pragma unroll
                for (int u = 0; u < 4; ++u)
                    c[u] = icol[islot(min(sb + threadIdx.x + (long long)u * kPT, s1 - 1))];
    #pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const long long q = sb + threadIdx.x + (long long)u * kPT;
                    const int cu = q < s1 ? c[u] - lo : -1;
                    wave_bitmap_or(bits, cu);
                    if (cu >= 0) {
                        cmin = min(cmin, cu);
                        cmax = max(cmax, cu);
                    }
                }
            }
            cmin = wave_min(cmin);
            cmax = wave_max(cmax);
            if (lane_id() == 0) {
                atomicMin(&smin, cmin);
                atomicMax(&smax, cmax);
            }
            for (int q = threadIdx.x; q < cnt; q += kPT) acc[q] = 0.0;
            __syncthreads();
            // only the words of the row's column span [wlo, whi) are scanned,
            // emitted and cleared (a row of few columns in a wide window
            // costs its span, not the window); an even number of words per
            // thread, pairs aligned
            const int wlo = (smin >> 5) & ~1, whi = (smax >> 5) + 1;
            const int wpt = (((whi - wlo + kPT - 1) / kPT) + 1) & ~1;
            const int w0 = wlo + threadIdx.x * wpt;
            const int w1 = min(w0 + wpt, whi);
            int mine = 0;
            for (int w = w0; w < w1; ++w) mine += __popc(bits[w]);
            int tot;
            int run = block_exscan_n<int, kPT>(mine, &tot, sh);
            if (cnt_out && threadIdx.x == 0) cnt_out[v] = tot;
            for (int w = w0; w < w1; w += 2) {
                pre[w >> 1] = (unsigned short)run;
                run += __popc(bits[w]) + __popc(bits[w + 1]);  // bits past whi are zero
            }
            __syncthreads();
            for (long long s = s0 + threadIdx.x; s < s1; s += 4 * kPT) {
                int c[4];
                float x[4];
    #pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const long long q = islot(min(s + (long long)u * kPT, s1 - 1));
                    c[u] = icol[q];
                    x[u] = ival[q];
                }
    #pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (s + (long long)u * kPT < s1) {
                        const int cu = c[u] - lo;
                        const int w = cu >> 5;
                        const int r = pre[w >> 1] + ((w & 1) ? __popc(bits[w - 1]) : 0) +
                                      __popc(bits[w] & ((1u << (cu & 31)) - 1u));
                        if (r < cnt) unsafeAtomicAdd(&acc[r], (double)x[u]);
                    }
                }
            }
            __syncthreads();
            // emit: word pairs interleaved over the threads (neighbouring
            // threads read neighbouring accumulators and write neighbouring
            // output positions); a thread clears the pair it emitted
            if (threadIdx.x == 0) {  // every thread has read the span (before the scan's barriers)
                smin = INT_MAX;
                smax = -1;
            }
            for (int pw = (wlo >> 1) + threadIdx.x; 2 * pw < whi; pw += kPT) {
                int r = pre[pw];
    #pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int w = 2 * pw + h;
                    unsigned b = bits[w];  // words past whi are zero
                    while (b) {
                        const int bit = __ffs(b) - 1;
                        b &= b - 1;
                        if (r < cnt) {
                            ccol[o + r] = lo + (w << 5) + bit;
                            cval[o + r] = (float)acc[r];
                        }
                        ++r;
                    }
                    bits[w] = 0u;
                }
            }
            __syncthreads();
            if (STAMP && threadIdx.x == 0) mark_done(dn, v);
        }
    }
}

// Rank rows whose window is at most 2^16 columns wide and whose count is at
// most 1024, one wave per row: the same bitmap / prefix-popcount / FP64
// accumulation as k_merge_rank, but a wave owns its bitmap (8 KB), so a CU
// keeps several rows in flight instead of one workgroup-synchronized row.
// A wave takes G <= 64 consecutive rows of the list (G = 64 when there are
// rows enough to fill the chip): each lane loads one row's descriptor (one
// memory round trip for G rows), and while a row is merged the next row's
// products are already being loaded (16 per lane held in registers; a
// longer row reads the rest in both passes).  No load sits under a
// lane-divergent branch and no loaded value is used before the row that
// needs it, so the loads issue back to back and are waited for once.
// Ranks come from the pair prefixes, and the columns are kept beside the
// accumulators so both outputs are written coalesced.  cnt_out (may be
// null) receives the row's count; a row whose count exceeds its output
// range writes only the range's first part.
constexpr int kRankWavePer = 16;  // products per lane held in registers

__device__ __forceinline__ long long readlane64(long long x, int r) {
    const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)x, r);
    const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)x >> 32), r);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// The first 16 * 64 products of a non-empty row [s0, s1): lane l gets
// products s0 + u * 64 + l; a position past the row reads the row's last
// product instead (its lane is masked where the value is used).
__device__ __forceinline__ void rank_wave_load(const int* __restrict__ icol,
                                               const float* __restrict__ ival, long long s0,
                                               long long s1, int* c, float* x) {
    const int lane = lane_id();
    const long long last = s1 - 1;
#pragma unroll
    for (int u = 0; u < kRankWavePer; ++u) {
        long long q = s0 + u * 64 + lane;
        q = islot(q < s1 ? q : last);
        c[u] = icol[q];
        x[u] = ival[q];
    }
}

// Rank of window column cc: set columns before it (quad prefix + the
// words of its quad before it + the bits of its word below it).
__device__ __forceinline__ int rank_of(const unsigned* bits, const unsigned short* pre, int cc) {
    const int w = cc >> 5, k = w & 3;
    const uint4 q = *(const uint4*)(bits + (w & ~3));
    const unsigned low = (1u << (cc & 31)) - 1u;
    const unsigned a = k == 0 ? q.x & low : q.x;
    const unsigned b = k == 1 ? q.y & low : (k > 1 ? q.y : 0u);
    const unsigned c = k == 2 ? q.z & low : (k > 2 ? q.z : 0u);
    const unsigned d = k == 3 ? q.w & low : 0u;
    return (int)pre[w >> 2] + __popc(a) + __popc(b) + __popc(c) + __popc(d);
}

template <int NW, int CAP, bool STAMP>
__global__ __launch_bounds__(64) void k_merge_rank_wave(
    const int* __restrict__ vlist, long long nlist, int G, unsigned long long* __restrict__ ctr,
    const long long* __restrict__ vbase,
    const int* __restrict__ vlo, const int* __restrict__ vwid, const int* __restrict__ icol,
    const float* __restrict__ ival, const long long* __restrict__ vout, int* __restrict__ ccol,
    float* __restrict__ cval, long long* __restrict__ cnt_out, const Done dn) {
    constexpr int PER = kRankWavePer;
    constexpr int NQ = NW / 256;  // 16-byte word quads per lane, at most
    static_assert(NW % 256 == 0, "words are handled in quads, the same number per lane");
    __shared__ __align__(16) unsigned bits[NW];
    __shared__ unsigned short pre[NW / 4];  // rank before each word quad
    __shared__ double acc[CAP];
    __shared__ int colr[CAP];
    const int lane = lane_id();
    for (int w = lane; w < NW; w += 64) bits[w] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const long long ngrp = (nlist + G - 1) / G;
    // groups of G rows: the first by the wave's index, the others from the
    // counter *ctr (zero at launch)
    auto ticket = [&]() -> long long {
        unsigned t = 0;
        if (lane == 0) t = (unsigned)atomicAdd(ctr, 1ull);
        return (long long)gridDim.x + (long long)__builtin_amdgcn_readfirstlane(t);
    };
    for (long long g = blockIdx.x; g < ngrp; g = ticket()) {
        // descriptors of the group's rows, one per lane
        const long long jl = g * G + lane;
        int mv = 0, mcnt = 0, mlo = 0, mnw = 0;
        long long ms0 = 0, ms1 = 0, mo = 0;
        if (lane < G && jl < nlist) {
            mv = vlist[jl];
            ms0 = vbase[mv];
            ms1 = vbase[mv + 1];
            mo = vout[mv];
            mcnt = (int)min((long long)CAP, vout[mv + 1] - mo);
            mlo = vlo[mv];
            mnw = (vwid[mv] + 31) >> 5;  // <= NW (class condition)
        }
        // the descriptors are complete before the row loop (else every row
        // would wait there for the previous row's stores as well)
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        const int nr = (int)min((long long)G, nlist - g * G);
        int c[PER];
        float x[PER];
        rank_wave_load(icol, ival, readlane64(ms0, 0), readlane64(ms1, 0), c, x);
        for (int r = 0; r < nr; ++r) {
            const int v = __builtin_amdgcn_readlane(mv, r);
            const long long s0 = readlane64(ms0, r), s1 = readlane64(ms1, r);
            const long long o0 = readlane64(mo, r);
            const int cnt = __builtin_amdgcn_readlane(mcnt, r);
            const int lo = __builtin_amdgcn_readlane(mlo, r);
            const int nw = __builtin_amdgcn_readlane(mnw, r);
            const long long base = s0 + lane;
            // the next row's products (the last row loads its own again)
            int cn[PER];
            float xn[PER];
            const int rn = r + 1 < nr ? r + 1 : r;
            rank_wave_load(icol, ival, readlane64(ms0, rn), readlane64(ms1, rn), cn, xn);
            // pass 1: the row's column bitmap
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                if (s0 + u * 64 < s1) {  // uniform
                    const bool ok = base + u * 64 < s1;
                    const int cc = c[u] - lo;
                    atomicOr(&bits[ok ? cc >> 5 : 0], ok ? 1u << (cc & 31) : 0u);
                }
            }
            for (long long q = s0 + PER * 64 + lane; q < s1; q += 64) {
                const int cc = icol[islot(q)] - lo;
                atomicOr(&bits[cc >> 5], 1u << (cc & 31));
            }
            for (int q = lane; q < cnt; q += 64) acc[q] = 0.0;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // prefix popcounts over word quads, a multiple of four words per
            // lane (words past nw are zero; 64 * wpt <= NW), held in
            // registers from the count to the prefix
            const int wpt = (((nw + 63) >> 6) + 3) & ~3;
            const int w0 = lane * wpt, nq = max(0, min(wpt, nw - w0) + 3) >> 2;
            uint4 qw[NQ];
            int mine = 0;
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                qw[i] = make_uint4(0u, 0u, 0u, 0u);
                if (i < nq) qw[i] = *(const uint4*)(bits + w0 + 4 * i);
                mine += __popc(qw[i].x) + __popc(qw[i].y) + __popc(qw[i].z) + __popc(qw[i].w);
            }
            int inc = mine;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(inc, o, 64);
                if (lane >= o) inc += y;
            }
            const int total = __shfl(inc, 63, 64);
            if (cnt_out && lane == 0) cnt_out[v] = total;
            int run = inc - mine;
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                if (i < nq) pre[(w0 >> 2) + i] = (unsigned short)run;
                run += __popc(qw[i].x) + __popc(qw[i].y) + __popc(qw[i].z) + __popc(qw[i].w);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // pass 2: ranks first (all reads in flight together), then the
            // FP64 accumulation and the column of every rank
            int rk[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                rk[u] = CAP;
                if (s0 + u * 64 < s1) {
                    const bool ok = base + u * 64 < s1;
                    const int r0 = rank_of(bits, pre, ok ? c[u] - lo : 0);
                    rk[u] = ok ? r0 : CAP;
                }
            }
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                if (s0 + u * 64 < s1 && rk[u] < cnt) {
                    unsafeAtomicAdd(&acc[rk[u]], (double)x[u]);
                    colr[rk[u]] = c[u];  // every product of a column stores the same value
                }
            }
            for (long long q = s0 + PER * 64 + lane; q < s1; q += 64) {
                const int cc = icol[islot(q)] - lo;
                const int r0 = rank_of(bits, pre, cc);
                if (r0 < cnt) {
                    unsafeAtomicAdd(&acc[r0], (double)ival[islot(q)]);
                    colr[r0] = cc + lo;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // the next row's products (this waits for their loads, not for
            // the stores below)
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                c[u] = cn[u];
                x[u] = xn[u];
            }
            const int ne = min(cnt, total);
            for (int q = lane; q < ne; q += 64) {
                ccol[o0 + q] = colr[q];
                cval[o0 + q] = (float)acc[q];
            }
            if (STAMP && lane == 0) mark_done(dn, v);
#pragma unroll
            for (int i = 0; i < NQ; ++i)  // a lane clears the words it counted
                if (i < nq) *(uint4*)(bits + w0 + 4 * i) = make_uint4(0u, 0u, 0u, 0u);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Where a merge writes: virtual row v's columns go to ocol/oval from
// obase[v] on, at most obase[v+1] - obase[v] of them; cnt (may be null)
// receives each row's distinct-column count.
struct MergeOut {
    const long long* obase;
    int* ocol;
    float* oval;
    long long* cnt;
    Done dn;  // completion times of batched products (null: none)
};

// The context's completion-time recording (SPG_BATCH_LATENCY), or none.
static Done done_of(const spg_context* c) {
    Done d{nullptr, nullptr};
    if (c->lat_on) {
        d.t = (unsigned long long*)c->lat_t.p;
        d.vprod = (const int*)c->vprod.p;
    }
    return d;
}

template <int LOG2H, bool STAMP>
static void launch_merge_hash(spg_context* c, const int* list, long long cnt, const MergeOut& o) {
    hipLaunchKernelGGL((k_merge_hash<LOG2H, STAMP>), dim3((unsigned)cnt), dim3(kBlock), 0, c->stream,
                       list, (const long long*)c->vbase.p, (const int*)inter_cols(c),
                       (const float*)inter_vals(c), o.obase, o.ocol, o.oval, o.cnt, o.dn);
}

// Every counter of the context zeroed (one fill kernel, also inside a graph
// capture) at the start of a merge phase, whose kernels then take pool slots.
static int reset_counters(spg_context* c) {
    c->ctr_pool = 0;
    return fill<unsigned long long>(c, (unsigned long long*)c->counters.p, 0ull, kCtrWords);
}

// A work counter of its own for one merge kernel: a pool slot, zero since
// the phase started.
static int zeroed_counter(spg_context* c, unsigned long long** ctr) {
    if (c->ctr_pool >= kCtrPoolSlots) return fail(c, SPG_EINVAL, "merge counters exhausted");
    *ctr = (unsigned long long*)c->counters.p + kCtrPoolBase + kCtrPoolStride * c->ctr_pool++;
    return SPG_OK;
}

// One merge kernel over the virtual rows list[0..cnt) of class b.  build:
// tile rows build their column bitmap themselves (no symbolic pass ran).
template <bool STAMP>
static int launch_merge_class(spg_context* c, int b, const int* list, long long cnt,
                              const MergeOut& o, bool build) {
    const long long* vbase = (const long long*)c->vbase.p;
    const int* vlo = (const int*)c->vlo.p;
    const int* icol = inter_cols(c);
    const uint16_t* icol16 = inter_cols16(c);
    const float* ival = inter_vals(c);
    const unsigned* tb = (const unsigned*)c->tilebits.p;
    if (b == kClassTileS) {
        // three waves per workgroup: 12 KB of LDS, so one fits beside the
        // two dense-tile workgroups of a CU (133 KB) and the medium class's
        // one-wave workgroup, and runs while the dense class does
        constexpr int WPB = kTileSWPB;
        const dim3 g(persistent_grid(c, (cnt + WPB - 1) / WPB, 8));
        unsigned long long* ctr = nullptr;
        SPG_TRY(zeroed_counter(c, &ctr));
        if (build)
            hipLaunchKernelGGL((k_merge_tile_wave<256, WPB, true, STAMP>), g, dim3(64 * WPB), 0, c->stream,
                               list, cnt, ctr, vbase, vlo, tb, icol16, ival, o.obase, o.ocol, o.oval,
                               o.cnt, o.dn);
        else
            hipLaunchKernelGGL((k_merge_tile_wave<256, WPB, false, STAMP>), g, dim3(64 * WPB), 0, c->stream,
                               list, cnt, ctr, vbase, vlo, tb, icol16, ival, o.obase, o.ocol, o.oval,
                               o.cnt, o.dn);
    } else if (b == kClassTileM) {
        // one wave per workgroup: 11.8 KB of LDS each (16-bit columns
        // beside the FP64 accumulators), so a CU holds 13 of them
        const dim3 g(persistent_grid(c, cnt, 13));
        unsigned long long* ctr = nullptr;
        SPG_TRY(zeroed_counter(c, &ctr));
        if (build)
            hipLaunchKernelGGL((k_merge_tile_wave<1024, 1, true, STAMP>), g, dim3(64), 0, c->stream, list,
                               cnt, ctr, vbase, vlo, tb, icol16, ival, o.obase, o.ocol, o.oval, o.cnt,
                               o.dn);
        else
            hipLaunchKernelGGL((k_merge_tile_wave<1024, 1, false, STAMP>), g, dim3(64), 0, c->stream,
                               list, cnt, ctr, vbase, vlo, tb, icol16, ival, o.obase, o.ocol, o.oval,
                               o.cnt, o.dn);
    } else if (b == kClassTile) {
        const dim3 g(persistent_grid(c, cnt, 2));
        if (build)
            hipLaunchKernelGGL((k_merge_tile<true, STAMP>), g, dim3(kPT), 0, c->stream, list, cnt, vbase,
                               vlo, (const int*)c->vwid.p, tb, icol16, ival, o.obase, o.ocol,
                               o.oval, o.cnt, o.dn);
        else
            hipLaunchKernelGGL((k_merge_tile<false, STAMP>), g, dim3(kPT), 0, c->stream, list, cnt, vbase,
                               vlo, (const int*)c->vwid.p, tb, icol16, ival, o.obase, o.ocol,
                               o.oval, o.cnt, o.dn);
    } else if (b == kClassRankW) {
        // groups of up to 64 rows, about a quarter of a wave's share of the
        // list (7 waves per CU), taken from a counter so that they balance
        const long long waves = (long long)c->num_cu * 7;
        const int G = (int)std::min<long long>(64, std::max<long long>(1, cnt / waves / 4));
        unsigned long long* ctr = nullptr;
        SPG_TRY(zeroed_counter(c, &ctr));
        hipLaunchKernelGGL((k_merge_rank_wave<kRankWaveWords, kRankWaveCap, STAMP>),
                           dim3(persistent_grid(c, (cnt + G - 1) / G, 7)), dim3(64), 0,
                           c->stream, list, cnt, G, ctr, vbase, vlo, (const int*)c->vwid.p, icol,
                           ival, o.obase, o.ocol, o.oval, o.cnt, o.dn);
    } else if (b == kClassRank) {
        hipLaunchKernelGGL(k_merge_rank<STAMP>,
                           dim3(persistent_grid(c, (cnt + kRankChunk - 1) / kRankChunk, 2)),
                           dim3(kPT), 0, c->stream, list, cnt,
                           (unsigned long long*)c->counters.p + 2, vbase, vlo, icol, ival,
                           o.obase, o.ocol, o.oval, o.cnt, o.dn);
    } else {
        const dim3 gw(persistent_grid(c, (cnt + kHashWaveWPB - 1) / kHashWaveWPB, 4));
        const dim3 bw(64 * kHashWaveWPB);
        switch (b + kHashMinLog2) {
            case 6:
                hipLaunchKernelGGL((k_merge_hash_wave<6, kHashWaveWPB, STAMP>), gw, bw, 0, c->stream,
                                   list, cnt, vbase, icol, ival, o.obase, o.ocol, o.oval, o.cnt, o.dn);
                break;
            case 7:
                hipLaunchKernelGGL((k_merge_hash_wave<7, kHashWaveWPB, STAMP>), gw, bw, 0, c->stream,
                                   list, cnt, vbase, icol, ival, o.obase, o.ocol, o.oval, o.cnt, o.dn);
                break;
            case 8:
                hipLaunchKernelGGL((k_merge_hash_wave<8, kHashWaveWPB, STAMP>), gw, bw, 0, c->stream,
                                   list, cnt, vbase, icol, ival, o.obase, o.ocol, o.oval, o.cnt, o.dn);
                break;
            case 9:
                hipLaunchKernelGGL((k_merge_hash_wave<9, 4, STAMP>),
                                   dim3(persistent_grid(c, (cnt + 3) / 4, 4)), dim3(256), 0,
                                   c->stream, list, cnt, vbase, icol, ival, o.obase, o.ocol,
                                   o.oval, o.cnt, o.dn);
                break;
            case 10: launch_merge_hash<10, STAMP>(c, list, cnt, o); break;
            case 11: launch_merge_hash<11, STAMP>(c, list, cnt, o); break;
            case 12: launch_merge_hash<12, STAMP>(c, list, cnt, o); break;
            case 13: launch_merge_hash<13, STAMP>(c, list, cnt, o); break;
            default: return fail(c, SPG_EINVAL, "bad merge class");
        }
    }
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

// The kernels that stamp completion times are separate instances: a stamp
// that is compiled in but never taken still slowed the wave hash merge by
// 70 % (1.0 -> 1.75 ms on the 64^3 stencil).
static int merge_class(spg_context* c, int b, const int* list, long long cnt, const MergeOut& o,
                       bool build) {
    return o.dn.t ? launch_merge_class<true>(c, b, list, cnt, o, build)
                  : launch_merge_class<false>(c, b, list, cnt, o, build);
}

// The merge classes are independent (disjoint rows, disjoint output
// ranges): the dense-tile class runs on the context's stream and the others
// beside it on two more, so they fill the issue slots and the memory
// pipeline its LDS-bound workgroups leave idle.  Forked from the context's
// stream by an event and joined back before the phase ends (also inside a
// graph capture).  false: everything stays on the context's stream.
static bool aux_streams(spg_context* c) {
    if (c->noalloc) return false;
    if (!c->aux_tried) {
        c->aux_tried = true;
        bool ok = true;
        for (int i = 0; i < 2 && ok; ++i)
            ok = hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 2 && ok; ++i)
            ok = hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming) == hipSuccess;
        if (!ok) (void)hipGetLastError();
        c->aux_ok = ok;
    }
    return c->aux_ok;
}

static int fork_aux(spg_context* c) {
    SPG_HIP(hipEventRecord(c->fork_ev, c->stream));
    for (int i = 0; i < 2; ++i) SPG_HIP(hipStreamWaitEvent(c->aux[i], c->fork_ev, 0));
    return SPG_OK;
}

static int join_aux(spg_context* c) {
    for (int i = 0; i < 2; ++i) {
        SPG_HIP(hipEventRecord(c->join_ev[i], c->aux[i]));
        SPG_HIP(hipStreamWaitEvent(c->stream, c->join_ev[i], 0));
    }
    return SPG_OK;
}

// Launches the merge kernels of every class with rows, the dense-tile class
// first (its workgroups take their LDS before the others fill the CUs).
// Only the two small wave-level tile classes run beside it (their
// workgroups fit in the LDS two dense-tile workgroups leave on a CU); the
// rank and hash classes follow on the context's stream once all three are
// done: run beside them they competed for the same CUs and took longer
// than in sequence (256 small products: merge 18.5 ms overlapped, 16 ms
// one after the other).
// Products below this many intermediate products merge on one stream: the
// fork and join cost tens of microseconds, more than the overlap saves.
constexpr long long kOverlapMaxRows = 1LL << 15;  // rank/hash classes run beside the dense one

static int launch_merge_classes(spg_context* c, const int* vl, const long long* off,
                                const MergeOut& o, bool build_all) {
    const bool par = c->nprod >= kForkMinProducts && aux_streams(c);
    if (par) SPG_TRY(fork_aux(c));
    const hipStream_t main = c->stream;
    int rc = SPG_OK;
    const long long nt = off[kClassTile + 1] - off[kClassTile];
    if (nt > 0) rc = merge_class(c, kClassTile, vl + off[kClassTile], nt, o, build_all);
    const int order[] = {kClassTileM, kClassTileS, kClassRankW, kClassRank, 7, 6, 5, 4, 3, 2, 1, 0};
    static_assert(sizeof(order) / sizeof(order[0]) == kNumVClasses - 1, "every class but the dense one");
    // classes of few rows also run beside the wave tile classes (after the
    // dense class, on its stream), the larger ones after the join
    bool joined = !par;
    for (int pass = 0; pass < 2 && rc == SPG_OK; ++pass) {
        for (const int b : order) {
            const long long cnt = off[b + 1] - off[b];
            if (rc != SPG_OK || cnt <= 0) continue;
            const bool beside = b == kClassTileM || b == kClassTileS || cnt <= kOverlapMaxRows;
            if (par && beside != (pass == 0)) continue;
            if (!par && pass == 1) continue;
            if (!joined && !beside) {
                // the others start once the tile classes are done, with the
                // GPU to themselves (persistent grids dispatched while the
                // CUs are held ran their fixed shares late)
                SPG_TRY(join_aux(c));
                joined = true;
            }
            // the two wave tile classes on the merge streams; the small rank
            // and hash classes follow the dense class on the context's stream
            if (!joined && (b == kClassTileM || b == kClassTileS))
                c->stream = b == kClassTileM ? c->aux[0] : c->aux[1];
            // without a symbolic pass every class builds its bitmaps; after
            // one, the small tile rows still do (it kept none for them)
            rc = merge_class(c, b, vl + off[b], cnt, o, build_all || b == kClassTileS);
            c->stream = main;
        }
    }
    if (!joined) SPG_TRY(join_aux(c));
    return rc;
}

int launch_merge(spg_context* c) {
    const long long V = c->V;
    SPG_TRY(reset_counters(c));
    if (V > 0) {
        hipLaunchKernelGGL(k_vclass_merge, dim3((unsigned)((V + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, c->stream, V, (const long long*)c->vcnt.p,
                           (const int*)c->vwid.p, c->W, c->cfg.rank_min_count,
                           (int*)c->vclass.p);
        SPG_HIP(hipGetLastError());
    }
    SPG_TRY(stable_partition(c, (const int32_t*)c->vclass.p, V, kNumVClasses,
                             (int32_t*)c->vlist.p, (int64_t*)c->vclass_off.p));
    long long off[kNumVClasses + 1];
    SPG_TRY(class_offsets(c, off));
    const int* vl = (const int*)c->vlist.p;
    const MergeOut o{(const long long*)c->vout.p, (int*)c->ccol.p, (float*)c->cval.p, nullptr,
                     done_of(c)};
    tmark(c, SPG_PHASE_MERGE, 0);
    SPG_TRY(launch_merge_classes(c, vl, off, o, false));
    tmark(c, SPG_PHASE_MERGE, 1);
    return SPG_OK;
}

// ---------------------------------------------------------- learned merge
// The symbolic pass is replaced by a predicted count per C row: every
// virtual row gets an output capacity from the prediction, is merged into a
// padded buffer by the kernel of the class that capacity selects (tile rows
// build their column bitmap themselves), and reports its true distinct
// count.  Rows whose count exceeds the capacity (or whose predicted-size
// hash table filled up) are merged again with their upper bound
// min(products, width) as capacity, which always fits; the exact counts
// then give C's row offsets and the rows are compacted into CSR.

// Adds the number of lanes with flag set to *ctr (one atomic per wave).
// Adds every thread's count to *ctr with one atomic per workgroup (kBlock
// threads, all of which call it): the atomics of a whole grid on one
// address serialize (one per wave cost 0.9 ms for 60K waves).
__device__ __forceinline__ void block_count_add(unsigned long long* ctr, unsigned n) {
    __shared__ unsigned long long part[kBlock / 64];
    unsigned long long x = n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane_id() == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += part[w];
        if (t) atomicAdd(ctr, t);
    }
}

__device__ __forceinline__ int hash_class_for(long long cap) {
    int lg = kHashMinLog2;
    while ((1LL << lg) < 2 * cap) ++lg;
    return lg - kHashMinLog2;
}

// Distinct columns expected when u products fall uniformly on w columns.
__device__ __forceinline__ double saturated(long long u, int w) {
    return (double)w * -expm1(-(double)u / (double)w);
}

// Per row split into column tiles: the sum over its tiles of saturated(),
// by which a tile's share of the row's prediction is weighted.
__global__ void k_learned_rowest(int m, const int* __restrict__ vstart,
                                 const long long* __restrict__ vub, const int* __restrict__ vwid,
                                 double* __restrict__ rowest) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int s = vstart[i], e = vstart[i + 1];
    double t = 0.0;
    if (e - s > 1)
        for (int v = s; v < e; ++v) t += saturated(vub[v], vwid[v]);
    rowest[i] = t;
}

// Class and output capacity of every virtual row.  The class is chosen from
// the row's predicted count (inflated by margin and slack for the sizes of
// the tile and hash classes; the raw prediction decides rank vs hash, as the
// exact count does in the plain merge); the capacity is
// the most that class's kernel can hold (256 or 1024 columns for the wave
// tile kernels, 1024 for the wave rank kernel, the hash table's half for
// hash rows) or the upper bound
// min(products, width) for the dense tile and rank kernels, so a row
// overflows only when its count exceeds its class (with tight, the inflated
// prediction itself, which saves memory).  Capacity costs memory,
// not bandwidth: the merge writes and the compaction reads only the row's
// count.  A tile of a long row gets the share of its row's prediction
// weighted by saturated(products, width).
__global__ void k_learned_class(long long V, const int* __restrict__ vrow_row,
                                const int* __restrict__ vstart, const long long* __restrict__ vub,
                                const int* __restrict__ vwid, const double* __restrict__ rowest,
                                const float* __restrict__ pred, int W, int rank_min, float margin,
                                int slack, int tight, long long* __restrict__ vcap,
                                int* __restrict__ cls, unsigned long long* __restrict__ nlive) {
    unsigned nl = 0;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (long long)gridDim.x * blockDim.x) {
        const long long u = vub[v];
        long long cap = 0;
        int k = -1;
        if (u > 0) {
            ++nl;
            const int width = vwid[v];
            const long long bound = min(u, (long long)width);
            const int i = vrow_row[v];
            double share = (double)pred[i];
            if (vstart[i + 1] - vstart[i] > 1) {
                const double tot = rowest[i];
                share = tot > 0.0 ? share * saturated(u, width) / tot : 0.0;
            }
            if (!(share > 0.0)) share = 0.0;  // negative or NaN
            double want = ceil(share * (1.0 + (double)margin)) + (double)slack;
            want = want < 1.0 ? 1.0 : (want > (double)bound ? (double)bound : want);
            const long long wl = (long long)want;
            const bool rank_ok = share > (double)rank_min;  // the raw prediction
            if (width <= W) {
                if (wl <= 256) {
                    k = kClassTileS;
                    cap = min(bound, 256LL);
                } else if (wl <= 1024) {
                    k = kClassTileM;
                    cap = min(bound, 1024LL);
                } else {
                    k = kClassTile;
                    cap = bound;
                }
            } else if (rank_ok && width <= kRankWaveWords * 32 && wl <= kRankWaveCap) {
                k = kClassRankW;
                cap = min(bound, (long long)kRankWaveCap);
            } else if (rank_ok && width <= kRankMaxWords * 32) {
                k = kClassRank;
                cap = bound;  // <= 4096 = the rank kernel's accumulator
            } else {
                k = hash_class_for(wl);  // wl <= u <= 4096: at most 8192 slots
                cap = min(bound, (1LL << (k + kHashMinLog2)) / 2);
            }
            if (tight) cap = min(cap, wl);  // least memory, more overflows
        }
        vcap[v] = cap;
        cls[v] = k;
    }
    block_count_add(nlive, nl);
}

// Rows that did not fit their capacity (count > capacity, or -1 from a full
// hash table) get the upper bound as fallback capacity (fcap) and a class
// that merges any count up to it; all other rows get 0 / -1.
__global__ void k_learned_check(long long V, const long long* __restrict__ vcnt,
                                const long long* __restrict__ pbase,
                                const long long* __restrict__ vub, const int* __restrict__ vwid,
                                int W, long long* __restrict__ fcap, int* __restrict__ cls,
                                unsigned long long* __restrict__ nover) {
    unsigned no = 0;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (long long)gridDim.x * blockDim.x) {
        const long long cap = pbase[v + 1] - pbase[v], cnt = vcnt[v];
        const bool over = cap > 0 && (cnt < 0 || cnt > cap);
        no += over ? 1u : 0u;
        long long f = 0;
        int k = -1;
        if (over) {
            const int width = vwid[v];
            f = min(vub[v], (long long)width);
            if (width <= W)
                k = kClassTile;
            else if (width <= kRankMaxWords * 32)
                k = kClassRank;
            else
                k = hash_class_for(f);
        }
        fcap[v] = f;
        cls[v] = k;
    }
    block_count_add(nover, no);
}

// Moves every virtual row's columns and values from the padded (or the
// fallback) buffer to its CSR position.  A wave takes 64 consecutive
// virtual rows, whose outputs are one contiguous range of C: each lane loads
// one row's descriptor, and every output position of the range finds its row
// by a binary search over the 64 row offsets (shuffles), so stores are
// coalesced and no per-row latency chain is serialized.
template <bool STAMP>
__global__ __launch_bounds__(kPT) void k_learned_compact(
    long long V, const long long* __restrict__ vcnt, const long long* __restrict__ vout,
    const long long* __restrict__ pbase, const long long* __restrict__ fbase,
    const int* __restrict__ pcol, const float* __restrict__ pval, const int* __restrict__ fcol,
    const float* __restrict__ fval, int* __restrict__ ccol, float* __restrict__ cval,
    unsigned long long* __restrict__ unrecovered, const Done dn) {
    const int wv = threadIdx.x >> 6, lane = lane_id();
    constexpr int NWV = kPT / 64;
    const long long ngroups = (V + 63) / 64;
    for (long long grp = (long long)blockIdx.x * NWV + wv; grp < ngroups;
         grp += (long long)gridDim.x * NWV) {
        const long long v = grp * 64 + lane;
        long long cnt = 0, src = 0, o = 0;
        int from_fb = 0;
        bool clamped = false;
        if (v < V) {
            cnt = vcnt[v];
            o = vout[v];
            const long long fb = fbase[v], fcap = fbase[v + 1] - fb;
            long long lim;
            if (fcap > 0) {
                from_fb = 1;
                src = fb;
                lim = fcap;
            } else {
                src = pbase[v];
                lim = pbase[v + 1] - src;
            }
            if (cnt > lim) {  // cannot happen: the fallback capacity is an upper bound
                atomicAdd(unrecovered, 1ull);
                cnt = lim;
                clamped = true;
            }
            if (cnt < 0) cnt = 0;
        } else if (V > 0) {
            o = vout[V];  // past the end: an empty row at the range's end
        }
        // a virtual row holds at most 8192 columns (a tile, or a row with at
        // most 4096 products), so 64 rows' offsets fit in int
        const long long o0 = __shfl(o, 0, 64);
        const int rel = (int)(o - o0);
        const int rend = rel + (int)cnt;  // end of the row's copied range
        const bool clean = __ballot(clamped) == 0;  // no row of the group was cut short
        const int total = __shfl((int)(o - o0 + cnt), 63, 64);
        if (total >= 64 * 256) {
            // long rows: one row after the other, the whole wave per row
            for (int qq = 0; qq < 64; ++qq) {
                const int cq = __shfl((int)cnt, qq, 64);
                const long long sq = __shfl(src, qq, 64);
                const long long oq = o0 + __shfl(rel, qq, 64);
                const bool fq = __shfl(from_fb, qq, 64) != 0;
                const int* __restrict__ sc = fq ? fcol + sq : pcol + sq;
                const float* __restrict__ sv = fq ? fval + sq : pval + sq;
                // whole passes without lane conditions (a store under one would
                // make every later store wait for the earlier ones), then the tail
                const int full = cq & ~(4 * 64 - 1);
                for (int e0 = 0; e0 < full; e0 += 4 * 64) {
                    int c4[4];
                    float x4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        c4[u] = sc[e0 + u * 64 + lane];
                        x4[u] = sv[e0 + u * 64 + lane];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        ccol[oq + e0 + u * 64 + lane] = c4[u];
                        cval[oq + e0 + u * 64 + lane] = x4[u];
                    }
                }
                if (full < cq) {
                    int c4[4];
                    float x4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int e = full + u * 64 + lane;
                        if (e < cq) {
                            c4[u] = sc[e];
                            x4[u] = sv[e];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int e = full + u * 64 + lane;
                        if (e < cq) {
                            ccol[oq + e] = c4[u];
                            cval[oq + e] = x4[u];
                        }
                    }
                }
            }
        if (STAMP && dn.t) {  // uniform: this group's rows are written; one atomic per product
            const int pv = v < V && cnt > 0 ? dn.vprod[v] : -1;
            const int pp = __shfl_up(pv, 1, 64);
            if (pv >= 0 && (lane == 0 || pp != pv))
                atomicMax(&dn.t[pv], (unsigned long long)wall_clock64());
        }
            continue;
        }
        for (int pb = 0; pb < total; pb += 4 * 64) {
            int p[4], q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                p[u] = pb + u * 64 + lane;
                q[u] = 0;  // largest lane whose range starts at or before p
            }
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int cand = q[u] + step;
                    if (__shfl(rel, cand, 64) <= p[u]) q[u] = cand;
                }
            }
            int cc[4];
            float xx[4];
            bool ok[4];
            long long sp[4];
            bool fp[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                // every shuffle runs with all lanes active: a lane that is
                // inactive in a bpermute does not supply its value
                sp[u] = __shfl(src, q[u], 64) + (p[u] - __shfl(rel, q[u], 64));
                fp[u] = __shfl(from_fb, q[u], 64) != 0;
                const int end = __shfl(rend, q[u], 64);
                ok[u] = p[u] < total && p[u] < end;
            }
            if (clean && pb + 4 * 64 <= total) {
                // a whole pass of a group without gaps: every position is
                // copied, with no lane condition (a store under one would make
                // every later store wait for the earlier ones)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cc[u] = (fp[u] ? fcol : pcol)[sp[u]];
                    xx[u] = (fp[u] ? fval : pval)[sp[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ccol[o0 + p[u]] = cc[u];
                    cval[o0 + p[u]] = xx[u];
                }
                continue;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cc[u] = 0;
                xx[u] = 0.f;
                if (ok[u]) {
                    cc[u] = fp[u] ? fcol[sp[u]] : pcol[sp[u]];
                    xx[u] = fp[u] ? fval[sp[u]] : pval[sp[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (ok[u]) {
                    ccol[o0 + p[u]] = cc[u];
                    cval[o0 + p[u]] = xx[u];
                }
            }
        }
        if (STAMP && dn.t) {  // uniform: this group's rows are written; one atomic per product
            const int pv = v < V && cnt > 0 ? dn.vprod[v] : -1;
            const int pp = __shfl_up(pv, 1, 64);
            if (pv >= 0 && (lane == 0 || pp != pv))
                atomicMax(&dn.t[pv], (unsigned long long)wall_clock64());
        }
    }
}

// Counter slots of the learned merge in c->counters.
constexpr int kCtrRank = 2, kCtrLive = 8, kCtrOver = 9, kCtrUnrec = 10;

static int readback2(spg_context* c, const void* a, const void* b, long long* out) {
    SPG_TRY(readback(c, a, sizeof(long long), out));
    out[1] = 0;
    if (b) SPG_TRY(readback(c, b, sizeof(long long), out + 1));
    return SPG_OK;
}

int launch_learned_merge(spg_context* c, spg_learned_stats* st) {
    const long long V = c->V;
    const size_t V1 = (size_t)V + 1;
    unsigned long long* ctr = (unsigned long long*)c->counters.p;
    SPG_TRY(ensure(c, c->vcap, sizeof(long long) * V1));
    SPG_TRY(ensure(c, c->pbase, sizeof(long long) * (V1 + 1)));
    SPG_TRY(ensure(c, c->fbase, sizeof(long long) * (V1 + 1)));
    SPG_TRY(ensure(c, c->vcnt, sizeof(long long) * V1));
    SPG_TRY(ensure(c, c->vout, sizeof(long long) * (V1 + 1)));
    SPG_TRY(ensure(c, c->crow, sizeof(long long) * ((size_t)c->m + 1)));
    long long* vcap = (long long*)c->vcap.p;
    long long* pbase = (long long*)c->pbase.p;
    long long* fbase = (long long*)c->fbase.p;
    long long* vcnt = (long long*)c->vcnt.p;
    // grid-stride: few workgroups, so few atomics on the row counters
    const unsigned gV = persistent_grid(c, (V + kBlock - 1) / kBlock, 4);

    tmark(c, SPG_PHASE_MERGE, 0);
    SPG_TRY(reset_counters(c));
    SPG_TRY(fill<long long>(c, vcnt, 0, V));
    if (V > 0) {
        // fbase doubles as the per-row tile weight sums until the fallback
        double* rowest = (double*)fbase;
        hipLaunchKernelGGL(k_learned_rowest, dim3((c->m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, c->m, (const int*)c->vstart.p, (const long long*)c->vub.p,
                           (const int*)c->vwid.p, rowest);
        SPG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_learned_class, dim3(gV), dim3(kBlock), 0, c->stream, V,
                           (const int*)c->vrow_row.p, (const int*)c->vstart.p,
                           (const long long*)c->vub.p, (const int*)c->vwid.p,
                           (const double*)rowest, (const float*)c->pred.p, c->W,
                           c->cfg.rank_min_count,
                           c->cfg.pred_margin, c->cfg.pred_slack, c->cfg.pred_tight, vcap,
                           (int*)c->vclass.p, ctr + kCtrLive);
        SPG_HIP(hipGetLastError());
    }
    scratch_reset(c);
    SPG_TRY(scan_i64(c, (const int64_t*)vcap, (int64_t*)pbase, V));
    long long rb[2];
    SPG_TRY(readback2(c, pbase + V, nullptr, rb));
    const long long padded = rb[0];
    SPG_TRY(ensure(c, c->pcol, sizeof(int) * (size_t)(padded + 16)));
    SPG_TRY(ensure(c, c->pval, sizeof(float) * (size_t)(padded + 16)));
    scratch_reset(c);
    SPG_TRY(stable_partition(c, (const int32_t*)c->vclass.p, V, kNumVClasses,
                             (int32_t*)c->vlist.p, (int64_t*)c->vclass_off.p));
    long long off[kNumVClasses + 1];
    SPG_TRY(class_offsets(c, off));
    const int* vl = (const int*)c->vlist.p;
    {
        const MergeOut o{pbase, (int*)c->pcol.p, (float*)c->pval.p, vcnt, Done{nullptr, nullptr}};
        SPG_TRY(launch_merge_classes(c, vl, off, o, true));
    }
    // exact fallback for the rows that did not fit
    if (V > 0) {
        hipLaunchKernelGGL(k_learned_check, dim3(gV), dim3(kBlock), 0, c->stream, V,
                           (const long long*)vcnt, (const long long*)pbase,
                           (const long long*)c->vub.p, (const int*)c->vwid.p, c->W, vcap,
                           (int*)c->vclass.p, ctr + kCtrOver);
        SPG_HIP(hipGetLastError());
    }
    scratch_reset(c);
    SPG_TRY(scan_i64(c, (const int64_t*)vcap, (int64_t*)fbase, V));
    SPG_TRY(readback2(c, fbase + V, ctr + kCtrOver, rb));
    const long long fslots = rb[0], nover = rb[1];
    if (nover > 0) {
        SPG_TRY(ensure(c, c->fcol, sizeof(int) * (size_t)(fslots + 16)));
        SPG_TRY(ensure(c, c->fval, sizeof(float) * (size_t)(fslots + 16)));
        scratch_reset(c);
        SPG_TRY(stable_partition(c, (const int32_t*)c->vclass.p, V, kNumVClasses,
                                 (int32_t*)c->vlist.p, (int64_t*)c->vclass_off.p));
        SPG_TRY(class_offsets(c, off));
        SPG_TRY(fill<unsigned long long>(c, ctr + kCtrRank, 0ull, 1));
        const MergeOut o{fbase, (int*)c->fcol.p, (float*)c->fval.p, vcnt, Done{nullptr, nullptr}};
        SPG_TRY(launch_merge_classes(c, vl, off, o, true));
    }
    tmark(c, SPG_PHASE_MERGE, 1);

    tmark(c, SPG_PHASE_COMPACT, 0);
    scratch_reset(c);
    SPG_TRY(scan_i64(c, (const int64_t*)vcnt, (int64_t*)c->vout.p, V));
    SPG_TRY(readback2(c, (const long long*)c->vout.p + V, nullptr, rb));
    c->nnzC = rb[0];
    SPG_TRY(ensure(c, c->ccol, sizeof(int) * (size_t)(c->nnzC + 1)));
    SPG_TRY(ensure(c, c->cval, sizeof(float) * (size_t)(c->nnzC + 1)));
    SPG_TRY(launch_rowptr(c));
    if (V > 0) {
        const int* fcol = nover > 0 ? (const int*)c->fcol.p : nullptr;
        const float* fval = nover > 0 ? (const float*)c->fval.p : nullptr;
        const Done dn = done_of(c);
        auto compact = dn.t ? k_learned_compact<true> : k_learned_compact<false>;
        hipLaunchKernelGGL(compact,
                           dim3(persistent_grid(c, (V + 64 * 8 - 1) / (64 * 8), 4)), dim3(kPT),
                           0, c->stream, V, (const long long*)vcnt, (const long long*)c->vout.p,
                           (const long long*)pbase, (const long long*)fbase,
                           (const int*)c->pcol.p, (const float*)c->pval.p, fcol, fval,
                           (int*)c->ccol.p, (float*)c->cval.p, ctr + kCtrUnrec, dn);
        SPG_HIP(hipGetLastError());
    }
    tmark(c, SPG_PHASE_COMPACT, 1);
    if (st) {
        long long live_unrec[2];
        SPG_TRY(readback2(c, ctr + kCtrLive, ctr + kCtrUnrec, live_unrec));
        st->vrows = live_unrec[0];
        st->overflow_vrows = nover;
        st->unrecovered = live_unrec[1];
        st->padded_slots = padded;
        st->fallback_slots = fslots;
        st->nnz = c->nnzC;
    }
    return SPG_OK;
}

// ------------------------------------------------ batch completion times
__global__ void k_vprod(long long V, const int* __restrict__ vrow_row,
                        const int* __restrict__ rprod, int* __restrict__ vprod) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) vprod[v] = rprod[vrow_row[v]];
}

__global__ void k_stamp(unsigned long long* t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t = (unsigned long long)wall_clock64();
}

int launch_latency_start(spg_context* c) {
    unsigned long long* t = (unsigned long long*)c->lat_t.p;
    SPG_TRY(fill<unsigned long long>(c, t, 0ull, c->batch_count));
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, c->stream, t + c->batch_count);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

int launch_vprod(spg_context* c) {
    SPG_TRY(ensure(c, c->vprod, sizeof(int) * ((size_t)c->V + 1)));
    if (c->V > 0) {
        hipLaunchKernelGGL(k_vprod, dim3((unsigned)((c->V + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, (long long)c->V, (const int*)c->vrow_row.p,
                           (const int*)c->rprod.p, (int*)c->vprod.p);
        SPG_HIP(hipGetLastError());
    }
    return SPG_OK;
}

// ------------------------------------------------ intermediate inspection
__global__ void k_inter_rowptr(const int* __restrict__ vstart,
                               const long long* __restrict__ vbase, int m,
                               long long* __restrict__ rp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= m) rp[i] = vbase[vstart[i]];
}

int launch_regroup_intermediate(spg_context* c, int64_t* d_rowptr) {
    hipLaunchKernelGGL(k_inter_rowptr, dim3((c->m + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       c->stream, (const int*)c->vstart.p, (const long long*)c->vbase.p, c->m,
                       (long long*)d_rowptr);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}

template int fill<int>(spg_context*, int*, int, long long);
template int fill<long long>(spg_context*, long long*, long long, long long);
template int fill<unsigned long long>(spg_context*, unsigned long long*, unsigned long long, long long);

}  // namespace spg
