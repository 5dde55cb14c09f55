// spg_api.hip - context management and the extern "C" phase entry points of
// libspgemm_amd.so (declared in include/spgemm.h).
//
// The phase sequence and its meaning follow the reference driver
// src/main.cu:20-37: COO2CM(A, COL_MAJOR) -> COO2CM(B, ROW_MAJOR) ->
// InitGEMM -> CategorizeGEMM -> BinGEMM | SimpleGEMM -> MergeGEMM.  All
// device state lives in the context (the reference keeps it in file-scope
// globals, src/mm.cu:19-41,259-261) and every call returns a status.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>

#include "spg_internal.h"

namespace spg {

int launch_init_vrows(spg_context* c);
int launch_rowptr(spg_context* c);

int fail(spg_context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int ensure(spg_context* c, DevBuf& b, size_t bytes) {
    if (b.fixed) {
        if (b.cap >= bytes) return SPG_OK;
        return fail(c, SPG_EOVERFLOW, "C needs " + std::to_string(bytes) +
                                          " bytes in an array the caller sized " + std::to_string(b.cap));
    }
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return SPG_OK;
    if (c->noalloc) {  // a new range of the arena (the old one is not reused)
        const size_t off = (c->arena_top + 255) & ~(size_t)255;
        if (off > c->arena_cap || bytes > c->arena_cap - off)
            return fail(c, SPG_ENOMEM, "the product needs more than the " +
                                           std::to_string(c->arena_cap >> 20) +
                                           " MiB of the allocation-free scratch");
        b.p = c->arena + off;
        b.cap = bytes;
        c->arena_top = off + bytes;
        return SPG_OK;
    }
    if (b.p) {
        SPG_HIP(hipStreamSynchronize(c->stream));
        SPG_HIP(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t cap = bytes + bytes / 8;  // headroom for slightly larger products
    hipError_t e = hipMalloc(&b.p, cap);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipMalloc(&b.p, bytes);
        cap = bytes;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return fail(c, SPG_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    b.cap = cap;
    return SPG_OK;
}

static void release(DevBuf& b) {
    if (b.p && !b.fixed) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    b.fixed = false;
}

// Scratch is a bump region sized before a phase starts (never reallocated
// while kernels of the phase may use it).
static int scratch_prepare(spg_context* c, long long max_n) {
    size_t need = (size_t)(4u << 20) + (size_t)16 * (size_t)(max_n + 1);
    SPG_TRY(ensure(c, c->scratch, need));
    c->scratch_top = 0;
    return SPG_OK;
}

void* scratch_alloc(spg_context* c, size_t bytes) {
    size_t off = (c->scratch_top + 255) & ~(size_t)255;
    if (off + bytes > c->scratch.cap) return nullptr;
    c->scratch_top = off + bytes;
    return (char*)c->scratch.p + off;
}

void scratch_reset(spg_context* c) { c->scratch_top = 0; }

void tmark(spg_context* c, int phase, int end) {
    if (!c->cfg.timing || c->plan_mode == 2) return;  // no event nodes in a captured run
    (void)hipEventRecord(c->timer.ev[2 * phase + end], c->stream);
    if (end) c->timer.recorded[phase] = true;
}

static int copy_d2h(spg_context* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SPG_OK;
    SPG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SPG_HIP(hipStreamSynchronize(c->stream));
    return SPG_OK;
}

int readback(spg_context* c, const void* d, size_t bytes, void* h) {
    if (c->noalloc) {  // pageable staging: no pinned allocation
        if (bytes > kReadbackMax) return fail(c, SPG_EINVAL, "readback too large");
        SPG_TRY(copy_d2h(c, c->host_stage, d, bytes));
        memcpy(h, c->host_stage, bytes);
        return SPG_OK;
    }
    if (bytes > c->pinned_small.cap) return fail(c, SPG_EINVAL, "readback too large");
    if (c->plan_mode == 2) {
        if (c->plan_pos + bytes > c->plan.size())
            return fail(c, SPG_EINVAL, "replay log exhausted (batch changed since it was recorded)");
        memcpy(h, c->plan.data() + c->plan_pos, bytes);
        c->plan_pos += bytes;
        return SPG_OK;
    }
    SPG_TRY(copy_d2h(c, c->pinned_small.p, d, bytes));
    memcpy(h, c->pinned_small.p, bytes);
    if (c->plan_mode == 1) {
        const unsigned char* b = (const unsigned char*)h;
        c->plan.insert(c->plan.end(), b, b + bytes);
    }
    return SPG_OK;
}

template <typename T>
static int read_scalar(spg_context* c, const void* src, T* out) {
    return readback(c, src, sizeof(T), out);
}

static void operand_release(Operand& o) {
    release(o.bptr);
    release(o.bidx);
    release(o.bval);
    o = Operand();
}

}  // namespace spg

using namespace spg;

extern "C" {

const char* spgGetErrorString(int status) {
    switch (status) {
        case SPG_OK: return "success";
        case SPG_EINVAL: return "invalid argument";
        case SPG_ENOMEM: return "out of memory";
        case SPG_EHIP: return "HIP runtime error";
        case SPG_EORDER: return "phase called before its prerequisites";
        case SPG_EIO: return "file could not be read or parsed";
        case SPG_ENODEV: return "no usable GPU";
        case SPG_EOVERFLOW: return "size exceeds a representable range";
        default: return "unknown status";
    }
}

const char* spgGetLastError(spg_context* c) { return c ? c->err.c_str() : ""; }

int spgDeviceCount(int* count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    if (count) *count = n;
    return n > 0 ? SPG_OK : SPG_ENODEV;
}

int spgCreate(spg_context** out, int device) {
    if (!out) return SPG_EINVAL;
    *out = nullptr;
    int n = 0;
    if (spgDeviceCount(&n) != SPG_OK || device < 0 || device >= n) return SPG_ENODEV;
    spg_context* c = new (std::nothrow) spg_context();
    if (!c) return SPG_ENOMEM;
    c->device = device;
    c->cfg.tile_log2 = kTileMaxLog2;
    c->cfg.item_products = 8192;
    c->cfg.long_row_ub = kHashMaxCount;
    c->cfg.xcd_swizzle = 1;
    c->cfg.timing = 1;
    c->cfg.expand_mode = SPG_EXPAND_ROWS;
    c->cfg.pred_margin = 0.25f;
    c->cfg.pred_slack = 8;
    c->cfg.heavy_factor = 50;
    c->cfg.rank_min_count = 128;
    c->cfg.transpose_mode = SPG_TRANSPOSE_AUTO;
    int rc = SPG_OK;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        rc = SPG_EHIP;
    }
    c->own_stream = rc == SPG_OK;
    if (rc == SPG_OK) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->num_cu = prop.multiProcessorCount;
    }
    for (int i = 0; rc == SPG_OK && i < 2 * SPG_NUM_PHASES; ++i)
        if (hipEventCreate(&c->timer.ev[i]) != hipSuccess) rc = SPG_EHIP;
    if (rc == SPG_OK && hipHostMalloc(&c->pinned_small.p, 4096, hipHostMallocDefault) != hipSuccess)
        rc = SPG_ENOMEM;
    if (rc != SPG_OK) {
        (void)hipGetLastError();
        spgDestroy(c);
        return rc;
    }
    c->pinned_small.cap = 4096;
    *out = c;
    return SPG_OK;
}

}  // extern "C"

namespace spg {

// Every device buffer of a context (the operands' storage aside).
template <typename F>
static void for_each_buffer(spg_context* c, F f) {
    DevBuf* bufs[] = {&c->ub, &c->ntile, &c->vstart, &c->vrow_row, &c->vlo, &c->vwid,
                      &c->counters, &c->arow_ptr, &c->arow_k, &c->arow_v, &c->tilebits, &c->feat,
                      &c->bsplit,
                      &c->vub,
                      &c->vbase, &c->cursor, &c->kbin, &c->kperm, &c->kbin_off, &c->kitems,
                      &c->kitem_off, &c->inter, &c->bpair, &c->ebase, &c->vclass, &c->vlist,
                      &c->vclass_off, &c->vcnt, &c->vout, &c->crow, &c->ccol, &c->cval,
                      &c->pred, &c->vcap, &c->pbase, &c->fbase, &c->pcol, &c->pval,
                      &c->fcol, &c->fval, &c->rwinb, &c->kwinb, &c->scratch,
                      &c->lat_t, &c->rprod, &c->vprod};
    for (DevBuf* b : bufs) f(*b);
}

void forget_buffers(spg_context* c) {
    for_each_buffer(c, [](DevBuf& b) { b = DevBuf(); });
    c->scratch_top = 0;
}

}  // namespace spg

extern "C" {

int spgDestroy(spg_context* c) {
    if (!c) return SPG_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    graph_release(c);
    operand_release(c->A);
    operand_release(c->B);
    for_each_buffer(c, [](DevBuf& b) { release(b); });
    if (c->pinned_small.p) (void)hipHostFree(c->pinned_small.p);
    for (auto& e : c->timer.ev)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
        if (c->join_ev[i]) (void)hipEventDestroy(c->join_ev[i]);
        if (c->aux[i]) (void)hipStreamDestroy(c->aux[i]);
    }
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SPG_OK;
}

int spgSetStream(spg_context* c, void* stream) {
    if (!c) return SPG_EINVAL;
    SPG_HIP(hipStreamSynchronize(c->stream));
    graph_release(c);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
    return SPG_OK;
}

int spgGetConfig(spg_context* c, spg_config* cfg) {
    if (!c || !cfg) return SPG_EINVAL;
    *cfg = c->cfg;
    return SPG_OK;
}

int spgSetConfig(spg_context* c, const spg_config* cfg) {
    if (!c || !cfg) return SPG_EINVAL;
    if (cfg->tile_log2 < 5 || cfg->tile_log2 > kTileMaxLog2)
        return fail(c, SPG_EINVAL, "tile_log2 must be in [5, 13]");
    if (cfg->item_products < 64 || cfg->item_products > (1 << 20))
        return fail(c, SPG_EINVAL, "item_products must be in [64, 2^20]");
    if (cfg->long_row_ub < 1 || cfg->long_row_ub > kHashMaxCount)
        return fail(c, SPG_EINVAL, "long_row_ub must be in [1, 4096]");
    if (cfg->expand_mode != SPG_EXPAND_KBINS && cfg->expand_mode != SPG_EXPAND_ROWS)
        return fail(c, SPG_EINVAL, "expand_mode must be SPG_EXPAND_KBINS or SPG_EXPAND_ROWS");
    if (!(cfg->pred_margin >= 0.f && cfg->pred_margin <= 100.f))
        return fail(c, SPG_EINVAL, "pred_margin must be in [0, 100]");
    if (cfg->pred_slack < 0 || cfg->pred_slack > (1 << 20))
        return fail(c, SPG_EINVAL, "pred_slack must be in [0, 2^20]");
    if (cfg->pred_tight != 0 && cfg->pred_tight != 1)
        return fail(c, SPG_EINVAL, "pred_tight must be 0 or 1");
    if (cfg->heavy_factor < 1 || cfg->heavy_factor > (1 << 20))
        return fail(c, SPG_EINVAL, "heavy_factor must be in [1, 2^20]");
    if (cfg->rank_min_count < 0 || cfg->rank_min_count > kHashMaxCount)
        return fail(c, SPG_EINVAL, "rank_min_count must be in [0, 4096]");
    if (cfg->transpose_mode < SPG_TRANSPOSE_AUTO || cfg->transpose_mode > SPG_TRANSPOSE_SORT)
        return fail(c, SPG_EINVAL, "transpose_mode must be AUTO, ATOMIC or SORT");
    if (cfg->rowchunk_ub < 0) return fail(c, SPG_EINVAL, "rowchunk_ub must be >= 0");
    c->cfg = *cfg;
    c->stage = std::min(c->stage, 0);
    graph_release(c);
    return SPG_OK;
}

int spgSynchronize(spg_context* c) {
    if (!c) return SPG_EINVAL;
    SPG_HIP(hipStreamSynchronize(c->stream));
    return SPG_OK;
}

static int set_operand_device_arrays(spg_context* c, int which, int nrows, int ncols, int nnz) {
    Operand& o = which == SPG_OPERAND_A ? c->A : c->B;
    // a new operand ends batch mode (spgBatchSetOperands sets it again after)
    c->rwin = nullptr;
    c->kwin = nullptr;
    c->batch_count = 0;
    graph_release(c);
    o.num_rows = nrows;
    o.num_cols = ncols;
    o.nnz = nnz;
    o.set = true;
    c->stage = 0;
    return SPG_OK;
}

// src/mm.cu:65-136 (cudaCOO2CM)
int spgCOO2CM(spg_context* c, coo mat, cm* out) {
    if (!c) return SPG_EINVAL;
    if (out) cmInit(out);
    if (mat.nnz < 0 || mat.num_rows < 0 || mat.num_cols < 0 || (mat.nnz > 0 && !mat.data))
        return fail(c, SPG_EINVAL, "invalid coo");
    const bool by_row = mat.type == ROW_MAJOR;
    // the input must be sorted by sortCOO and in range
    for (int e = 0; e < mat.nnz; ++e) {
        const triplet& t = mat.data[e];
        if (t.ridx < 0 || t.ridx >= mat.num_rows || t.cidx < 0 || t.cidx >= mat.num_cols)
            return fail(c, SPG_EINVAL, "coo index out of range at entry " + std::to_string(e));
        if (e > 0) {
            const triplet& p = mat.data[e - 1];
            int pm = by_row ? p.ridx : p.cidx, tm = by_row ? t.ridx : t.cidx;
            int pn = by_row ? p.cidx : p.ridx, tn = by_row ? t.cidx : t.ridx;
            if (tm < pm || (tm == pm && tn < pn))
                return fail(c, SPG_EINVAL, "coo is not sorted (call sortCOO first)");
        }
    }
    SPG_HIP(hipSetDevice(c->device));
    const int which = by_row ? SPG_OPERAND_B : SPG_OPERAND_A;
    Operand& o = which == SPG_OPERAND_A ? c->A : c->B;
    const int nmajor = by_row ? mat.num_rows : mat.num_cols;
    SPG_TRY(ensure(c, o.bptr, sizeof(int) * ((size_t)nmajor + 1)));
    SPG_TRY(ensure(c, o.bidx, sizeof(int) * (size_t)mat.nnz));
    SPG_TRY(ensure(c, o.bval, sizeof(float) * (size_t)mat.nnz));
    SPG_TRY(scratch_prepare(c, (long long)nmajor + mat.nnz + (long long)mat.nnz * 12 / 16));
    triplet* d_trip = (triplet*)scratch_alloc(c, sizeof(triplet) * (size_t)mat.nnz + 16);
    if (!d_trip) return fail(c, SPG_ENOMEM, "coo2cm staging");
    if (mat.nnz > 0)
        SPG_HIP(hipMemcpyAsync(d_trip, mat.data, sizeof(triplet) * (size_t)mat.nnz,
                               hipMemcpyHostToDevice, c->stream));
    SPG_TRY(launch_coo2cm(c, d_trip, mat.nnz, mat.type, nmajor, (int*)o.bptr.p,
                          (int*)o.bidx.p, (float*)o.bval.p));
    o.ptr = (const int*)o.bptr.p;
    o.idx = (const int*)o.bidx.p;
    o.val = (const float*)o.bval.p;
    o.owned = true;
    SPG_TRY(set_operand_device_arrays(c, which, mat.num_rows, mat.num_cols, mat.nnz));
    SPG_HIP(hipStreamSynchronize(c->stream));
    if (out) {
        out->ptr = (int*)malloc(sizeof(int) * ((size_t)nmajor + 1));
        out->idx = (int*)malloc(sizeof(int) * ((size_t)mat.nnz + 1));
        out->val = (float*)malloc(sizeof(float) * ((size_t)mat.nnz + 1));
        if (!out->ptr || !out->idx || !out->val) {
            cmFree(out);
            return fail(c, SPG_ENOMEM, "host copy of the compressed matrix");
        }
        SPG_TRY(copy_d2h(c, out->ptr, o.ptr, sizeof(int) * ((size_t)nmajor + 1)));
        SPG_TRY(copy_d2h(c, out->idx, o.idx, sizeof(int) * (size_t)mat.nnz));
        SPG_TRY(copy_d2h(c, out->val, o.val, sizeof(float) * (size_t)mat.nnz));
        out->nnz = mat.nnz;
        out->num_rows = mat.num_rows;
        out->num_cols = mat.num_cols;
        out->type = mat.type;
    }
    return SPG_OK;
}

int spgSetOperand(spg_context* c, int which, const cm* mat) {
    if (!c || !mat || (which != SPG_OPERAND_A && which != SPG_OPERAND_B))
        return fail(c, SPG_EINVAL, "bad operand");
    const bool want_csc = which == SPG_OPERAND_A;
    if ((mat->type == COL_MAJOR) != want_csc)
        return fail(c, SPG_EINVAL, want_csc ? "operand A must be CSC (COL_MAJOR)"
                                            : "operand B must be CSR (ROW_MAJOR)");
    const int nmajor = want_csc ? mat->num_cols : mat->num_rows;
    const int nminor = want_csc ? mat->num_rows : mat->num_cols;
    if (nmajor < 0 || nminor < 0 || mat->nnz < 0 || !mat->ptr ||
        (mat->nnz > 0 && (!mat->idx || !mat->val)))
        return fail(c, SPG_EINVAL, "invalid compressed matrix");
    if (mat->ptr[0] != 0 || mat->ptr[nmajor] != mat->nnz)
        return fail(c, SPG_EINVAL, "ptr does not span nnz");
    for (int q = 0; q < nmajor; ++q) {
        if (mat->ptr[q + 1] < mat->ptr[q]) return fail(c, SPG_EINVAL, "ptr not monotone");
        for (int e = mat->ptr[q]; e < mat->ptr[q + 1]; ++e) {
            if (mat->idx[e] < 0 || mat->idx[e] >= nminor)
                return fail(c, SPG_EINVAL, "index out of range");
            if (e > mat->ptr[q] && mat->idx[e] < mat->idx[e - 1])
                return fail(c, SPG_EINVAL, "indices not sorted inside a row/column");
        }
    }
    SPG_HIP(hipSetDevice(c->device));
    Operand& o = want_csc ? c->A : c->B;
    SPG_TRY(ensure(c, o.bptr, sizeof(int) * ((size_t)nmajor + 1)));
    SPG_TRY(ensure(c, o.bidx, sizeof(int) * (size_t)mat->nnz));
    SPG_TRY(ensure(c, o.bval, sizeof(float) * (size_t)mat->nnz));
    SPG_HIP(hipMemcpyAsync(o.bptr.p, mat->ptr, sizeof(int) * ((size_t)nmajor + 1),
                           hipMemcpyHostToDevice, c->stream));
    if (mat->nnz > 0) {
        SPG_HIP(hipMemcpyAsync(o.bidx.p, mat->idx, sizeof(int) * (size_t)mat->nnz,
                               hipMemcpyHostToDevice, c->stream));
        SPG_HIP(hipMemcpyAsync(o.bval.p, mat->val, sizeof(float) * (size_t)mat->nnz,
                               hipMemcpyHostToDevice, c->stream));
    }
    SPG_HIP(hipStreamSynchronize(c->stream));
    o.ptr = (const int*)o.bptr.p;
    o.idx = (const int*)o.bidx.p;
    o.val = (const float*)o.bval.p;
    o.owned = true;
    return set_operand_device_arrays(c, which, mat->num_rows, mat->num_cols, mat->nnz);
}

int spgSetOperandDevice(spg_context* c, int which, const int* d_ptr, const int* d_idx,
                        const float* d_val, int num_rows, int num_cols, int nnz) {
    if (!c || (which != SPG_OPERAND_A && which != SPG_OPERAND_B) || !d_ptr || num_rows < 0 ||
        num_cols < 0 || nnz < 0 || (nnz > 0 && (!d_idx || !d_val)))
        return fail(c, SPG_EINVAL, "bad device operand");
    Operand& o = which == SPG_OPERAND_A ? c->A : c->B;
    o.ptr = d_ptr;
    o.idx = d_idx;
    o.val = d_val;
    o.owned = false;
    return set_operand_device_arrays(c, which, num_rows, num_cols, nnz);
}

// src/mm.cu:150-193 (cudaInitGEMM)
int spgInitGEMM(spg_context* c, cm* C, int64_t* nprod) {
    if (!c) return SPG_EINVAL;
    if (!c->A.set || !c->B.set) return fail(c, SPG_EORDER, "operands A and B must be set first");
    if (c->A.num_cols != c->B.num_rows)
        return fail(c, SPG_EINVAL, "A.num_cols != B.num_rows");
    SPG_HIP(hipSetDevice(c->device));
    c->stage = 0;
    c->have_pred = false;
    c->m = c->A.num_rows;
    c->K = c->A.num_cols;
    c->n = c->B.num_cols;
    c->W = 1 << c->cfg.tile_log2;
    c->T = std::max(1, (c->n + c->W - 1) / c->W);
    if (c->batch_count > 0) {  // tiles of the widest product's column window
        c->T = 1;
        for (int32_t w : c->b_ncols) c->T = std::max(c->T, (w + c->W - 1) / c->W);
    }
    if (c->T > 1024)
        return fail(c, SPG_EINVAL, "B has more than 1024 column tiles; raise tile_log2");
    for (auto& r : c->timer.recorded) r = false;
    tmark(c, SPG_PHASE_INIT, 0);
    {
        // the transpose of A (scratch_prepare gives 4 MB + 16 B per element)
        const long long tr = (long long)(transpose_scratch(c, c->m, c->A.nnz) / 16) + 1;
        SPG_TRY(scratch_prepare(c, std::max<long long>(std::max<long long>(c->m, c->K) + 1, tr)));
    }
    SPG_TRY(ensure(c, c->ub, sizeof(int64_t) * ((size_t)c->m + 1)));
    SPG_TRY(ensure(c, c->ntile, sizeof(int32_t) * ((size_t)c->m + 1)));
    SPG_TRY(ensure(c, c->vstart, sizeof(int32_t) * ((size_t)c->m + 1)));
    SPG_TRY(ensure(c, c->arow_ptr, sizeof(int32_t) * ((size_t)c->m + 1)));
    SPG_TRY(ensure(c, c->arow_k, sizeof(int32_t) * ((size_t)c->A.nnz + 1)));
    SPG_TRY(ensure(c, c->arow_v, sizeof(float) * ((size_t)c->A.nnz + 1)));
    SPG_TRY(ensure(c, c->feat, sizeof(float) * 4 * ((size_t)c->m + 1)));
    SPG_TRY(launch_init(c));
    int32_t V = 0;
    SPG_TRY(read_scalar<int32_t>(c, (const int32_t*)c->vstart.p + c->m, &V));
    c->V = V;
    c->have_long = c->V > c->m;
    SPG_TRY(ensure(c, c->vrow_row, sizeof(int32_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->vlo, sizeof(int32_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->vwid, sizeof(int32_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->counters, sizeof(uint64_t) * kCtrWords));
    SPG_TRY(ensure(c, c->vub, sizeof(int64_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->vbase, sizeof(int64_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->cursor, sizeof(int32_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->vclass, sizeof(int32_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->vlist, sizeof(int32_t) * ((size_t)c->V + 1)));
    SPG_TRY(ensure(c, c->vclass_off, sizeof(int64_t) * (kNumVClasses + 1)));
    if (c->have_long)
        SPG_TRY(ensure(c, c->bsplit, sizeof(int32_t) * (size_t)c->K * (size_t)(c->T + 1)));
    SPG_TRY(scratch_prepare(c, std::max<long long>(c->V, c->K) + 1));
    SPG_TRY(launch_init_vrows(c));
    unsigned long long* d_lim = (unsigned long long*)scratch_alloc(c, 4 * sizeof(uint64_t));
    if (!d_lim) return fail(c, SPG_ENOMEM, "init scratch");
    SPG_TRY(launch_limits(c, d_lim));
    int64_t lim[4];
    SPG_TRY(readback(c, d_lim, sizeof(lim), lim));
    if (spgSlotLimitStatus(lim[0], lim[1], lim[2]) != SPG_OK)
        return fail(c, SPG_EOVERFLOW,
                    "a virtual row (bound " + std::to_string(lim[0]) + ", untiled " +
                        std::to_string(lim[1]) + ") or an expansion chunk (longest B row " +
                        std::to_string(lim[2]) +
                        ") reaches 2^31 products; raise long_row_ub / lower tile_log2 or "
                        "split the product");
    const int64_t np = lim[3];
    c->nprod = np;
    SPG_TRY(ensure(c, c->inter, inter_bytes(np)));
    tmark(c, SPG_PHASE_INIT, 1);
    if (C) {
        cmInit(C);
        C->type = ROW_MAJOR;
        C->num_rows = c->m;
        C->num_cols = c->n;
        C->nnz = np > INT_MAX ? INT_MAX : (int)np;
    }
    if (nprod) *nprod = np;
    c->stage = 1;
    return SPG_OK;
}

// src/mm.cu:263-286 (cudaCategorizeGEMM)
int spgCategorizeGEMM(spg_context* c, int64_t* bin_counts) {
    if (!c) return SPG_EINVAL;
    if (c->stage < 1) return fail(c, SPG_EORDER, "InitGEMM must run first");
    SPG_HIP(hipSetDevice(c->device));
    tmark(c, SPG_PHASE_CATEGORIZE, 0);
    SPG_TRY(scratch_prepare(c, (long long)c->K + 1));
    const size_t K1 = (size_t)c->K + 1;
    SPG_TRY(ensure(c, c->kbin, sizeof(int32_t) * K1));
    SPG_TRY(ensure(c, c->kperm, sizeof(int32_t) * K1));
    SPG_TRY(ensure(c, c->kbin_off, sizeof(int64_t) * (SPG_NUM_KBINS + 1)));
    SPG_TRY(ensure(c, c->kitems, sizeof(int32_t) * K1));
    SPG_TRY(ensure(c, c->kitem_off, sizeof(int64_t) * (K1 + 1)));
    SPG_TRY(launch_categorize(c));
    tmark(c, SPG_PHASE_CATEGORIZE, 1);
    if (bin_counts)
        for (int b = 0; b < SPG_NUM_KBINS; ++b)
            bin_counts[b] = c->kbin_off_h[b + 1] - c->kbin_off_h[b];
    c->stage = 2;
    return SPG_OK;
}

static int expand_common(spg_context* c, bool binned) {
    if (!c) return SPG_EINVAL;
    if (c->stage < (binned ? 2 : 1))
        return fail(c, SPG_EORDER, binned ? "CategorizeGEMM must run first"
                                          : "InitGEMM must run first");
    SPG_HIP(hipSetDevice(c->device));
    SPG_TRY(scratch_prepare(c, c->V + 1));
    if (binned && c->cfg.expand_mode == SPG_EXPAND_ROWS) {
        SPG_TRY(launch_expand_rows(c));  // records its own timer around the kernels
    } else {
        SPG_TRY(spg::fill<int>(c, (int*)c->cursor.p, 0, c->V));
        tmark(c, SPG_PHASE_EXPAND, 0);
        SPG_TRY(binned ? launch_expand_binned(c) : launch_expand_simple(c));
        tmark(c, SPG_PHASE_EXPAND, 1);
    }
    c->stage = 3;
    return SPG_OK;
}

// src/mm.cu:314-327 (cudaBinGEMM)
int spgBinGEMM(spg_context* c) { return expand_common(c, true); }
// src/mm.cu:356-364 (cudaSimpleGEMM)
int spgSimpleGEMM(spg_context* c) { return expand_common(c, false); }

int spgSymbolicGEMM(spg_context* c, int64_t* nnz) {
    if (!c) return SPG_EINVAL;
    if (c->stage < 3) return fail(c, SPG_EORDER, "BinGEMM or SimpleGEMM must run first");
    SPG_HIP(hipSetDevice(c->device));
    SPG_TRY(scratch_prepare(c, c->V + 1));
    const size_t V1 = (size_t)c->V + 1;
    SPG_TRY(ensure(c, c->vclass, sizeof(int32_t) * V1));
    SPG_TRY(ensure(c, c->vlist, sizeof(int32_t) * V1));
    SPG_TRY(ensure(c, c->vclass_off, sizeof(int64_t) * (kNumVClasses + 1)));
    SPG_TRY(ensure(c, c->vcnt, sizeof(int64_t) * V1));
    SPG_TRY(ensure(c, c->vout, sizeof(int64_t) * (V1 + 1)));
    SPG_TRY(ensure(c, c->tilebits, sizeof(uint32_t) * 256 * V1));
    SPG_TRY(launch_symbolic(c));
    int64_t total = 0;
    SPG_TRY(read_scalar<int64_t>(c, (const int64_t*)c->vout.p + c->V, &total));
    c->nnzC = total;
    SPG_TRY(ensure(c, c->crow, sizeof(int64_t) * ((size_t)c->m + 1)));
    SPG_TRY(ensure(c, c->ccol, sizeof(int32_t) * (size_t)total));
    SPG_TRY(ensure(c, c->cval, sizeof(float) * (size_t)total));
    SPG_TRY(launch_rowptr(c));
    if (nnz) *nnz = total;
    c->stage = 4;
    return SPG_OK;
}

// src/mm.cu:407-450 (cudaMergeGEMM).  Runs the symbolic pass first when the
// caller follows the reference order (Bin -> Merge) without it.
int spgMergeGEMM(spg_context* c) {
    if (!c) return SPG_EINVAL;
    if (c->stage == 3) SPG_TRY(spgSymbolicGEMM(c, nullptr));
    if (c->stage < 4) return fail(c, SPG_EORDER, "expansion must run first");
    SPG_HIP(hipSetDevice(c->device));
    SPG_TRY(scratch_prepare(c, c->V + 1));
    SPG_TRY(launch_merge(c));
    c->stage = 5;
    return SPG_OK;
}

int spgMultiply(spg_context* c) {
    SPG_TRY(spgInitGEMM(c, nullptr, nullptr));
    SPG_TRY(spgCategorizeGEMM(c, nullptr));
    SPG_TRY(spgBinGEMM(c));
    SPG_TRY(spgSymbolicGEMM(c, nullptr));
    return spgMergeGEMM(c);
}

int spgPredictRows(spg_context* c, const spg_predictor* p, float* pred) {
    if (!c || !p) return SPG_EINVAL;
    if (c->stage < 1) return fail(c, SPG_EORDER, "InitGEMM must run first");
    if (predictor_device(p) != c->device)
        return fail(c, SPG_EINVAL, "predictor and context are on different devices");
    SPG_HIP(hipSetDevice(c->device));
    SPG_TRY(ensure(c, c->pred, sizeof(float) * ((size_t)c->m + 1)));
    tmark(c, SPG_PHASE_PREDICT, 0);
    SPG_TRY(predictor_run(c, p, 1.f, (float*)c->pred.p));
    tmark(c, SPG_PHASE_PREDICT, 1);
    c->have_pred = true;
    if (pred) SPG_TRY(copy_d2h(c, pred, c->pred.p, sizeof(float) * (size_t)c->m));
    return SPG_OK;
}

int spgSetRowPredictions(spg_context* c, const float* pred) {
    if (!c || (!pred && c->m > 0)) return SPG_EINVAL;
    if (c->stage < 1) return fail(c, SPG_EORDER, "InitGEMM must run first");
    SPG_HIP(hipSetDevice(c->device));
    SPG_TRY(ensure(c, c->pred, sizeof(float) * ((size_t)c->m + 1)));
    if (c->m > 0) {
        SPG_HIP(hipMemcpyAsync(c->pred.p, pred, sizeof(float) * (size_t)c->m,
                               hipMemcpyHostToDevice, c->stream));
        SPG_HIP(hipStreamSynchronize(c->stream));
    }
    c->have_pred = true;
    return SPG_OK;
}

// Merge sized by the row predictions instead of the symbolic pass; rows
// that overflow their predicted capacity are merged again exactly.
int spgLearnedMergeGEMM(spg_context* c, spg_learned_stats* st) {
    if (!c) return SPG_EINVAL;
    if (c->stage < 3) return fail(c, SPG_EORDER, "BinGEMM or SimpleGEMM must run first");
    if (!c->have_pred)
        return fail(c, SPG_EORDER, "spgPredictRows or spgSetRowPredictions must run first");
    SPG_HIP(hipSetDevice(c->device));
    SPG_TRY(scratch_prepare(c, c->V + 1));
    SPG_TRY(launch_learned_merge(c, st));
    c->stage = 5;
    return SPG_OK;
}

int spgMultiplyLearned(spg_context* c, const spg_predictor* p, spg_learned_stats* st) {
    SPG_TRY(spgInitGEMM(c, nullptr, nullptr));
    SPG_TRY(spgCategorizeGEMM(c, nullptr));
    SPG_TRY(spgBinGEMM(c));
    SPG_TRY(spgPredictRows(c, p, nullptr));
    return spgLearnedMergeGEMM(c, st);
}

int spgGetC(spg_context* c, spg_csr64* out) {
    if (!c || !out) return SPG_EINVAL;
    if (c->stage < 5) return fail(c, SPG_EORDER, "MergeGEMM must run first");
    SPG_HIP(hipSetDevice(c->device));
    out->ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)c->m + 1));
    out->idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c->nnzC + 1));
    out->val = (float*)malloc(sizeof(float) * (size_t)(c->nnzC + 1));
    if (!out->ptr || !out->idx || !out->val) {
        spgFreeCSR64(out);
        return fail(c, SPG_ENOMEM, "host C");
    }
    SPG_TRY(copy_d2h(c, out->ptr, c->crow.p, sizeof(int64_t) * ((size_t)c->m + 1)));
    SPG_TRY(copy_d2h(c, out->idx, c->ccol.p, sizeof(int32_t) * (size_t)c->nnzC));
    SPG_TRY(copy_d2h(c, out->val, c->cval.p, sizeof(float) * (size_t)c->nnzC));
    out->nnz = c->nnzC;
    out->num_rows = c->m;
    out->num_cols = c->n;
    return SPG_OK;
}

int spgGetCDevice(spg_context* c, const int64_t** ptr, const int32_t** idx, const float** val,
                  int64_t* nnz) {
    if (!c) return SPG_EINVAL;
    if (c->stage < 5) return fail(c, SPG_EORDER, "MergeGEMM must run first");
    if (ptr) *ptr = (const int64_t*)c->crow.p;
    if (idx) *idx = (const int32_t*)c->ccol.p;
    if (val) *val = (const float*)c->cval.p;
    if (nnz) *nnz = c->nnzC;
    return SPG_OK;
}

int spgGetUpperBound(spg_context* c, int64_t* ub) {
    if (!c || !ub) return SPG_EINVAL;
    if (c->stage < 1) return fail(c, SPG_EORDER, "InitGEMM must run first");
    return copy_d2h(c, ub, c->ub.p, sizeof(int64_t) * (size_t)c->m);
}

int spgGetKBins(spg_context* c, int32_t* bin_of_k, int32_t* perm, int64_t* offsets) {
    if (!c) return SPG_EINVAL;
    if (c->stage < 2) return fail(c, SPG_EORDER, "CategorizeGEMM must run first");
    if (bin_of_k) SPG_TRY(copy_d2h(c, bin_of_k, c->kbin.p, sizeof(int32_t) * (size_t)c->K));
    if (perm)
        SPG_TRY(copy_d2h(c, perm, c->kperm.p,
                         sizeof(int32_t) * (size_t)c->kbin_off_h[SPG_NUM_KBINS]));
    if (offsets)
        for (int b = 0; b <= SPG_NUM_KBINS; ++b) offsets[b] = c->kbin_off_h[b];
    return SPG_OK;
}

int spgGetRowCounts(spg_context* c, int64_t* cnt) {
    if (!c || !cnt) return SPG_EINVAL;
    if (c->stage < 4) return fail(c, SPG_EORDER, "SymbolicGEMM must run first");
    std::vector<int64_t> rp((size_t)c->m + 1);
    SPG_TRY(copy_d2h(c, rp.data(), c->crow.p, sizeof(int64_t) * ((size_t)c->m + 1)));
    for (int i = 0; i < c->m; ++i) cnt[i] = rp[i + 1] - rp[i];
    return SPG_OK;
}

int spgGetIntermediateByRow(spg_context* c, int64_t* row_ptr, int32_t* cols, float* vals) {
    if (!c) return SPG_EINVAL;
    if (c->stage < 3) return fail(c, SPG_EORDER, "expansion must run first");
    SPG_HIP(hipSetDevice(c->device));
    if (row_ptr) {
        SPG_TRY(scratch_prepare(c, c->m + 1));
        int64_t* d = (int64_t*)scratch_alloc(c, sizeof(int64_t) * ((size_t)c->m + 1));
        if (!d) return fail(c, SPG_ENOMEM, "scratch");
        SPG_TRY(launch_regroup_intermediate(c, d));
        SPG_TRY(copy_d2h(c, row_ptr, d, sizeof(int64_t) * ((size_t)c->m + 1)));
    }
    if (cols || vals) {  // the grouped layout, split into the two arrays
        const long long np = c->nprod;
        std::vector<int32_t> w((size_t)((np + kIGroup - 1) / kIGroup) * (size_t)(2 * kIGroup));
        if (!w.empty()) SPG_TRY(copy_d2h(c, w.data(), c->inter.p, sizeof(int32_t) * w.size()));
        for (long long s = 0; s < np; ++s) {
            const long long q = islot(s);
            if (cols) cols[s] = w[(size_t)q];
            if (vals) memcpy(&vals[s], &w[(size_t)(q + kIGroup)], sizeof(float));
        }
        if (cols && np > 0 && c->V > 0) {  // narrow rows: 16-bit columns, lo + (c16 - lo) mod 2^16
            const size_t V = (size_t)c->V;
            std::vector<uint16_t> c16((size_t)np);
            std::vector<int64_t> vb(V), vu(V);
            std::vector<int32_t> vl(V), vw(V);
            SPG_TRY(copy_d2h(c, c16.data(), inter_cols16(c), sizeof(uint16_t) * c16.size()));
            SPG_TRY(copy_d2h(c, vb.data(), c->vbase.p, sizeof(int64_t) * V));
            SPG_TRY(copy_d2h(c, vu.data(), c->vub.p, sizeof(int64_t) * V));
            SPG_TRY(copy_d2h(c, vl.data(), c->vlo.p, sizeof(int32_t) * V));
            SPG_TRY(copy_d2h(c, vw.data(), c->vwid.p, sizeof(int32_t) * V));
            for (size_t v = 0; v < V; ++v) {
                if (vw[v] > c->W) continue;
                const uint32_t lo = (uint32_t)vl[v];
                for (int64_t s = vb[v]; s < vb[v] + vu[v]; ++s)
                    cols[s] = (int32_t)(lo + ((c16[(size_t)s] - lo) & 0xFFFFu));
            }
        }
    }
    return SPG_OK;
}

int spgGetPhaseTimes(spg_context* c, double* ms) {
    if (!c || !ms) return SPG_EINVAL;
    SPG_HIP(hipStreamSynchronize(c->stream));
    for (int p = 0; p < SPG_NUM_PHASES; ++p) {
        ms[p] = 0.0;
        if (!c->cfg.timing || !c->timer.recorded[p]) continue;
        float t = 0.f;
        if (hipEventElapsedTime(&t, c->timer.ev[2 * p], c->timer.ev[2 * p + 1]) == hipSuccess)
            ms[p] = t;
        else
            (void)hipGetLastError();
    }
    return SPG_OK;
}

}  // extern "C"
