// spg_noalloc.hip - spgMultiplyNoAlloc: the plain product (InitGEMM,
// CategorizeGEMM, BinGEMM, SymbolicGEMM, MergeGEMM) for callers that cannot
// allocate device memory.  Every intermediate buffer of the product (row
// bounds, virtual rows, k-bins, the intermediate products, symbolic counts,
// the scan and partition scratch) is a range of one file-scope device array,
// `scratch`, of SPG_NOALLOC_SCRATCH_BYTES per GPU; the operands and C are the
// caller's device arrays.  No hipMalloc, hipFree, pinned or other allocating
// call is made on the way: readbacks go through a static pageable host
// buffer, and the per-device context is a static object.
//
// Each GPU the library runs on holds its copy of `scratch` (the code object's
// device data) until the process ends.  Calls on one GPU share it and the
// per-device host context, so calls that overlap on one GPU (two host threads,
// or two streams) race; a call returns only after its work has finished on
// the stream.  Everything the kernels need at zero is cleared by fill kernels
// in every call (nothing is assumed about the previous call's contents).
#include <cstring>

#include "spg_internal.h"

// The allocation-free scratch.  In host code this name is only a symbol
// handle: its device address is taken once per device with
// hipGetSymbolAddress and kept in g_scratch, which is what host code passes.
static __device__ __align__(256) char scratch[SPG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

void* g_scratch[kMaxDevices] = {};
int g_num_cu[kMaxDevices] = {};
bool g_ready[kMaxDevices] = {};
spg_context g_ctx[kMaxDevices];
unsigned char g_stage[kMaxDevices][spg::kReadbackMax];

void set_err(char* err, size_t err_len, const char* msg) {
    if (!err || err_len == 0) return;
    strncpy(err, msg, err_len - 1);
    err[err_len - 1] = '\0';
}

// Once per device, with that device current: the device address of
// `scratch` and the CU count.
int setup_device(int device, char* err, size_t err_len) {
    if (g_ready[device]) return SPG_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        set_err(err, err_len, "hipGetSymbolAddress of the allocation-free scratch failed");
        return SPG_EHIP;
    }
    hipDeviceProp_t prop;
    int cu = 256;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        cu = prop.multiProcessorCount;
    else
        (void)hipGetLastError();
    g_scratch[device] = p;
    g_num_cu[device] = cu;
    g_ready[device] = true;
    return SPG_OK;
}

}  // namespace

extern "C" int spgMultiplyNoAlloc(int device, void* stream, const int* a_ptr, const int* a_idx,
                                  const float* a_val, int m, int K, int nnz_a, const int* b_ptr,
                                  const int* b_idx, const float* b_val, int n, int nnz_b,
                                  int64_t* c_ptr, int32_t* c_idx, float* c_val,
                                  int64_t c_capacity, int64_t* c_nnz, char* err,
                                  size_t err_len) {
    set_err(err, err_len, "");
    if (c_nnz) *c_nnz = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) {
        (void)hipGetLastError();
        ndev = 0;
    }
    if (device < 0 || device >= ndev || device >= kMaxDevices) {
        set_err(err, err_len, "no such device");
        return SPG_ENODEV;
    }
    if (m < 0 || K < 0 || n < 0 || nnz_a < 0 || nnz_b < 0 || c_capacity < 0 || !c_nnz ||
        !a_ptr || !b_ptr || !c_ptr || (nnz_a > 0 && (!a_idx || !a_val)) ||
        (nnz_b > 0 && (!b_idx || !b_val)) || (c_capacity > 0 && (!c_idx || !c_val))) {
        set_err(err, err_len, "invalid argument");
        return SPG_EINVAL;
    }
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        set_err(err, err_len, "hipSetDevice failed");
        return SPG_EHIP;
    }
    int rc = setup_device(device, err, err_len);
    if (rc != SPG_OK) return rc;

    spg_context* c = &g_ctx[device];
    spg::forget_buffers(c);
    c->device = device;
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
    c->num_cu = g_num_cu[device];
    c->noalloc = true;
    c->arena = (char*)g_scratch[device];
    c->arena_cap = SPG_NOALLOC_SCRATCH_BYTES;
    c->arena_top = 0;
    c->host_stage = g_stage[device];
    c->cfg = spg_config{};
    c->cfg.tile_log2 = spg::kTileMaxLog2;
    c->cfg.item_products = 8192;
    c->cfg.long_row_ub = spg::kHashMaxCount;
    c->cfg.xcd_swizzle = 1;
    c->cfg.timing = 0;  // no HIP events
    c->cfg.expand_mode = SPG_EXPAND_ROWS;
    c->cfg.pred_margin = 0.25f;
    c->cfg.pred_slack = 8;
    c->cfg.heavy_factor = 50;
    c->cfg.rank_min_count = 128;
    c->cfg.transpose_mode = SPG_TRANSPOSE_AUTO;
    c->rwin = c->kwin = nullptr;
    c->batch_count = 0;
    c->plan_mode = 0;
    c->stage = 0;
    c->nnzC = 0;
    c->have_pred = false;
    c->A = spg::Operand();
    c->A.ptr = a_ptr;
    c->A.idx = a_idx;
    c->A.val = a_val;
    c->A.num_rows = m;
    c->A.num_cols = K;
    c->A.nnz = nnz_a;
    c->A.set = true;
    c->B = spg::Operand();
    c->B.ptr = b_ptr;
    c->B.idx = b_idx;
    c->B.val = b_val;
    c->B.num_rows = K;
    c->B.num_cols = n;
    c->B.nnz = nnz_b;
    c->B.set = true;
    // C is the caller's: written in place, never replaced
    c->crow = {c_ptr, sizeof(int64_t) * ((size_t)m + 1), true};
    c->ccol = {c_idx, sizeof(int32_t) * (size_t)c_capacity, true};
    c->cval = {c_val, sizeof(float) * (size_t)c_capacity, true};

    rc = spgInitGEMM(c, nullptr, nullptr);
    if (rc == SPG_OK) rc = spgCategorizeGEMM(c, nullptr);
    if (rc == SPG_OK) rc = spgBinGEMM(c);
    int64_t total = 0;
    if (rc == SPG_OK) {
        rc = spgSymbolicGEMM(c, &total);
        if (rc == SPG_EOVERFLOW && c->nnzC > c_capacity) total = c->nnzC;  // C too small
    }
    if (rc == SPG_OK) rc = spgMergeGEMM(c);
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == SPG_OK && se != hipSuccess) {
        c->err = hipGetErrorString(se);
        rc = SPG_EHIP;
    }
    *c_nnz = total;
    if (rc != SPG_OK) set_err(err, err_len, c->err.c_str());
    spg::forget_buffers(c);  // nothing of this call survives but the arena's contents
    c->noalloc = false;
    return rc;
}
