"""spgMultiplyNoAlloc (the allocation-free entry point) through the C-ABI,
checked against spgMultiply on the same operands (the same phases and
kernels with the context's own buffers): structure bit-exact and values
within VAL_RTOL, on two calls in a row with the second product smaller than
the first, then a C array that is too small (SPG_EOVERFLOW and the required
size) and a product larger than the fixed scratch (SPG_ENOMEM, nothing
written).  Device arrays come from the HIP runtime directly (hipMalloc in the
test, never in the entry point)."""
import ctypes

import numpy as np
import pytest

from helpers import csc_of, csr_of

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-5
SPG_OK, SPG_ENOMEM, SPG_EOVERFLOW = 0, 2, 7
H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(max(a.nbytes, 8))
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D) == 0
        return p

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype)
        if out.nbytes:
            assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), D2H) == 0
        return out

    def free_all(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def no_alloc(spg, hip, A, Ash, B, Bsh, capacity):
    m, K = Ash
    n = Bsh[1]
    a_ptr, a_idx, a_val = csc_of(A, Ash)
    b_ptr, b_idx, b_val = csr_of(B, Bsh)
    d = [hip.upload(x) for x in (a_ptr, a_idx, a_val, b_ptr, b_idx, b_val)]
    c_ptr = hip.alloc(8 * (m + 1))
    c_idx = hip.alloc(4 * max(capacity, 1))
    c_val = hip.alloc(4 * max(capacity, 1))
    nnz = ctypes.c_int64(-1)
    err = ctypes.create_string_buffer(512)
    rc = spg.lib().spgMultiplyNoAlloc(0, None, d[0], d[1], d[2], m, K, len(a_idx), d[3], d[4], d[5],
                                      n, len(b_idx), c_ptr, c_idx, c_val, capacity,
                                      ctypes.byref(nnz), err, len(err))
    C = None
    if rc == SPG_OK:
        C = (hip.download(c_ptr, m + 1, np.int64), hip.download(c_idx, nnz.value, np.int32),
             hip.download(c_val, nnz.value, np.float32))
    return rc, nnz.value, C, err.value.decode()


def with_context(spg, ctx, A, Ash, B, Bsh):
    ctx.set_operand(spg.OPERAND_A, *csc_of(A, Ash), *Ash)
    ctx.set_operand(spg.OPERAND_B, *csr_of(B, Bsh), *Bsh)
    ctx.multiply()
    return ctx.get_c()


def same(C, R):
    assert np.array_equal(C[0], R[0]) and np.array_equal(C[1], R[1])
    a, b = C[2].astype(np.float64), R[2].astype(np.float64)
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    err[(np.abs(a) <= 1e-30) & (np.abs(b) <= 1e-30)] = 0.0
    assert err.max(initial=0.0) <= VAL_RTOL
    return float(np.mean(C[2] == R[2])) if len(C[2]) else 1.0


def coo(M):
    return (M.rows, M.cols, M.vals), (M.num_rows, M.num_cols)


def test_no_alloc_matches_multiply(spg, gpu_ctx):
    hip = Hip()
    try:
        # 1: a rectangular product with column-tiled rows (B has 2^14 columns)
        print("case no-alloc 1: R-MAT 2^12 x 2^14 times 2^14 x 2^14", flush=True)
        A, Ash = coo(spg.gen_rmat(12, 14, 8 << 12, seed=41))
        B, Bsh = coo(spg.gen_rmat(14, 14, 4 << 14, seed=42))
        R = with_context(spg, gpu_ctx, A, Ash, B, Bsh)
        rc, nnz, C, msg = no_alloc(spg, hip, A, Ash, B, Bsh, int(R[0][-1]) + 100)
        assert rc == SPG_OK, msg
        assert nnz == R[0][-1]
        print(f"  nnz {nnz}, bitwise-equal values {same(C, R):.6f}")
        # 2: a smaller product right after (the scratch keeps call 1's contents)
        print("case no-alloc 2: R-MAT scale 10, A*A", flush=True)
        A2, sh2 = coo(spg.gen_rmat(10, 10, 8 << 10, seed=43))
        R2 = with_context(spg, gpu_ctx, A2, sh2, A2, sh2)
        rc, nnz, C2, msg = no_alloc(spg, hip, A2, sh2, A2, sh2, int(R2[0][-1]))
        assert rc == SPG_OK, msg
        assert nnz == R2[0][-1]
        same(C2, R2)
        # 3: C one entry too small: the required size comes back
        print("case no-alloc 3: C too small", flush=True)
        rc, nnz, _, msg = no_alloc(spg, hip, A2, sh2, A2, sh2, int(R2[0][-1]) - 1)
        assert rc == SPG_EOVERFLOW and nnz == R2[0][-1], (rc, nnz, msg)
        # 4: 2^15 rows of 64 columns: 2^27 products need 1 GiB of intermediates
        print("case no-alloc 4: larger than the scratch", flush=True)
        A4, sh4 = coo(spg.gen_uniform_rows(1 << 15, 1 << 15, 64, 44))
        rc, nnz, _, msg = no_alloc(spg, hip, A4, sh4, A4, sh4, 1)
        assert rc == SPG_ENOMEM and "allocation-free scratch" in msg, (rc, msg)
        # 5: the entry point still works after the failed call
        print("case no-alloc 5: again after the failures", flush=True)
        rc, nnz, C5, msg = no_alloc(spg, hip, A, Ash, B, Bsh, int(R[0][-1]))
        assert rc == SPG_OK, msg
        same(C5, R)
    finally:
        hip.free_all()
