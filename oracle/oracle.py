"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  See oracle.h for
the reference file:line each function restates.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_double, c_float, c_int, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libspgref.so")

NUM_KBINS = 9
ORC_ROW_MAJOR, ORC_COL_MAJOR = 0, 1
(FAULT_NONE, FAULT_DROP_LAST_K, FAULT_TRANSPOSE_B, FAULT_OFF_BY_ONE_UB, FAULT_VALUE_SIGN,
 FAULT_SORT_MINOR, FAULT_PTR_UNSHIFTED, FAULT_CSR_ROW_IDX) = range(8)

_lib = None


def build():
    """Compiles the oracle (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = POINTER
        sig = {
            "orc_set_fault": (None, [c_int]),
            "orc_get_fault": (c_int, []),
            "orc_sort_coo": (None, [c_int64, c_void_p, c_int]),
            "orc_coo2cm": (None, [c_int64, c_void_p, c_int, c_int32, P(c_int32), P(c_int32), P(c_float)]),
            "orc_csc_to_csr": (None, [c_int32, c_int32, P(c_int32), P(c_int32), P(c_float),
                                      P(c_int32), P(c_int32), P(c_float)]),
            "orc_upper_bound": (c_int64, [c_int32, c_int32, P(c_int32), P(c_int32), P(c_int32),
                                          P(c_int64), P(c_int64)]),
            "orc_kbins": (c_int32, [c_int32, P(c_int32), P(c_int32), c_int64, P(c_int32), P(c_int32),
                                    P(c_int64)]),
            "orc_expand": (None, [c_int32, c_int32, P(c_int32), P(c_int32), P(c_float), P(c_int32),
                                  P(c_int32), P(c_float), P(c_int64), P(c_int32), P(c_float)]),
            "orc_symbolic_from_slots": (c_int64, [c_int32, c_int32, P(c_int64), P(c_int32), P(c_int64)]),
            "orc_merge": (None, [c_int32, c_int32, P(c_int64), P(c_int32), P(c_float), P(c_int64),
                                 P(c_int32), P(c_float)]),
            "orc_spgemm_restated": (c_int64, [c_int32, c_int32, c_int32, P(c_int32), P(c_int32), P(c_float),
                                              P(c_int32), P(c_int32), P(c_float), P(P(c_int64)),
                                              P(P(c_int32)), P(P(c_float)), P(c_int64)]),
            "orc_gustavson": (c_int64, [c_int32, c_int32, P(c_int32), P(c_int32), P(c_float), P(c_int32),
                                        P(c_int32), P(c_float), P(c_int32), c_int32, c_int,
                                        P(P(c_int64)), P(P(c_int32)), P(P(c_double)), P(P(c_double))]),
            "orc_cpu_baseline": (c_int64, [c_int32, c_int32, c_int32, P(c_int32), P(c_int32), P(c_float),
                                           P(c_int32), P(c_int32), P(c_float), c_int32, c_int32, c_int,
                                           P(c_double), P(c_int64), P(c_double)]),
            "orc_free": (None, [c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def set_fault(f):
    lib().orc_set_fault(int(f))


TRIPLET_DTYPE = np.dtype([("ridx", np.int32), ("cidx", np.int32), ("val", np.float32)])


def sort_coo(rows, cols, vals, type):
    t = np.empty(len(rows), TRIPLET_DTYPE)
    t["ridx"], t["cidx"], t["val"] = rows, cols, vals
    lib().orc_sort_coo(len(t), t.ctypes.data, int(type))
    return t["ridx"].copy(), t["cidx"].copy(), t["val"].copy()


def coo2cm(rows, cols, vals, type, nmajor):
    t = np.empty(len(rows), TRIPLET_DTYPE)
    t["ridx"], t["cidx"], t["val"] = rows, cols, vals
    ptr = np.zeros(nmajor + 1, np.int32)
    idx = np.zeros(max(len(t), 1), np.int32)
    val = np.zeros(max(len(t), 1), np.float32)
    lib().orc_coo2cm(len(t), t.ctypes.data, int(type), nmajor, _p(ptr, c_int32), _p(idx, c_int32),
                     _p(val, c_float))
    return ptr, idx[:len(t)], val[:len(t)]


def csc_to_csr(m, K, a_ptr, a_idx, a_val):
    a_ptr, a_idx, a_val = _i32(a_ptr), _i32(a_idx), _f32(a_val)
    nnz = int(a_ptr[K])
    r_ptr = np.zeros(m + 1, np.int32)
    r_idx = np.zeros(max(nnz, 1), np.int32)
    r_val = np.zeros(max(nnz, 1), np.float32)
    lib().orc_csc_to_csr(m, K, _p(a_ptr, c_int32), _p(a_idx, c_int32), _p(a_val, c_float),
                         _p(r_ptr, c_int32), _p(r_idx, c_int32), _p(r_val, c_float))
    return r_ptr, r_idx[:nnz], r_val[:nnz]


def upper_bound(m, K, a_ptr, a_idx, b_ptr):
    a_ptr, a_idx, b_ptr = _i32(a_ptr), _i32(a_idx), _i32(b_ptr)
    ub = np.zeros(max(m, 1), np.int64)
    base = np.zeros(m + 1, np.int64)
    nprod = lib().orc_upper_bound(m, K, _p(a_ptr, c_int32), _p(a_idx, c_int32), _p(b_ptr, c_int32),
                                  _p(ub, c_int64), _p(base, c_int64))
    return ub[:m], base, nprod


def kbins(K, a_ptr, b_ptr, nprod):
    a_ptr, b_ptr = _i32(a_ptr), _i32(b_ptr)
    bins = np.zeros(max(K, 1), np.int32)
    perm = np.zeros(max(K, 1), np.int32)
    offs = np.zeros(NUM_KBINS + 1, np.int64)
    n = lib().orc_kbins(K, _p(a_ptr, c_int32), _p(b_ptr, c_int32), int(nprod), _p(bins, c_int32),
                        _p(perm, c_int32), _p(offs, c_int64))
    return bins[:K], perm[:n], offs


def expand(m, K, A, B, base):
    a_ptr, a_idx, a_val = _i32(A[0]), _i32(A[1]), _f32(A[2])
    b_ptr, b_idx, b_val = _i32(B[0]), _i32(B[1]), _f32(B[2])
    base = np.ascontiguousarray(base, np.int64)
    nprod = int(base[m])
    icol = np.zeros(max(nprod, 1), np.int32)
    ival = np.zeros(max(nprod, 1), np.float32)
    lib().orc_expand(m, K, _p(a_ptr, c_int32), _p(a_idx, c_int32), _p(a_val, c_float),
                     _p(b_ptr, c_int32), _p(b_idx, c_int32), _p(b_val, c_float), _p(base, c_int64),
                     _p(icol, c_int32), _p(ival, c_float))
    return icol[:nprod], ival[:nprod]


def symbolic_from_slots(m, n, base, icol):
    base = np.ascontiguousarray(base, np.int64)
    icol = _i32(icol)
    cnt = np.zeros(max(m, 1), np.int64)
    tot = lib().orc_symbolic_from_slots(m, n, _p(base, c_int64), _p(icol, c_int32), _p(cnt, c_int64))
    return cnt[:m], tot


def spgemm_restated(m, n, K, A, B):
    """The restated method: (row_ptr int64, col int32, val f32), nprod."""
    a_ptr, a_idx, a_val = _i32(A[0]), _i32(A[1]), _f32(A[2])
    b_ptr, b_idx, b_val = _i32(B[0]), _i32(B[1]), _f32(B[2])
    cp, ci, cv = POINTER(c_int64)(), POINTER(c_int32)(), POINTER(c_float)()
    nprod = c_int64()
    nnz = lib().orc_spgemm_restated(m, n, K, _p(a_ptr, c_int32), _p(a_idx, c_int32), _p(a_val, c_float),
                                    _p(b_ptr, c_int32), _p(b_idx, c_int32), _p(b_val, c_float),
                                    byref(cp), byref(ci), byref(cv), byref(nprod))
    ptr = np.ctypeslib.as_array(cp, shape=(m + 1,)).copy()
    idx = np.ctypeslib.as_array(ci, shape=(nnz + 1,))[:nnz].copy()
    val = np.ctypeslib.as_array(cv, shape=(nnz + 1,))[:nnz].copy()
    for q in (cp, ci, cv):
        lib().orc_free(ctypes.cast(q, c_void_p))
    return (ptr, idx, val), nprod.value


def gustavson(m, n, A_csr, B, rows=None, prod32=False):
    """Row-wise Gustavson on A in CSR.  Returns (ptr, idx, val64, sabs)."""
    r_ptr, r_idx, r_val = _i32(A_csr[0]), _i32(A_csr[1]), _f32(A_csr[2])
    b_ptr, b_idx, b_val = _i32(B[0]), _i32(B[1]), _f32(B[2])
    sel = None if rows is None else _i32(rows)
    nr = m if sel is None else len(sel)
    cp, ci = POINTER(c_int64)(), POINTER(c_int32)()
    cv, cs = POINTER(c_double)(), POINTER(c_double)()
    nnz = lib().orc_gustavson(m, n, _p(r_ptr, c_int32), _p(r_idx, c_int32), _p(r_val, c_float),
                              _p(b_ptr, c_int32), _p(b_idx, c_int32), _p(b_val, c_float),
                              None if sel is None else _p(sel, c_int32), nr, 1 if prod32 else 0,
                              byref(cp), byref(ci), byref(cv), byref(cs))
    ptr = np.ctypeslib.as_array(cp, shape=(nr + 1,)).copy()
    idx = np.ctypeslib.as_array(ci, shape=(nnz + 1,))[:nnz].copy()
    val = np.ctypeslib.as_array(cv, shape=(nnz + 1,))[:nnz].copy()
    sab = np.ctypeslib.as_array(cs, shape=(nnz + 1,))[:nnz].copy()
    for q in (cp, ci, cv, cs):
        lib().orc_free(ctypes.cast(q, c_void_p))
    return ptr, idx, val, sab


def cpu_baseline(m, n, K, A, B, row_begin, row_end, nthreads):
    """Times the restated method on rows [row_begin, row_end)."""
    a_ptr, a_idx, a_val = _i32(A[0]), _i32(A[1]), _f32(A[2])
    b_ptr, b_idx, b_val = _i32(B[0]), _i32(B[1]), _f32(B[2])
    t = np.zeros(5, np.float64)
    nnz = c_int64()
    cs = c_double()
    nprod = lib().orc_cpu_baseline(m, n, K, _p(a_ptr, c_int32), _p(a_idx, c_int32), _p(a_val, c_float),
                                   _p(b_ptr, c_int32), _p(b_idx, c_int32), _p(b_val, c_float),
                                   int(row_begin), int(row_end), int(nthreads), _p(t, c_double),
                                   byref(nnz), byref(cs))
    return {"nprod": nprod, "nnz": nnz.value, "checksum": cs.value,
            "ms": dict(zip(("ub", "bins", "expand", "merge", "compact"), t.tolist()))}
