"""spgemm_amd - MI355X outer-product SpGEMM (host-side mirror of the C-ABI).

The native library ``spgemm_amd/lib/libspgemm_amd.so`` (HIP, gfx950) does all
of the work; this module only binds it with ctypes and moves numpy arrays in
and out.  The phase methods of :class:`Context` mirror the reference driver
(fast-squid/spGEMM src/main.cu:20-37):

    ctx.coo2cm(A_coo_col_major)   # cudaCOO2CM(A)      src/mm.cu:65-136
    ctx.coo2cm(B_coo_row_major)   # cudaCOO2CM(B)
    ctx.init_gemm()               # cudaInitGEMM       src/mm.cu:150-193
    ctx.categorize_gemm()         # cudaCategorizeGEMM src/mm.cu:263-286
    ctx.bin_gemm()                # cudaBinGEMM        src/mm.cu:314-327
    ctx.merge_gemm()              # cudaMergeGEMM      src/mm.cu:407-450
    C = ctx.get_c()

There is no CPU fallback: every compute call goes through the HIP library and
raises :class:`SpgError` if it is missing or fails.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_bool, c_char_p, c_double, c_float, c_size_t,
                    c_int, c_int32, c_int64, c_uint64, c_void_p)

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspgemm_amd.so")

ROW_MAJOR = 0
COL_MAJOR = 1
OPERAND_A = 0
OPERAND_B = 1
NUM_KBINS = 9
PHASES = ("init", "categorize", "expand", "symbolic", "merge", "predict", "compact")
MTX_COMPAT_VALUES = 1
EXPAND_KBINS = 0
EXPAND_ROWS = 1
TRANSPOSE_AUTO = 0
TRANSPOSE_ATOMIC = 1
TRANSPOSE_SORT = 2
BATCH_GRAPH = 1
BATCH_LATENCY = 2

STATUS = {0: "SPG_OK", 1: "SPG_EINVAL", 2: "SPG_ENOMEM", 3: "SPG_EHIP", 4: "SPG_EORDER",
          5: "SPG_EIO", 6: "SPG_ENODEV", 7: "SPG_EOVERFLOW"}


class SpgError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {msg}")


class Triplet(Structure):
    _fields_ = [("ridx", c_int), ("cidx", c_int), ("val", c_float)]


class COO(Structure):
    _fields_ = [("data", POINTER(Triplet)), ("nnz", c_int), ("num_cols", c_int),
                ("num_rows", c_int), ("type", c_bool)]


class CM(Structure):
    _fields_ = [("ptr", POINTER(c_int)), ("idx", POINTER(c_int)), ("val", POINTER(c_float)),
                ("nnz", c_int), ("num_rows", c_int), ("num_cols", c_int), ("type", c_bool)]


class CSR64(Structure):
    _fields_ = [("ptr", POINTER(c_int64)), ("idx", POINTER(c_int32)), ("val", POINTER(c_float)),
                ("nnz", c_int64), ("num_rows", c_int32), ("num_cols", c_int32)]


class Config(Structure):
    _fields_ = [("tile_log2", c_int32), ("item_products", c_int32), ("long_row_ub", c_int64),
                ("xcd_swizzle", c_int32), ("timing", c_int32), ("expand_mode", c_int32),
                ("pred_margin", c_float), ("pred_slack", c_int32), ("pred_tight", c_int32),
                ("heavy_factor", c_int32), ("rank_min_count", c_int32),
                ("transpose_mode", c_int32), ("rowchunk_ub", c_int32)]


class LearnedStats(Structure):
    _fields_ = [("vrows", c_int64), ("overflow_vrows", c_int64), ("unrecovered", c_int64),
                ("padded_slots", c_int64), ("fallback_slots", c_int64), ("nnz", c_int64)]


TRIPLET_DTYPE = np.dtype([("ridx", np.int32), ("cidx", np.int32), ("val", np.float32)])

_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "initCOO": (None, [POINTER(COO)]),
    "freeCOO": (None, [POINTER(COO)]),
    "setCOOtype": (None, [POINTER(COO), c_bool]),
    "copyCOO": (COO, [COO]),
    "sortCOO": (None, [COO]),
    "readCOO": (None, [POINTER(COO), c_char_p]),
    "readMTX": (None, [POINTER(COO), c_char_p]),
    "spgReadMTX": (c_int, [POINTER(COO), c_char_p, c_int]),
    "spgWriteMTX": (c_int, [POINTER(CSR64), c_char_p]),
    "cmInit": (None, [POINTER(CM)]),
    "cmFree": (None, [POINTER(CM)]),
    "cmGetNumRows": (c_int, [CM]),
    "cmGetNumCols": (c_int, [CM]),
    "cmGetNNZ": (c_int, [CM]),
    "cmSetNumRows": (None, [POINTER(CM), c_int]),
    "cmSetNumCols": (None, [POINTER(CM), c_int]),
    "cmSetNNZ": (None, [POINTER(CM), c_int]),
    "cmSetType": (None, [POINTER(CM), c_bool]),
    "spgFreeCSR64": (None, [POINTER(CSR64)]),
    "spgGenUniformRows": (c_int, [POINTER(COO), c_int, c_int, c_int, c_uint64]),
    "spgGenRMAT": (c_int, [POINTER(COO), c_int, c_int, c_int64, c_double, c_double, c_double,
                           c_int, c_int, c_uint64]),
    "spgGenStencil27": (c_int, [POINTER(COO), c_int, c_int, c_int, c_uint64]),
    "spgCreate": (c_int, [POINTER(c_void_p), c_int]),
    "spgDestroy": (c_int, [c_void_p]),
    "spgSetStream": (c_int, [c_void_p, c_void_p]),
    "spgGetConfig": (c_int, [c_void_p, POINTER(Config)]),
    "spgSetConfig": (c_int, [c_void_p, POINTER(Config)]),
    "spgGetErrorString": (c_char_p, [c_int]),
    "spgGetLastError": (c_char_p, [c_void_p]),
    "spgSynchronize": (c_int, [c_void_p]),
    "spgDeviceCount": (c_int, [POINTER(c_int)]),
    "spgCOO2CM": (c_int, [c_void_p, COO, POINTER(CM)]),
    "spgSetOperand": (c_int, [c_void_p, c_int, POINTER(CM)]),
    "spgSetOperandDevice": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                    c_int]),
    "spgInitGEMM": (c_int, [c_void_p, POINTER(CM), POINTER(c_int64)]),
    "spgCategorizeGEMM": (c_int, [c_void_p, POINTER(c_int64)]),
    "spgSimpleGEMM": (c_int, [c_void_p]),
    "spgBinGEMM": (c_int, [c_void_p]),
    "spgSymbolicGEMM": (c_int, [c_void_p, POINTER(c_int64)]),
    "spgMergeGEMM": (c_int, [c_void_p]),
    "spgGetC": (c_int, [c_void_p, POINTER(CSR64)]),
    "spgGetCDevice": (c_int, [c_void_p, POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p),
                              POINTER(c_int64)]),
    "spgMultiply": (c_int, [c_void_p]),
    "spgGetUpperBound": (c_int, [c_void_p, POINTER(c_int64)]),
    "spgGetKBins": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int64)]),
    "spgGetRowCounts": (c_int, [c_void_p, POINTER(c_int64)]),
    "spgGetIntermediateByRow": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int32),
                                        POINTER(c_float)]),
    "spgGetPhaseTimes": (c_int, [c_void_p, POINTER(c_double)]),
    "spgSlotLimitStatus": (c_int, [c_int64, c_int64, c_int64]),
    "spgMultiplyNoAlloc": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                   c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                   c_void_p, c_void_p, c_int64, POINTER(c_int64), c_char_p,
                                   c_size_t]),
    "spgPredictorCreate": (c_int, [POINTER(c_void_p), c_int, c_uint64]),
    "spgPredictorDestroy": (c_int, [c_void_p]),
    "spgPredictorGetLastError": (c_char_p, [c_void_p]),
    "spgPredictorNumParams": (c_int, []),
    "spgPredictorAddSamples": (c_int, [c_void_p, c_void_p]),
    "spgPredictorTrain": (c_int, [c_void_p, c_int, c_int, c_float, POINTER(c_float)]),
    "spgPredictorGradient": (c_int, [c_void_p, POINTER(c_float), POINTER(c_float)]),
    "spgPredictorGetWeights": (c_int, [c_void_p, POINTER(c_float), POINTER(c_float)]),
    "spgPredictorSetWeights": (c_int, [c_void_p, POINTER(c_float), POINTER(c_float)]),
    "spgPredictorSave": (c_int, [c_void_p, c_char_p]),
    "spgPredictorLoad": (c_int, [c_void_p, c_char_p]),
    "spgPredictRows": (c_int, [c_void_p, c_void_p, POINTER(c_float)]),
    "spgSetRowPredictions": (c_int, [c_void_p, POINTER(c_float)]),
    "spgLearnedMergeGEMM": (c_int, [c_void_p, POINTER(LearnedStats)]),
    "spgMultiplyLearned": (c_int, [c_void_p, c_void_p, POINTER(LearnedStats)]),
    "spgBatchSetOperands": (c_int, [c_void_p, c_int, POINTER(CM), POINTER(CM)]),
    "spgBatchGetLayout": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    "spgBatchMultiply": (c_int, [c_void_p, c_void_p, c_int, POINTER(LearnedStats)]),
    "spgBatchGetC": (c_int, [c_void_p, c_int, POINTER(CSR64)]),
    "spgBatchGetLatency": (c_int, [c_void_p, POINTER(c_double)]),
    "spgTune": (c_int, [c_void_p, c_void_p, c_int, c_uint64, c_char_p, POINTER(Config),
                        POINTER(c_double)]),
    "spgTuneLookup": (c_int, [c_void_p, c_char_p, c_int, POINTER(c_int)]),
}


def exported_symbols():
    """Names the C-ABI exports (every function include/spgemm.h declares)."""
    return sorted(_SIGS)


def lib():
    """Loads the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpgError(6, f"native library missing: {LIB_PATH} (run python -m spgemm_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(POINTER(ctype))


def device_count():
    n = c_int(0)
    lib().spgDeviceCount(byref(n))
    return n.value


# ------------------------------------------------------------------ COO helpers
class HostCOO:
    """numpy-backed COO (rows, cols, vals) with the reference's type flag."""

    def __init__(self, rows, cols, vals, num_rows, num_cols, type=ROW_MAJOR):
        self.rows = np.ascontiguousarray(rows, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float32)
        self.num_rows = int(num_rows)
        self.num_cols = int(num_cols)
        self.type = type

    @property
    def nnz(self):
        return int(self.rows.shape[0])

    def triplets(self):
        t = np.empty(self.nnz, dtype=TRIPLET_DTYPE)
        t["ridx"], t["cidx"], t["val"] = self.rows, self.cols, self.vals
        return t

    def as_ctypes(self):
        """ctypes COO viewing a triplet array (kept alive on the object)."""
        self._trip = self.triplets()
        c = COO()
        c.data = self._trip.ctypes.data_as(POINTER(Triplet))
        c.nnz, c.num_rows, c.num_cols, c.type = self.nnz, self.num_rows, self.num_cols, bool(self.type)
        return c

    def sorted(self, type):
        """sortCOO order (major, minor, value) for the given type."""
        if type == ROW_MAJOR:
            o = np.lexsort((self.vals, self.cols, self.rows))
        else:
            o = np.lexsort((self.vals, self.rows, self.cols))
        return HostCOO(self.rows[o], self.cols[o], self.vals[o], self.num_rows, self.num_cols, type)

    def to_compressed(self, type):
        """(ptr, idx, val) in CSR (ROW_MAJOR) or CSC (COL_MAJOR), host numpy."""
        s = self.sorted(type)
        major = s.rows if type == ROW_MAJOR else s.cols
        minor = s.cols if type == ROW_MAJOR else s.rows
        nmajor = self.num_rows if type == ROW_MAJOR else self.num_cols
        ptr = np.zeros(nmajor + 1, dtype=np.int32)
        np.cumsum(np.bincount(major, minlength=nmajor), out=ptr[1:])
        return ptr, minor.copy(), s.vals.copy()


def _coo_from_lib(c):
    L = lib()
    n = c.nnz
    if n > 0:
        buf = np.ctypeslib.as_array(ctypes.cast(c.data, POINTER(c_int32)), shape=(n * 3,)).copy()
        t = buf.view(TRIPLET_DTYPE)
        out = HostCOO(t["ridx"], t["cidx"], t["val"], c.num_rows, c.num_cols, int(c.type))
    else:
        z = np.zeros(0, np.int32)
        out = HostCOO(z, z, np.zeros(0, np.float32), c.num_rows, c.num_cols, int(c.type))
    L.freeCOO(byref(c))
    return out


def _check(status, ctx=None, what=""):
    if status != 0:
        msg = what
        if ctx is not None:
            m = lib().spgGetLastError(ctx)
            if m:
                msg = f"{what}: {m.decode()}"
        raise SpgError(status, msg)


def gen_uniform_rows(n_rows, n_cols, nnz_per_row, seed):
    c = COO()
    _check(lib().spgGenUniformRows(byref(c), n_rows, n_cols, nnz_per_row, seed), what="gen_uniform_rows")
    return _coo_from_lib(c)


def gen_rmat(row_bits, col_bits, n_draws, a=0.57, b=0.19, c=0.19, seed=0, clip_rows=0, clip_cols=0):
    cc = COO()
    _check(lib().spgGenRMAT(byref(cc), row_bits, col_bits, int(n_draws), a, b, c, clip_rows,
                            clip_cols, seed), what="gen_rmat")
    return _coo_from_lib(cc)


def gen_stencil27(nx, ny, nz, seed):
    c = COO()
    _check(lib().spgGenStencil27(byref(c), nx, ny, nz, seed), what="gen_stencil27")
    return _coo_from_lib(c)


def read_mtx(path, compat_values=False):
    c = COO()
    _check(lib().spgReadMTX(byref(c), path.encode(), MTX_COMPAT_VALUES if compat_values else 0),
           what=f"read_mtx {path}")
    return _coo_from_lib(c)


def slot_limit_status(max_vrow_bound, max_untiled_bound, max_b_row):
    """spgSlotLimitStatus: OK, or OVERFLOW when a count of the expansion
    would not fit its 32-bit slot cursor or chunk offsets."""
    return lib().spgSlotLimitStatus(int(max_vrow_bound), int(max_untiled_bound), int(max_b_row))


def sort_coo(coo, type):
    """Sorts through the native sortCOO (lib/coo.h:39)."""
    cc = coo.as_ctypes()
    cc.type = bool(type)
    lib().sortCOO(cc)
    t = coo._trip
    return HostCOO(t["ridx"].copy(), t["cidx"].copy(), t["val"].copy(), coo.num_rows,
                   coo.num_cols, type)


# ------------------------------------------------------------------ context
class Context:
    """One GPU context (all device state of one product at a time)."""

    def __init__(self, device=0):
        h = c_void_p()
        _check(lib().spgCreate(byref(h), device), what="spgCreate")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().spgDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, status, what):
        _check(status, self.h, what)

    # configuration
    def get_config(self):
        cfg = Config()
        self._c(lib().spgGetConfig(self.h, byref(cfg)), "spgGetConfig")
        return {f: getattr(cfg, f) for f, _ in Config._fields_}

    def set_config(self, **kw):
        cfg = Config()
        self._c(lib().spgGetConfig(self.h, byref(cfg)), "spgGetConfig")
        for k, v in kw.items():
            setattr(cfg, k, v)
        self._c(lib().spgSetConfig(self.h, byref(cfg)), "spgSetConfig")

    # operands
    def coo2cm(self, coo, return_host=False):
        """cudaCOO2CM: coo must be sorted (sortCOO order) for its type."""
        cc = coo.as_ctypes()
        out = CM()
        self._c(lib().spgCOO2CM(self.h, cc, byref(out) if return_host else None), "spgCOO2CM")
        if not return_host:
            return None
        nmajor = coo.num_rows if coo.type == ROW_MAJOR else coo.num_cols
        ptr = np.ctypeslib.as_array(out.ptr, shape=(nmajor + 1,)).copy()
        idx = np.ctypeslib.as_array(out.idx, shape=(max(coo.nnz, 1),))[:coo.nnz].copy()
        val = np.ctypeslib.as_array(out.val, shape=(max(coo.nnz, 1),))[:coo.nnz].copy()
        lib().cmFree(byref(out))
        return ptr, idx, val

    def set_operand(self, which, ptr, idx, val, num_rows, num_cols):
        """Upload a host CSC (which=OPERAND_A) or CSR (OPERAND_B)."""
        ptr = np.ascontiguousarray(ptr, np.int32)
        idx = np.ascontiguousarray(idx, np.int32)
        val = np.ascontiguousarray(val, np.float32)
        m = CM()
        m.ptr, m.idx, m.val = _ptr(ptr, c_int), _ptr(idx, c_int), _ptr(val, c_float)
        m.nnz, m.num_rows, m.num_cols = int(idx.shape[0]), int(num_rows), int(num_cols)
        m.type = which == OPERAND_A
        self._c(lib().spgSetOperand(self.h, which, byref(m)), "spgSetOperand")

    def set_operands_from_coo(self, A, B):
        """Host conversion: A -> CSC, B -> CSR, then upload."""
        self.set_operand(OPERAND_A, *A.to_compressed(COL_MAJOR), A.num_rows, A.num_cols)
        self.set_operand(OPERAND_B, *B.to_compressed(ROW_MAJOR), B.num_rows, B.num_cols)

    # phases
    def init_gemm(self):
        np_ = c_int64(0)
        self._c(lib().spgInitGEMM(self.h, None, byref(np_)), "spgInitGEMM")
        return np_.value

    def categorize_gemm(self):
        counts = np.zeros(NUM_KBINS, np.int64)
        self._c(lib().spgCategorizeGEMM(self.h, _ptr(counts, c_int64)), "spgCategorizeGEMM")
        return counts

    def bin_gemm(self):
        self._c(lib().spgBinGEMM(self.h), "spgBinGEMM")

    def simple_gemm(self):
        self._c(lib().spgSimpleGEMM(self.h), "spgSimpleGEMM")

    def symbolic_gemm(self):
        n = c_int64(0)
        self._c(lib().spgSymbolicGEMM(self.h, byref(n)), "spgSymbolicGEMM")
        return n.value

    def merge_gemm(self):
        self._c(lib().spgMergeGEMM(self.h), "spgMergeGEMM")

    def multiply(self):
        self._c(lib().spgMultiply(self.h), "spgMultiply")

    def synchronize(self):
        self._c(lib().spgSynchronize(self.h), "spgSynchronize")

    def get_c(self):
        """Returns (row_ptr int64[m+1], col int32[nnz], val float32[nnz])."""
        out = CSR64()
        self._c(lib().spgGetC(self.h, byref(out)), "spgGetC")
        m, nnz = out.num_rows, out.nnz
        ptr = np.ctypeslib.as_array(out.ptr, shape=(m + 1,)).copy()
        idx = np.ctypeslib.as_array(out.idx, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(out.val, shape=(max(nnz, 1),))[:nnz].copy()
        lib().spgFreeCSR64(byref(out))
        return ptr, idx, val

    def get_c_device(self):
        p, i, v = c_void_p(), c_void_p(), c_void_p()
        n = c_int64()
        self._c(lib().spgGetCDevice(self.h, byref(p), byref(i), byref(v), byref(n)), "spgGetCDevice")
        return p.value, i.value, v.value, n.value

    # inspection
    def upper_bound(self, m):
        ub = np.zeros(m, np.int64)
        self._c(lib().spgGetUpperBound(self.h, _ptr(ub, c_int64)), "spgGetUpperBound")
        return ub

    def kbins(self, K):
        bins = np.zeros(max(K, 1), np.int32)
        offs = np.zeros(NUM_KBINS + 1, np.int64)
        self._c(lib().spgGetKBins(self.h, _ptr(bins, c_int32), None, _ptr(offs, c_int64)), "spgGetKBins")
        perm = np.zeros(max(int(offs[-1]), 1), np.int32)
        self._c(lib().spgGetKBins(self.h, None, _ptr(perm, c_int32), None), "spgGetKBins")
        return bins[:K], perm[:int(offs[-1])], offs

    def row_counts(self, m):
        cnt = np.zeros(max(m, 1), np.int64)
        self._c(lib().spgGetRowCounts(self.h, _ptr(cnt, c_int64)), "spgGetRowCounts")
        return cnt[:m]

    def intermediate_by_row(self, m, nprod):
        rp = np.zeros(m + 1, np.int64)
        cols = np.zeros(max(nprod, 1), np.int32)
        vals = np.zeros(max(nprod, 1), np.float32)
        self._c(lib().spgGetIntermediateByRow(self.h, _ptr(rp, c_int64), _ptr(cols, c_int32),
                                              _ptr(vals, c_float)), "spgGetIntermediateByRow")
        return rp, cols[:nprod], vals[:nprod]

    def phase_times(self):
        ms = np.zeros(len(PHASES), np.float64)
        self._c(lib().spgGetPhaseTimes(self.h, _ptr(ms, c_double)), "spgGetPhaseTimes")
        return dict(zip(PHASES, ms.tolist()))

    # learned allocation
    def predict_rows(self, predictor, m=None):
        """Runs the MLP over the row features (after init_gemm) and keeps the
        predicted row counts for learned_merge_gemm; returns them (m floats)
        when m is given, else nothing is copied to the host."""
        if m is None:
            self._c(lib().spgPredictRows(self.h, predictor.h, None), "spgPredictRows")
            return None
        pred = np.zeros(max(m, 1), np.float32)
        self._c(lib().spgPredictRows(self.h, predictor.h, _ptr(pred, c_float)), "spgPredictRows")
        return pred[:m]

    def set_row_predictions(self, pred):
        pred = np.ascontiguousarray(pred, np.float32)
        self._c(lib().spgSetRowPredictions(self.h, _ptr(pred, c_float)), "spgSetRowPredictions")

    def learned_merge_gemm(self):
        st = LearnedStats()
        self._c(lib().spgLearnedMergeGEMM(self.h, byref(st)), "spgLearnedMergeGEMM")
        return {f: getattr(st, f) for f, _ in LearnedStats._fields_}

    # batched products
    def batch_set_operands(self, pairs):
        """pairs: list of (A, B) HostCOO operands; A is converted to CSC and
        B to CSR on the host, then the batch is stacked on the GPU."""
        n = len(pairs)
        As, Bs = (CM * n)(), (CM * n)()
        keep = []
        for b, (A, B) in enumerate(pairs):
            for arr, M, typ in ((As, A, COL_MAJOR), (Bs, B, ROW_MAJOR)):
                ptr, idx, val = M.to_compressed(typ)
                keep += [ptr, idx, val]
                x = arr[b]
                x.ptr, x.idx, x.val = _ptr(ptr, c_int), _ptr(idx, c_int), _ptr(val, c_float)
                x.nnz, x.num_rows, x.num_cols, x.type = M.nnz, M.num_rows, M.num_cols, bool(typ)
        self._c(lib().spgBatchSetOperands(self.h, n, As, Bs), "spgBatchSetOperands")
        self.batch_size = n

    def batch_layout(self):
        n = self.batch_size
        ro = np.zeros(n + 1, np.int64)
        co = np.zeros(n, np.int64)
        self._c(lib().spgBatchGetLayout(self.h, _ptr(ro, c_int64), _ptr(co, c_int64)),
                "spgBatchGetLayout")
        return ro, co

    def batch_multiply(self, predictor=None, graph=False, latency=False):
        st = LearnedStats()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_LATENCY if latency else 0)
        self._c(lib().spgBatchMultiply(self.h, predictor.h if predictor else None, flags, byref(st)),
                "spgBatchMultiply")
        return {f: getattr(st, f) for f, _ in LearnedStats._fields_} if predictor else None

    def batch_latency(self, count):
        """Per-product completion times (ms from the start of the run) of the
        last batch_multiply(latency=True)."""
        ms = np.zeros(count, np.float64)
        self._c(lib().spgBatchGetLatency(self.h, _ptr(ms, c_double)), "spgBatchGetLatency")
        return ms

    def batch_get_c(self, b):
        out = CSR64()
        self._c(lib().spgBatchGetC(self.h, b, byref(out)), "spgBatchGetC")
        m, nnz = out.num_rows, out.nnz
        ptr = np.ctypeslib.as_array(out.ptr, shape=(m + 1,)).copy()
        idx = np.ctypeslib.as_array(out.idx, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(out.val, shape=(max(nnz, 1),))[:nnz].copy()
        lib().spgFreeCSR64(byref(out))
        return ptr, idx, val

    # autotuner
    def tune(self, predictor=None, reps=3, seed=0, table=None):
        """Picks the fastest tunables for the operands on this context (and
        keeps them set); returns (config dict, median ms)."""
        cfg = Config()
        ms = c_double(0.0)
        self._c(lib().spgTune(self.h, predictor.h if predictor else None, reps, seed,
                              table.encode() if table else None, byref(cfg), byref(ms)), "spgTune")
        return {f: getattr(cfg, f) for f, _ in Config._fields_}, ms.value

    def tune_lookup(self, table, learned=False):
        found = c_int(0)
        self._c(lib().spgTuneLookup(self.h, table.encode(), int(learned), byref(found)),
                "spgTuneLookup")
        return bool(found.value)

    def multiply_learned(self, predictor):
        st = LearnedStats()
        self._c(lib().spgMultiplyLearned(self.h, predictor.h, byref(st)), "spgMultiplyLearned")
        return {f: getattr(st, f) for f, _ in LearnedStats._fields_}


class Predictor:
    """The 4-64-64-1 row-count MLP (weights on the GPU)."""

    def __init__(self, device=0, seed=0):
        h = c_void_p()
        _check(lib().spgPredictorCreate(byref(h), device, seed), what="spgPredictorCreate")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().spgPredictorDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, status, what):
        if status != 0:
            m = lib().spgPredictorGetLastError(self.h)
            raise SpgError(status, f"{what}: {m.decode() if m else ''}")

    @staticmethod
    def num_params():
        return int(lib().spgPredictorNumParams())

    def add_samples(self, ctx):
        self._c(lib().spgPredictorAddSamples(self.h, ctx.h), "spgPredictorAddSamples")

    def train(self, epochs, batch=4096, lr=1e-3):
        loss = c_float(0.0)
        self._c(lib().spgPredictorTrain(self.h, epochs, batch, lr, byref(loss)), "spgPredictorTrain")
        return loss.value

    def gradient(self):
        """(loss, grad) of the mean squared error over all samples."""
        g = np.zeros(self.num_params(), np.float32)
        loss = c_float(0.0)
        self._c(lib().spgPredictorGradient(self.h, _ptr(g, c_float), byref(loss)),
                "spgPredictorGradient")
        return loss.value, g

    def get_weights(self):
        w = np.zeros(self.num_params(), np.float32)
        norm = np.zeros(8, np.float32)
        self._c(lib().spgPredictorGetWeights(self.h, _ptr(w, c_float), _ptr(norm, c_float)),
                "spgPredictorGetWeights")
        return w, norm

    def set_weights(self, w, norm):
        w = np.ascontiguousarray(w, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        self._c(lib().spgPredictorSetWeights(self.h, _ptr(w, c_float), _ptr(norm, c_float)),
                "spgPredictorSetWeights")

    def save(self, path):
        self._c(lib().spgPredictorSave(self.h, path.encode()), "spgPredictorSave")

    def load(self, path):
        self._c(lib().spgPredictorLoad(self.h, path.encode()), "spgPredictorLoad")



def spgemm(A, B, device=0, ctx=None, simple=False):
    """C = A @ B for HostCOO operands through the full GPU path; returns
    (row_ptr, col, val).  A is converted to CSC and B to CSR on the GPU."""
    own = ctx is None
    ctx = ctx or Context(device)
    try:
        ctx.coo2cm(A.sorted(COL_MAJOR))
        ctx.coo2cm(B.sorted(ROW_MAJOR))
        ctx.init_gemm()
        if simple:
            ctx.simple_gemm()
        else:
            ctx.categorize_gemm()
            ctx.bin_gemm()
        ctx.symbolic_gemm()
        ctx.merge_gemm()
        return ctx.get_c()
    finally:
        if own:
            ctx.close()
