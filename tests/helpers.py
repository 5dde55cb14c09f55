"""Shared helpers of the tests: fixture loading and independent numpy
conversions (no product or oracle code)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_prod(name):
    z = np.load(os.path.join(GOLDEN, f"prod_{name}.npz"))
    A = (z["A_rows"], z["A_cols"], z["A_vals"])
    B = (z["B_rows"], z["B_cols"], z["B_vals"])
    return A, tuple(z["A_shape"]), B, tuple(z["B_shape"]), z


PROD_FIXTURES = ["tiny", "rect", "powerlaw", "dense_rowcol", "n1", "empty", "rand_real"]


def compress(rows, cols, vals, nmajor, by_row):
    """(ptr, idx, val) with entries sorted by (major, minor, value)."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float32)
    major, minor = (rows, cols) if by_row else (cols, rows)
    o = np.lexsort((vals, minor, major))
    ptr = np.zeros(nmajor + 1, np.int32)
    np.cumsum(np.bincount(major, minlength=nmajor), out=ptr[1:])
    return ptr, minor[o].astype(np.int32), vals[o]


def csc_of(A, shape):
    return compress(A[0], A[1], A[2], shape[1], by_row=False)


def csr_of(M, shape):
    return compress(M[0], M[1], M[2], shape[0], by_row=True)


def rel_err(x, ref, floor=None):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.abs(ref) if floor is None else np.maximum(np.abs(ref), floor)
    den = np.where(den == 0, 1.0, den)
    return np.abs(x - ref) / den
