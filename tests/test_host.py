"""Host side of the boundary (no GPU): the C-ABI library loads and exports
every declared symbol; the COO/CM containers, the Matrix Market loader and
the seeded generators behave as the reference's lib/coo.h + src/coo.cpp
(with defect D11 fixed) and as their documented closed forms."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spgemm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)  # preprocessor lines
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_header_functions_are_exported(spg):
    L = ctypes.CDLL(spg.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 40
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's functions
    assert sorted(spg.exported_symbols()) == decl


def test_struct_layouts(spg):
    assert ctypes.sizeof(spg.Triplet) == 12
    assert ctypes.sizeof(spg.COO) == 24
    assert ctypes.sizeof(spg.CM) == 40


MIRRORS = [("Triplet", "triplet"), ("COO", "coo"), ("CM", "cm"), ("CSR64", "spg_csr64"),
           ("Config", "spg_config"), ("LearnedStats", "spg_learned_stats")]


def test_struct_mirrors_match_header(spg, tmp_path):
    """Every ctypes mirror has the C struct's size and, field by field, its
    offsets (a C program built from include/spgemm.h prints them): the
    library writes whole structs into these mirrors (spgGetConfig,
    spgGetLearnedStats, ...)."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "spgemm.h"', "int main(void) {"]
    for py, c in MIRRORS:
        lines.append(f'    printf("{py} size %zu\\n", sizeof({c}));')
        for f, _ in getattr(spg, py)._fields_:
            lines.append(f'    printf("{py} {f} %zu\\n", offsetof({c}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            py, f, v = ln.split()
            got[(py, f)] = int(v)
    for py, _ in MIRRORS:
        cls = getattr(spg, py)
        assert got[(py, "size")] == ctypes.sizeof(cls), py
        for f, _ in cls._fields_:
            assert got[(py, f)] == getattr(cls, f).offset, (py, f)


def load_mtx_expect(name):
    z = np.load(os.path.join(GOLDEN, f"mtx_{name}.npz"))
    return z


def canon(rows, cols, vals):
    o = np.lexsort((vals, cols, rows))
    return np.asarray(rows)[o], np.asarray(cols)[o], np.asarray(vals, np.float32)[o]


@pytest.mark.parametrize("name", ["general", "symmetric", "skew", "pattern", "integer", "long_line"])
def test_read_mtx_matches_scipy(spg, name):
    M = spg.read_mtx(os.path.join(GOLDEN, f"mtx_{name}.mtx"))
    z = load_mtx_expect(name)
    assert (M.num_rows, M.num_cols) == tuple(z["shape"])
    got = canon(M.rows, M.cols, M.vals)
    exp = canon(z["rows"], z["cols"], z["vals"])
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)


def test_read_mtx_compat_values(spg):
    M = spg.read_mtx(os.path.join(GOLDEN, "mtx_skew.mtx"), compat_values=True)
    assert np.all(M.vals == np.float32(1.1)) and M.nnz == 4


def test_read_mtx_errors(spg, tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    with pytest.raises(spg.SpgError):
        spg.read_mtx(str(bad))
    with pytest.raises(spg.SpgError):
        spg.read_mtx(str(tmp_path / "missing.mtx"))
    oob = tmp_path / "oob.mtx"
    oob.write_text("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n")
    with pytest.raises(spg.SpgError):
        spg.read_mtx(str(oob))


def test_sort_coo_order(spg):
    rng = np.random.default_rng(0)
    r = rng.integers(0, 50, 2000).astype(np.int32)
    c = rng.integers(0, 40, 2000).astype(np.int32)
    v = rng.integers(-3, 3, 2000).astype(np.float32)
    A = spg.HostCOO(r, c, v, 50, 40)
    for t in (spg.ROW_MAJOR, spg.COL_MAJOR):
        S = spg.sort_coo(A, t)
        maj, mnr = (S.rows, S.cols) if t == spg.ROW_MAJOR else (S.cols, S.rows)
        key = maj.astype(np.int64) * 100 + mnr
        assert np.all(np.diff(key) >= 0)
        tie = np.diff(key) == 0
        assert np.all(np.diff(S.vals)[tie] >= 0)
        assert sorted(zip(S.rows, S.cols, S.vals)) == sorted(zip(r, c, v))


def test_copy_coo_and_cm_accessors(spg):
    L = spg.lib()
    c = spg.COO()
    L.spgGenUniformRows(ctypes.byref(c), 10, 20, 3, 5)
    L.setCOOtype(ctypes.byref(c), True)
    d = L.copyCOO(c)
    assert d.nnz == 30 and d.type and d.num_cols == 20
    assert all(d.data[i].cidx == c.data[i].cidx for i in range(30))
    L.freeCOO(ctypes.byref(d))
    L.freeCOO(ctypes.byref(c))
    assert c.nnz == 0 and not c.data
    m = spg.CM()
    L.cmInit(ctypes.byref(m))
    L.cmSetNumRows(ctypes.byref(m), 7)
    L.cmSetNumCols(ctypes.byref(m), 9)
    L.cmSetNNZ(ctypes.byref(m), 11)
    L.cmSetType(ctypes.byref(m), True)
    assert (L.cmGetNumRows(m), L.cmGetNumCols(m), L.cmGetNNZ(m), m.type) == (7, 9, 11, True)
    L.cmFree(ctypes.byref(m))


def test_uniform_rows_generator(spg):
    U = spg.gen_uniform_rows(4096, 4096, 8, seed=1)
    assert U.nnz == 32768
    cnt = np.bincount(U.rows, minlength=4096)
    assert np.all(cnt == 8)
    key = U.rows.astype(np.int64) * 4096 + U.cols
    assert np.all(np.diff(key) > 0)  # sorted, distinct columns per row
    assert U.vals.min() >= -1 and U.vals.max() < 1
    U2 = spg.gen_uniform_rows(4096, 4096, 8, seed=1)
    assert np.array_equal(U.cols, U2.cols) and np.array_equal(U.vals, U2.vals)


def test_stencil_generator(spg):
    S = spg.gen_stencil27(6, 5, 4, seed=3)
    assert S.nnz == (3 * 6 - 2) * (3 * 5 - 2) * (3 * 4 - 2)
    d = S.rows == S.cols
    assert np.all(S.vals[d] == 26.0)
    assert np.all(np.abs(S.vals[~d] + 1.0) <= 0.01 + 1e-7)


def test_rmat_generator(spg):
    R = spg.gen_rmat(12, 12, 16 * 4096, seed=0)
    assert R.num_rows == R.num_cols == 4096
    key = R.rows.astype(np.int64) * 4096 + R.cols
    assert np.all(np.diff(key) > 0)  # duplicates merged, sorted
    assert 0.7 * 16 * 4096 < R.nnz <= 16 * 4096
    deg = np.bincount(R.rows, minlength=4096)
    assert deg[0] > 20 * deg.mean()  # power-law skew toward low indices
    # rectangular with clipping
    Q = spg.gen_rmat(11, 10, 5000, seed=2, clip_rows=1500, clip_cols=700)
    assert Q.num_rows == 1500 and Q.num_cols == 700
    assert Q.rows.max() < 1500 and Q.cols.max() < 700


def test_reference_loader_parity(spg, tmp_path):
    """oracle/_ref/libspgref.so is the reference's own src/coo.cpp (built from
    its sources by oracle/Makefile).  On real general / symmetric files its
    readMTX (values forced to 1.1) + sortCOO must give our compat-mode COO."""
    from oracle import oracle as O
    if not os.path.exists(O.REF_LIB_PATH):
        pytest.skip("reference host library not built (reference tree absent)")
    R = ctypes.CDLL(O.REF_LIB_PATH)
    R.readMTX.argtypes = [ctypes.POINTER(spg.COO), ctypes.c_char_p]
    R.sortCOO.argtypes = [spg.COO]
    R.initCOO.argtypes = [ctypes.POINTER(spg.COO)]
    for name in ("symmetric", "general"):
        path = os.path.join(GOLDEN, f"mtx_{name}.mtx")
        if name == "general":
            # the reference reads the size line as "cols rows": use a square file
            path = str(tmp_path / "sq.mtx")
            with open(path, "w") as f:
                f.write("%%MatrixMarket matrix coordinate real general\n4 4 5\n"
                        "1 1 1.5\n2 3 -2.25\n4 2 3e-1\n3 1 7\n2 3 0.5\n")
        for t in (spg.ROW_MAJOR, spg.COL_MAJOR):
            rc = spg.COO()
            R.initCOO(ctypes.byref(rc))
            rc.type = bool(t)
            R.readMTX(ctypes.byref(rc), path.encode())
            R.sortCOO(rc)
            n = rc.nnz
            ref = [(rc.data[i].ridx, rc.data[i].cidx, rc.data[i].val) for i in range(n)]
            ours = spg.sort_coo(spg.read_mtx(path, compat_values=True), t)
            got = list(zip(ours.rows.tolist(), ours.cols.tolist(), ours.vals.tolist()))
            assert [(a, b, np.float32(v)) for a, b, v in ref] == [(a, b, np.float32(v)) for a, b, v in got]


def test_slot_limit_status(spg):
    """The 32-bit expansion limits become SPG_EOVERFLOW instead of wrapping
    (src/mm.cu:173-180): synthetic virtual-row bounds reduced on the host the
    way InitGEMM reduces them on the GPU, just below and at each limit."""
    OK, OVER, INVAL = 0, 7, 1
    lim = 2 ** 31 - 1 - 64
    rng = np.random.default_rng(0)

    def status(vub, untiled, blen):
        vub = np.asarray(vub, np.int64)
        untiled = np.asarray(untiled, bool)
        mu = int(vub[untiled].max(initial=0))
        return spg.slot_limit_status(int(vub.max(initial=0)), mu, int(np.max(blen, initial=0)))

    vub = rng.integers(0, 1 << 20, 1000)
    untiled = rng.random(1000) < 0.5
    blen = rng.integers(0, 5000, 300)
    assert status(vub, untiled, blen) == OK
    # one tiled virtual row at the cursor limit, then one past it
    v2 = vub.copy()
    v2[np.nonzero(~untiled)[0][0]] = lim
    assert status(v2, untiled, blen) == OK
    v2[np.nonzero(~untiled)[0][0]] = lim + 1
    assert status(v2, untiled, blen) == OVER
    v2[np.nonzero(~untiled)[0][0]] = 3 * 2 ** 31
    assert status(v2, untiled, blen) == OVER
    # 256-entry chunks of untiled rows: bounded by min(longest B row, row bound)
    per = lim // 256
    v3 = vub.copy()
    v3[np.nonzero(untiled)[0][0]] = per + 1
    assert status(v3, untiled, [per]) == OK          # B rows short enough
    assert status(v3, untiled, [per + 1]) == OVER    # both past: a chunk could wrap
    assert status(vub, untiled, [10 ** 9]) == OK     # long B row, small untiled rows
    assert spg.slot_limit_status(-1, 0, 0) == INVAL
