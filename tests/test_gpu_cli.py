"""The output leg of the reference driver (src/main.cu:11-45 with
src/mm.cu:438-449): the command-line program spgemm_outer (the counterpart
of bin/outer) on Matrix Market files, and the sweep of bin/launch.sh
(./outer X.mtx X.mtx over a list of matrices) as bench.py --mtx-list.
Checked against the oracle: the printed counts, and C written with
spgWriteMTX and read back with scipy.io.mmread (structure bit-exact, values
within VAL_RTOL; %.9g round-trips FP32 exactly)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from helpers import GOLDEN, csc_of, csr_of

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "spgemm_amd", "lib", "spgemm_outer")
VAL_RTOL = 1e-5
SQUARE = ["mtx_integer", "mtx_long_line", "mtx_skew", "mtx_symmetric"]


def oracle_square(orc, spg, path, compat=False):
    A = spg.read_mtx(path, compat_values=compat)
    n = A.num_rows
    T = (A.rows, A.cols, A.vals)
    (ptr, idx, val), nprod = orc.spgemm_restated(n, n, n, csc_of(T, (n, n)), csr_of(T, (n, n)))
    return A, ptr, idx, val, nprod


def run_cli(args):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("nprod ")
    nprod = int(lines[0].split()[1])
    nnz, nnz_f = (int(x) for x in lines[1].split())
    return nprod, nnz, nnz_f


def read_c(path, n):
    M = scipy.io.mmread(path)
    M = sp.coo_matrix(M)
    o = np.lexsort((M.col, M.row))
    rows, cols, vals = M.row[o], M.col[o], M.data[o]
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))])
    return ptr, cols.astype(np.int32), vals.astype(np.float32)


def rmat_file(spg, tmp_path, s=10):
    R = spg.gen_rmat(s, s, 8 << s, seed=17)
    M = sp.coo_matrix((R.vals.astype(np.float64), (R.rows, R.cols)), shape=(R.num_rows, R.num_cols))
    p = str(tmp_path / f"rmat_s{s}.mtx")
    scipy.io.mmwrite(p, M, precision=9)
    return p


def check_against(orc_out, C_file, n):
    _, ptr, idx, val, _ = orc_out
    gptr, gidx, gval = read_c(C_file, n)
    assert np.array_equal(gptr, ptr) and np.array_equal(gidx, idx)
    err = np.abs(gval.astype(np.float64) - val) / np.maximum(np.abs(val.astype(np.float64)), 1e-30)
    err[(np.abs(gval) <= 1e-30) & (np.abs(val) <= 1e-30)] = 0.0
    assert err.max(initial=0.0) <= VAL_RTOL


@pytest.mark.parametrize("name", SQUARE + ["rmat_s10"])
def test_cli_prints_counts_and_writes_c(spg, orc, tmp_path, name):
    """spgemm_outer A.mtx A.mtx --out C.mtx: the reference's two printed
    counts (structural nnz, values > 1e-4) and C itself."""
    print(f"case cli {name}", flush=True)
    path = rmat_file(spg, tmp_path) if name == "rmat_s10" else os.path.join(GOLDEN, name + ".mtx")
    ref = oracle_square(orc, spg, path)
    out = str(tmp_path / "C.mtx")
    nprod, nnz, nnz_f = run_cli([path, path, "--out", out])
    assert nprod == ref[4] and nnz == ref[1][-1]
    val = ref[3]
    # a value within rounding of the threshold may fall either side of it
    near = np.abs(val.astype(np.float64) - 1e-4) <= 1e-5 * 1e-4
    assert abs(nnz_f - int(np.count_nonzero(val > 1e-4))) <= int(near.sum())
    check_against(ref, out, ref[0].num_rows)


def test_cli_rectangular(spg, orc, tmp_path):
    """Two different files: the 4 x 3 general file times the 3 x 5 pattern
    file (pattern entries read as 1.1, src/coo.cpp:151-168)."""
    print("case cli rectangular", flush=True)
    pa, pb = os.path.join(GOLDEN, "mtx_general.mtx"), os.path.join(GOLDEN, "mtx_pattern.mtx")
    A, B = spg.read_mtx(pa), spg.read_mtx(pb)
    m, K, n = A.num_rows, A.num_cols, B.num_cols
    (ptr, idx, val), nprod = orc.spgemm_restated(
        m, n, K, csc_of((A.rows, A.cols, A.vals), (m, K)), csr_of((B.rows, B.cols, B.vals), (K, n)))
    out = str(tmp_path / "C.mtx")
    g_nprod, nnz, _ = run_cli([pa, pb, "--out", out])
    assert g_nprod == nprod and nnz == ptr[-1]
    check_against((A, ptr, idx, val, nprod), out, m)


def test_cli_compat_values_known_answer(spg, orc, tmp_path):
    """--compat-values reads every value as 1.1 (src/coo.cpp:151-168): each
    entry of C is then the FP32 rounding of (number of contributing
    products) * fl32(1.1 * 1.1), the count being the pattern product (the
    reference's int accumulator truncates it to the count itself, D1)."""
    print("case cli compat values", flush=True)
    path = rmat_file(spg, tmp_path, s=9)
    A = spg.read_mtx(path, compat_values=True)
    assert np.all(A.vals == np.float32(1.1))
    n = A.num_rows
    P = sp.csr_matrix((np.ones(A.nnz), (A.rows, A.cols)), shape=(n, n))
    cnt = (P @ P).tocsr()
    cnt.sort_indices()
    out = str(tmp_path / "C.mtx")
    nprod, nnz, _ = run_cli([path, path, "--compat-values", "--out", out])
    assert nnz == cnt.nnz and nprod == int(cnt.data.sum())
    gptr, gidx, gval = read_c(out, n)
    assert np.array_equal(gptr, cnt.indptr) and np.array_equal(gidx, cnt.indices)
    p11 = float(np.float32(np.float32(1.1) * np.float32(1.1)))
    assert np.array_equal(gval, (cnt.data * p11).astype(np.float32))


def test_cli_simple_and_errors(spg, orc, tmp_path):
    """--simple (SimpleGEMM instead of Categorize + Bin) gives the same C; a
    missing file and a shape mismatch end with status 1 and a message."""
    print("case cli simple + errors", flush=True)
    path = rmat_file(spg, tmp_path, s=9)
    ref = oracle_square(orc, spg, path)
    out = str(tmp_path / "C.mtx")
    nprod, nnz, _ = run_cli([path, path, "--simple", "--out", out])
    assert nnz == ref[1][-1]
    check_against(ref, out, ref[0].num_rows)
    r = subprocess.run([CLI, str(tmp_path / "missing.mtx"), path], capture_output=True, text=True)
    assert r.returncode == 1 and "read A" in r.stderr
    pat = os.path.join(GOLDEN, "mtx_pattern.mtx")  # 3 x 5: pattern times pattern is undefined
    r = subprocess.run([CLI, pat, pat], capture_output=True, text=True)
    assert r.returncode == 1 and "InitGEMM" in r.stderr


def test_mtx_list_sweep(spg, orc, tmp_path):
    """bench.py --mtx-list: one JSON line per matrix of the list (the
    golden square files, a generated R-MAT file, a non-square file that is
    reported and skipped), nprod and nnz(C) equal to the oracle's."""
    print("case mtx-list sweep", flush=True)
    rm = rmat_file(spg, tmp_path)
    names = [os.path.join(GOLDEN, n + ".mtx") for n in SQUARE] + [rm]
    lst = tmp_path / "list.txt"
    odd = [os.path.join(GOLDEN, "mtx_general.mtx"), os.path.join(GOLDEN, "mtx_pattern.mtx")]
    lst.write_text("# matrices\n" + "\n".join(names + odd) + "\n")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--mtx-list", str(lst),
                        "--steps", "2", "--warmup", "1"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    recs = [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert r.returncode == 1, r.stderr  # the non-square files are error lines
    assert len(recs) == len(names) + len(odd), r.stderr[-3000:]
    for p, rec in zip(names, recs):
        _, ptr, _, val, nprod = oracle_square(orc, spg, p)
        assert rec["matrix"] == p and "error" not in rec
        assert rec["nprod"] == nprod and rec["nnz_C"] == ptr[-1]
        assert rec["plain"]["ms"] > 0 and rec["learned"]["ms"] > 0
    assert all("not square" in rec["error"] for rec in recs[len(names):])
