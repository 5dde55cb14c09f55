"""Pins the CPU oracle before it is trusted as the GPU path's checker.

Each oracle function is checked against something it does not compute
itself: golden fixtures made from dense numpy / boolean pattern products
(tests/golden/make_golden.py), closed forms of the 27-point stencil and the
fixed-degree generator, brute-force dense products, scipy, invariants, and a
second independent implementation (row-wise Gustavson vs. the k-driven
expansion + merge).  The fault-injection tests show that a plausible mistake
(a dropped term, a transposed operand, an off-by-one bound, a wrong sign)
makes at least one of these pins fail.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import PROD_FIXTURES, csc_of, csr_of, load_prod, rel_err


def run_restated(O, A, Ash, B, Bsh):
    m, K = Ash
    n = Bsh[1]
    return O.spgemm_restated(m, n, K, csc_of(A, Ash), csr_of(B, Bsh))


def check_against_truth(O, name):
    A, Ash, B, Bsh, z = load_prod(name)
    (ptr, idx, val), nprod = run_restated(O, A, Ash, B, Bsh)
    ok = (np.array_equal(ptr, z["C_ptr"]) and np.array_equal(idx, z["C_idx"])
          and nprod == int(z["nprod"]))
    if not ok:
        return False
    # FP32 products summed in FP64 and rounded once vs the float64 dense
    # product: |err| <= 2^-24 |c| + n*2^-24*sum|ab| (products rounded to f32)
    err = np.abs(val.astype(np.float64) - z["C_val"])
    bound = 2.0 ** -24 * np.abs(z["C_val"]) + 1.001 * 2.0 ** -24 * z["C_sabs"] + 1e-300
    return bool(np.all(err <= bound))


@pytest.mark.parametrize("name", PROD_FIXTURES)
def test_restated_matches_golden(orc, name):
    assert check_against_truth(orc, name)


@pytest.mark.parametrize("name", PROD_FIXTURES)
def test_gustavson_fp64_matches_golden(orc, name):
    A, Ash, B, Bsh, z = load_prod(name)
    m, K = Ash
    ptr, idx, val, sab = orc.gustavson(m, Bsh[1], csr_of(A, Ash), csr_of(B, Bsh))
    assert np.array_equal(ptr, z["C_ptr"]) and np.array_equal(idx, z["C_idx"])
    assert np.allclose(val, z["C_val"], rtol=1e-12, atol=1e-12 * (1 + z["C_sabs"].max(initial=0)))
    assert np.allclose(sab, z["C_sabs"], rtol=1e-12)


def test_tiny_known_answers(orc):
    # hand-checkable entries of the tiny fixture: C(0,3) = 1*2 + 1*(-2) = 0
    # stays structural (numerical zeros kept); the duplicate A(3,4) is summed
    A, Ash, B, Bsh, z = load_prod("tiny")
    (ptr, idx, val), _ = run_restated(orc, A, Ash, B, Bsh)
    row0 = dict(zip(idx[ptr[0]:ptr[1]], val[ptr[0]:ptr[1]]))
    assert 3 in row0 and row0[3] == 0.0
    row3 = dict(zip(idx[ptr[3]:ptr[4]], val[ptr[3]:ptr[4]]))
    # A(3,4) = 0.25 + 0.25, B(4,1) = -1, B(4,6) = 3; A(3,2) = -1.5, B(2,3) = -2, B(2,5) = 0.125
    assert row3 == {1: -0.5, 3: 3.0, 5: -0.1875, 6: 1.5}
    assert ptr[2] == ptr[1] and ptr[5] == ptr[4]  # empty rows 1 and 4 of A


@pytest.mark.parametrize("L", [3, 5, 8, 11])
def test_stencil_closed_forms(orc, spg, L):
    S = spg.gen_stencil27(L, L, L, seed=3)
    n = L ** 3
    assert S.nnz == (3 * L - 2) ** 3
    A = (S.rows, S.cols, S.vals)
    (ptr, idx, val), nprod = run_restated(orc, A, (n, n), A, (n, n))
    assert nprod == (9 * L - 10) ** 3
    assert ptr[-1] == (5 * L - 6) ** 3
    assert np.diff(ptr).max() == min(5, L) ** 3
    ub, base, np2 = orc.upper_bound(n, n, *csc_of(A, (n, n))[:2], csr_of(A, (n, n))[0])
    assert np2 == nprod and ub.sum() == nprod


def test_uniform_rows_closed_form(orc, spg):
    U = spg.gen_uniform_rows(512, 512, 8, seed=1)
    A = (U.rows, U.cols, U.vals)
    (ptr, idx, val), nprod = run_restated(orc, A, (512, 512), A, (512, 512))
    assert nprod == 8 * U.nnz == 8 * 8 * 512


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bruteforce_dense(orc, seed):
    rng = np.random.default_rng(seed)
    m, K, n = 37, 53, 41
    Ad = (rng.random((m, K)) < 0.15) * rng.uniform(-1, 1, (m, K))
    Bd = (rng.random((K, n)) < 0.15) * rng.uniform(-1, 1, (K, n))
    Ad = Ad.astype(np.float32)
    Bd = Bd.astype(np.float32)
    A = np.nonzero(Ad)
    B = np.nonzero(Bd)
    Ac = (A[0].astype(np.int32), A[1].astype(np.int32), Ad[A])
    Bc = (B[0].astype(np.int32), B[1].astype(np.int32), Bd[B])
    (ptr, idx, val), nprod = run_restated(orc, Ac, (m, K), Bc, (K, n))
    S = (Ad != 0).astype(np.int64) @ (Bd != 0).astype(np.int64)
    V = Ad.astype(np.float64) @ Bd.astype(np.float64)
    assert nprod == S.sum()
    for i in range(m):
        js = np.nonzero(S[i])[0]
        assert np.array_equal(idx[ptr[i]:ptr[i + 1]], js)
        assert np.allclose(val[ptr[i]:ptr[i + 1]], V[i, js], rtol=1e-5, atol=1e-6)


def test_scipy_values(orc):
    rng = np.random.default_rng(7)
    A = sp.random(300, 200, density=0.03, random_state=rng, format="coo", dtype=np.float32)
    B = sp.random(200, 250, density=0.03, random_state=rng, format="coo", dtype=np.float32)
    Ac = (A.row.astype(np.int32), A.col.astype(np.int32), A.data)
    Bc = (B.row.astype(np.int32), B.col.astype(np.int32), B.data)
    (ptr, idx, val), _ = run_restated(orc, Ac, A.shape, Bc, B.shape)
    C = (A.astype(np.float64) @ B.astype(np.float64)).tocsr()
    C.sort_indices()
    # scipy drops nothing here (positive data: no cancellation), so its
    # structure must equal the pattern product
    assert np.array_equal(ptr, C.indptr) and np.array_equal(idx, C.indices)
    assert np.all(rel_err(val, C.data) < 1e-6)


def test_restated_equals_gustavson_prod32_bitwise(orc, spg):
    R = spg.gen_rmat(10, 10, 8 * 1024, seed=11)
    A = (R.rows, R.cols, R.vals)
    n = 1024
    (ptr, idx, val), _ = run_restated(orc, A, (n, n), A, (n, n))
    gp, gi, gv, _ = orc.gustavson(n, n, csr_of(A, (n, n)), csr_of(A, (n, n)), prod32=True)
    assert np.array_equal(ptr, gp) and np.array_equal(idx, gi)
    assert np.array_equal(val, gv.astype(np.float32))


def test_invariants_and_bins(orc, spg):
    R = spg.gen_rmat(11, 11, 16 * 2048, seed=5)
    n = 2048
    A = (R.rows, R.cols, R.vals)
    a_ptr, a_idx, a_val = csc_of(A, (n, n))
    b_ptr, b_idx, b_val = csr_of(A, (n, n))
    ub, base, nprod = orc.upper_bound(n, n, a_ptr, a_idx, b_ptr)
    alen, blen = np.diff(a_ptr).astype(np.int64), np.diff(b_ptr).astype(np.int64)
    assert nprod == int((alen * blen).sum()) == int(ub.sum())
    bins, perm, offs = orc.kbins(n, a_ptr, b_ptr, nprod)
    # bins partition every k with a non-empty B row, k ascending in a bin
    assert np.array_equal(np.sort(perm), np.nonzero(blen > 0)[0])
    for b in range(orc.NUM_KBINS):
        ks = perm[offs[b]:offs[b + 1]]
        assert np.all(bins[ks] == b) and np.all(np.diff(ks) > 0)
    heavy = alen * blen > 50 * (nprod // n)
    assert np.array_equal(bins == 0, heavy & (blen > 0))
    light = (~heavy) & (blen > 0)
    exp = np.where(blen >= 128, 8, np.floor(np.log2(np.maximum(blen, 1))).astype(np.int64) + 1)
    assert np.array_equal(bins[light], exp[light])
    icol, ival = orc.expand(n, n, (a_ptr, a_idx, a_val), (b_ptr, b_idx, b_val), base)
    cnt, tot = orc.symbolic_from_slots(n, n, base, icol)
    assert np.all(cnt <= ub) and np.all(cnt <= n)
    (ptr, idx, val), _ = orc.spgemm_restated(n, n, n, (a_ptr, a_idx, a_val), (b_ptr, b_idx, b_val))
    assert np.array_equal(np.diff(ptr), cnt)
    for i in range(0, n, 97):
        assert np.all(np.diff(idx[ptr[i]:ptr[i + 1]]) > 0)


def test_cpu_baseline_matches_restated(orc, spg):
    R = spg.gen_rmat(11, 11, 8 * 2048, seed=9)
    n = 2048
    A = (R.rows, R.cols, R.vals)
    Acsc, Bcsr = csc_of(A, (n, n)), csr_of(A, (n, n))
    (ptr, idx, val), nprod = orc.spgemm_restated(n, n, n, Acsc, Bcsr)
    for threads in (1, 3):
        r = orc.cpu_baseline(n, n, n, Acsc, Bcsr, 0, n, threads)
        assert r["nprod"] == nprod and r["nnz"] == ptr[-1]
        assert abs(r["checksum"] - float(val.astype(np.float64).sum())) <= 1e-6 * np.abs(val).sum()
    r = orc.cpu_baseline(n, n, n, Acsc, Bcsr, 100, 900, 2)
    assert r["nnz"] == ptr[900] - ptr[100]


@pytest.mark.parametrize("fault", [1, 2, 3, 4])
def test_faults_are_caught(orc, fault):
    """A planted oracle mistake must fail at least one golden fixture."""
    try:
        orc.set_fault(fault)
        caught = [not check_against_truth(orc, nm) for nm in ("tiny", "rect", "powerlaw")]
    finally:
        orc.set_fault(orc.FAULT_NONE)
    assert any(caught)
    assert all(check_against_truth(orc, nm) for nm in ("tiny", "rect", "powerlaw"))


# ---------------------------------------------------------------- predictor
def _mlp_setup(seed, N=40):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.3, mlp.NP)
    x = rng.normal(0, 1, (N, 4)) * [1, 2, 1.5, 0.5] + [1, 3, 2, 1]
    t = -np.abs(rng.normal(0, 1, N))
    return mlp, w, mlp.normalization(x), x, t


@pytest.mark.parametrize("seed", [0, 1])
def test_mlp_gradient_matches_finite_differences(seed):
    """Pins the backpropagation the training kernel is checked against: a
    dropped term, a transposed weight or a wrong sign moves some
    coordinate's central difference away from the analytic gradient."""
    mlp, w, norm, x, t = _mlp_setup(seed)
    _, g = mlp.loss_and_grad(w, norm, x, t)
    rng = np.random.default_rng(seed + 10)
    idx = np.concatenate([rng.choice(256, 12, replace=False),              # W1
                          256 + rng.choice(64, 6, replace=False),          # b1
                          320 + rng.choice(4096, 24, replace=False),       # W2
                          4416 + rng.choice(64, 6, replace=False),         # b2
                          4480 + rng.choice(64, 8, replace=False), [4544]])  # W3, b3
    h = 1e-6
    for q in idx:
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (mlp.loss_and_grad(wp, norm, x, t)[0] - mlp.loss_and_grad(wm, norm, x, t)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 + 1e-4 * abs(fd), (q, fd, g[q])


def test_mlp_forward_and_loss_agree(seed=3):
    mlp, w, norm, x, t = _mlp_setup(seed)
    y = mlp.forward(w, norm, x)
    loss, _ = mlp.loss_and_grad(w, norm, x, t)
    assert np.isclose(loss, np.mean((y - t) ** 2), rtol=1e-12)
    # prediction transform: y = target reproduces the count, clamped to
    # [0, min(ub, n)]; the uniform baseline is ub for ub << n and n for ub >> n
    ub = np.array([0, 10, 100, 1000])
    cnt = np.array([0, 7, 100, 30])
    n = 50
    tt = mlp.target(cnt, ub, n)
    base = mlp.log1p_uniform_count(ub, n)
    p = np.where(ub > 0, np.clip(np.expm1(tt + base), 0, np.minimum(ub, n)), 0)
    assert np.allclose(p, [0, 7, 50, 30])
    assert np.isclose(mlp.log1p_uniform_count(1, 1e9), np.log1p(1.0), rtol=1e-6)
    assert np.isclose(mlp.log1p_uniform_count(1e6, 50), np.log1p(50.0))


def test_mlp_features_match_row_loop():
    from oracle import mlp
    rng = np.random.default_rng(4)
    m, K, n = 30, 25, 40
    ar = rng.integers(0, m, 200)
    ac = rng.integers(0, K, 200)
    blen = rng.integers(0, 9, K)
    bptr = np.concatenate([[0], np.cumsum(blen)])
    x, ub = mlp.features(m, ar, ac, bptr)
    for i in range(m):
        bs = [blen[k] for r, k in zip(ar, ac) if r == i]
        if not bs:
            assert np.all(x[i] == 0) and ub[i] == 0
            continue
        exp = np.log1p([len(bs), sum(bs), max(bs), np.std(bs)])
        assert np.allclose(x[i], exp, rtol=1e-12, atol=1e-12)
        assert ub[i] == sum(bs)


# ------------------------------------------------------ COO helpers (a3, a5)
def _coo_with_ties(seed, nrows=60, ncols=45, nnz=3000):
    """Random COO with repeated (row, col) pairs, exact duplicate triplets,
    ties in value, and empty rows and columns (the last 5 of each)."""
    rng = np.random.default_rng(seed)
    r = rng.integers(0, nrows - 5, nnz).astype(np.int32)
    c = rng.integers(0, ncols - 5, nnz).astype(np.int32)
    v = rng.integers(-3, 4, nnz).astype(np.float32) * 0.5
    r[-40:], c[-40:], v[-40:] = r[:40], c[:40], v[:40]  # exact duplicates
    return r, c, v, nrows, ncols


def _lexsorted(r, c, v, by_row):
    major, minor = (r, c) if by_row else (c, r)
    o = np.lexsort((v, minor, major))
    return r[o], c[o], v[o]


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("by_row", [True, False], ids=["row_major", "col_major"])
def test_sort_coo_matches_lexsort(orc, seed, by_row):
    """src/coo.cpp:43-79: (major, minor, value) order, against numpy's
    lexsort; the sorted triplets are the input's as a multiset."""
    r, c, v, _, _ = _coo_with_ties(seed)
    t = orc.ORC_ROW_MAJOR if by_row else orc.ORC_COL_MAJOR
    got = orc.sort_coo(r, c, v, t)
    exp = _lexsorted(r, c, v, by_row)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


@pytest.mark.parametrize("by_row", [True, False], ids=["row_major", "col_major"])
def test_coo2cm_matches_bincount(orc, by_row):
    """src/mm.cu:43-63,100-108: ptr = exclusive scan of the major-index
    histogram (empty majors included), idx/val in the sorted COO's order."""
    r, c, v, nrows, ncols = _coo_with_ties(2)
    r, c, v = _lexsorted(r, c, v, by_row)
    nmajor = nrows if by_row else ncols
    major, minor = (r, c) if by_row else (c, r)
    t = orc.ORC_ROW_MAJOR if by_row else orc.ORC_COL_MAJOR
    ptr, idx, val = orc.coo2cm(r, c, v, t, nmajor)
    exp_ptr = np.concatenate([[0], np.cumsum(np.bincount(major, minlength=nmajor))])
    assert np.array_equal(ptr, exp_ptr)
    assert ptr[-1] == len(r) and np.all(np.diff(ptr)[-5:] == 0)
    assert np.array_equal(idx, minor) and np.array_equal(val, v)
    # (ptr, idx, val) read back by scipy is the triplets' matrix
    M = (sp.csr_matrix if by_row else sp.csc_matrix)(
        (val.astype(np.float64), idx, ptr), shape=(nrows, ncols))
    dense = np.zeros((nrows, ncols))
    np.add.at(dense, (r, c), v.astype(np.float64))
    assert np.array_equal(M.toarray(), dense)


def test_csc_to_csr_matches_scipy(orc):
    """CSC -> CSR with k ascending inside every row (duplicates kept), as
    scipy's tocsr() of a CSC matrix orders them."""
    r, c, v, m, K = _coo_with_ties(3)
    a_ptr, a_idx, a_val = csc_of((r, c, v), (m, K))
    got_ptr, got_idx, got_val = orc.csc_to_csr(m, K, a_ptr, a_idx, a_val)
    S = sp.csc_matrix((a_val, a_idx, a_ptr), shape=(m, K)).tocsr()
    assert np.array_equal(got_ptr, S.indptr) and np.array_equal(got_idx, S.indices)
    assert np.array_equal(got_val, S.data)
    assert np.all(np.diff(got_ptr)[-5:] == 0)
    # and as the test helper's independent lexsort build
    h_ptr, h_idx, _ = csr_of((r, c, v), (m, K))
    assert np.array_equal(got_ptr, h_ptr) and np.array_equal(got_idx, h_idx)


def _coo_pins_hold(orc):
    ok = True
    for by_row in (True, False):
        r, c, v, nrows, ncols = _coo_with_ties(4)
        t = orc.ORC_ROW_MAJOR if by_row else orc.ORC_COL_MAJOR
        got = orc.sort_coo(r, c, v, t)
        ok &= all(np.array_equal(g, e) for g, e in zip(got, _lexsorted(r, c, v, by_row)))
        rs, cs, vs = _lexsorted(r, c, v, by_row)
        nmajor = nrows if by_row else ncols
        ptr, idx, _ = orc.coo2cm(rs, cs, vs, t, nmajor)
        major = rs if by_row else cs
        ok &= np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(major, minlength=nmajor))]))
    a_ptr, a_idx, a_val = csc_of((r, c, v), (nrows, ncols))
    gp, gi, _ = orc.csc_to_csr(nrows, ncols, a_ptr, a_idx, a_val)
    S = sp.csc_matrix((a_val, a_idx, a_ptr), shape=(nrows, ncols)).tocsr()
    ok &= np.array_equal(gp, S.indptr) and np.array_equal(gi, S.indices)
    return bool(ok)


@pytest.mark.parametrize("fault", [5, 6, 7], ids=["sort_minor", "ptr_unshifted", "csr_row_idx"])
def test_coo_helper_faults_are_caught(orc, fault):
    """A planted mistake in sort_coo, coo2cm or csc_to_csr fails its pin."""
    assert _coo_pins_hold(orc)
    try:
        orc.set_fault(fault)
        caught = not _coo_pins_hold(orc)
    finally:
        orc.set_fault(orc.FAULT_NONE)
    assert caught
