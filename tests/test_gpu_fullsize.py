"""The benchmark workloads at their full sizes.  Sampled rows (the hub rows
with the largest bounds, rows whose tiles straddle tile boundaries, seeded
random rows) are compared with the oracle's row-wise product
(oracle.gustavson(rows=...)): structure bit-exact, values within VAL_RTOL.
The whole products are checked through properties that do not need the
oracle to run all of them (it would take minutes to hours):

* structure invariants: row pointers non-decreasing, columns in range and
  strictly increasing inside every row, nnz(C_i) <= min(ub_i, n), and
  sum(ub) = nprod = sum_k nnz(A col k) * nnz(B row k) (closed form on the
  host);
* closed forms where the workload has them: the 27-point stencil on 64^3 has
  nprod = 566^3 and nnz(C) = 314^3 (every product of two 27-point stencils
  reaches the 125-point neighbourhood clipped to the grid);
* linearity (a checksum of every row): C x = A (B x) for a random x, in
  float64 on the host, within VAL_RTOL of |A| (|B| |x|), the bound of the
  value contract (FP32 products, FP64 sums, one FP32 rounding);
* agreement of independent kernel paths on the same product: the learned
  path (different merge kernels, capacities from the predictor, fallback,
  compaction) gives the plain path's structure exactly and its values
  within VAL_RTOL; the batched path gives every product's one-product
  result.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-5


def csr_host(M, shape):
    """A generated HostCOO as a float64 scipy CSR (duplicates summed: the
    same linear map)."""
    return sp.csr_matrix((M.vals.astype(np.float64), (M.rows, M.cols)), shape=shape)


def check_structure(C, ub, n, nprod=None):
    ptr, col, _ = C
    m = len(ptr) - 1
    assert ptr[0] == 0 and ptr[-1] == len(col)
    cnt = np.diff(ptr)
    assert np.all(cnt >= 0)
    assert np.all(cnt <= np.minimum(ub, n))
    if nprod is not None:
        assert int(ub.sum()) == nprod
    if len(col):
        assert int(col.min()) >= 0 and int(col.max()) < n
        first = np.zeros(len(col), bool)
        first[ptr[:-1][cnt > 0]] = True  # the first entry of every non-empty row
        assert np.all((col[1:] > col[:-1]) | first[1:])
    return m


def check_linear(C, A, B, n, seed):
    ptr, col, val = C
    m = len(ptr) - 1
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    Cm = sp.csr_matrix((val.astype(np.float64), col, ptr), shape=(m, n))
    y1 = Cm @ x
    y2 = A @ (B @ x)
    bound = abs(A) @ (abs(B) @ np.abs(x))
    err = np.abs(y1 - y2)
    assert np.all(err <= VAL_RTOL * bound + 1e-30), float((err / np.maximum(bound, 1e-300)).max())


def same_c(C1, C2):
    assert np.array_equal(C1[0], C2[0])
    assert np.array_equal(C1[1], C2[1])
    a, b = C1[2].astype(np.float64), C2[2].astype(np.float64)
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    err[(np.abs(a) <= 1e-30) & (np.abs(b) <= 1e-30)] = 0
    assert err.max(initial=0.0) <= VAL_RTOL


def nprod_of(A, B):
    """sum_k nnz(A col k) * nnz(B row k), from the operands as the library
    holds them (duplicates kept)."""
    ka = np.bincount(A.cols, minlength=A.num_cols).astype(np.int64)
    kb = np.bincount(B.rows, minlength=B.num_rows).astype(np.int64)
    return int((ka * kb).sum())


def predictor(spg, ctx):
    from spgemm_amd import train_predictor as tp
    P, _ = tp.load_or_train(ctx)
    return P


def sample_rows(ub, long_row_ub, seed, n_top=50, n_edge=500, n_tiled=500, n_rand=1000):
    """Rows to check against the oracle: the n_top largest bounds (hub rows,
    split into the most products per tile), tiled rows with the smallest
    bounds above the tiling threshold, random tiled rows and random rows."""
    rng = np.random.default_rng(seed)
    order = np.argsort(ub, kind="stable")
    tiled = np.nonzero(ub > long_row_ub)[0]
    pick = [order[-n_top:]]
    if len(tiled):
        pick.append(tiled[np.argsort(ub[tiled], kind="stable")[:n_edge]])
        pick.append(rng.choice(tiled, min(n_tiled, len(tiled)), replace=False))
    pick.append(rng.choice(len(ub), min(n_rand, len(ub)), replace=False))
    return np.unique(np.concatenate(pick)).astype(np.int32)


def check_rows_vs_oracle(orc, C, A_csr, B_csr, m, n, rows):
    """Rows of C against the oracle's row-wise Gustavson with FP32 products
    (src/mm.cu:366-404's merge restated): structure bit-exact, values within
    VAL_RTOL of the FP64 sum rounded to FP32."""
    ptr, idx, val = C
    gp, gi, gv, _ = orc.gustavson(m, n, A_csr, B_csr, rows=rows, prod32=True)
    cnt = (ptr[rows + 1] - ptr[rows]).astype(np.int64)
    assert np.array_equal(cnt, np.diff(gp)), "row counts differ from the oracle"
    sel = np.concatenate([np.arange(ptr[i], ptr[i + 1]) for i in rows]) if len(rows) else []
    assert np.array_equal(idx[sel], gi), "columns differ from the oracle"
    got = val[sel].astype(np.float64)
    ref = gv.astype(np.float32).astype(np.float64)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    err[(np.abs(got) <= 1e-30) & (np.abs(ref) <= 1e-30)] = 0.0
    assert err.max(initial=0.0) <= VAL_RTOL, float(err.max())
    return int(cnt.sum())


def coo_csr(M):
    from helpers import csr_of
    return csr_of((M.rows, M.cols, M.vals), (M.num_rows, M.num_cols))


BENCH_CFG = dict(tile_log2=13, item_products=8192, long_row_ub=512, xcd_swizzle=1, timing=1,
                 heavy_factor=50, rank_min_count=1024)


@pytest.mark.parametrize("rank_min", [1024, 512, 128])
def test_rmat_s18_bench_config_vs_oracle(spg, orc, gpu_ctx, rank_min):
    """R-MAT scale 18, edge factor 16 with the settings the benchmark times
    (row-chunk expansion at 32 tiles per long row, long_row_ub 512, and
    each rank_min_count the tuner has chosen for the plain or the learned
    path): about 2k rows of the plain and of the learned C against the
    oracle, the hub rows (up to 2.5M products, 32 tiles) among them."""
    print(f"case full R-MAT s18 ef16, benchmark settings, rank_min_count {rank_min}, "
          "rows vs oracle", flush=True)
    ctx = gpu_ctx
    A = spg.gen_rmat(18, 18, 16 << 18, seed=0)
    ctx.set_config(expand_mode=spg.EXPAND_ROWS, **dict(BENCH_CFG, rank_min_count=rank_min))
    ctx.set_operands_from_coo(A, A)
    ctx.multiply()
    ub = ctx.upper_bound(A.num_rows)
    Cp = ctx.get_c()
    check_structure(Cp, ub, A.num_cols, nprod_of(A, A))
    Acsr = coo_csr(A)
    rows = sample_rows(ub, BENCH_CFG["long_row_ub"], 0)
    nchk = check_rows_vs_oracle(orc, Cp, Acsr, Acsr, A.num_rows, A.num_cols, rows)
    print(f"  plain: {len(rows)} rows, {nchk} entries match the oracle")
    P = predictor(spg, ctx)
    try:
        st = ctx.multiply_learned(P)
        Cl = ctx.get_c()
    finally:
        P.close()
    assert st["unrecovered"] == 0
    same_c(Cl, Cp)
    check_rows_vs_oracle(orc, Cl, Acsr, Acsr, A.num_rows, A.num_cols, rows)


def test_rect_learned_rows_vs_oracle(spg, orc, gpu_ctx):
    """The rectangular product through the learned path with the row-chunk
    expansion: sampled rows against the oracle."""
    print("case full rectangular R-MAT, learned, rows vs oracle", flush=True)
    ctx = gpu_ctx
    A = spg.gen_rmat(20, 18, 8 << 20, seed=4)
    B = spg.gen_rmat(18, 18, 8 << 18, seed=5)
    ctx.set_config(expand_mode=spg.EXPAND_ROWS, **BENCH_CFG)
    ctx.set_operands_from_coo(A, B)
    P = predictor(spg, ctx)
    try:
        st = ctx.multiply_learned(P)
        Cl = ctx.get_c()
    finally:
        P.close()
    assert st["unrecovered"] == 0
    ub = ctx.upper_bound(A.num_rows)
    check_structure(Cl, ub, B.num_cols, nprod_of(A, B))
    rows = sample_rows(ub, BENCH_CFG["long_row_ub"], 1)
    check_rows_vs_oracle(orc, Cl, coo_csr(A), coo_csr(B), A.num_rows, B.num_cols, rows)


def test_rmat_s18_full(spg, gpu_ctx):
    """R-MAT scale 18, edge factor 16: plain and learned paths."""
    print("case full R-MAT s18 ef16", flush=True)
    ctx = gpu_ctx
    A = spg.gen_rmat(18, 18, 16 << 18, seed=0)
    ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=512, xcd_swizzle=1, timing=1,
                   expand_mode=spg.EXPAND_KBINS, heavy_factor=50, rank_min_count=128)
    ctx.set_operands_from_coo(A, A)
    ctx.multiply()
    ub = ctx.upper_bound(A.num_rows)
    Cp = ctx.get_c()
    check_structure(Cp, ub, A.num_cols, nprod_of(A, A))
    Ah = csr_host(A, (A.num_rows, A.num_cols))
    check_linear(Cp, Ah, Ah, A.num_cols, 1)
    P = predictor(spg, ctx)
    try:
        ctx.set_config(long_row_ub=1024)
        st = ctx.multiply_learned(P)
        Cl = ctx.get_c()
    finally:
        P.close()
    print(f"  nnz {Cp[0][-1]}, learned stats {st}")
    assert st["unrecovered"] == 0
    same_c(Cl, Cp)


def test_stencil_64_full(spg, gpu_ctx):
    """27-point stencil on 64^3: closed forms and linearity, both paths."""
    print("case full stencil 64^3", flush=True)
    ctx = gpu_ctx
    S = spg.gen_stencil27(64, 64, 64, 3)
    ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=4096, xcd_swizzle=1, timing=1,
                   expand_mode=spg.EXPAND_ROWS, heavy_factor=50, rank_min_count=128)
    ctx.set_operands_from_coo(S, S)
    nprod = ctx.init_gemm()
    assert nprod == 566 ** 3
    ctx.multiply()
    Cp = ctx.get_c()
    assert Cp[0][-1] == 314 ** 3
    check_structure(Cp, ctx.upper_bound(S.num_rows), S.num_cols, 566 ** 3)
    Sh = csr_host(S, (S.num_rows, S.num_cols))
    check_linear(Cp, Sh, Sh, S.num_cols, 2)
    P = predictor(spg, ctx)
    try:
        st = ctx.multiply_learned(P)
        Cl = ctx.get_c()
    finally:
        P.close()
    assert st["unrecovered"] == 0
    same_c(Cl, Cp)


def test_rect_rmat_full(spg, gpu_ctx):
    """A = 2^20 x 2^18 R-MAT (8 * 2^20 draws), B = R-MAT scale 18, ef 8: the
    learned path (the workload's headline) checked by linearity and against
    the plain path."""
    print("case full rectangular R-MAT", flush=True)
    ctx = gpu_ctx
    A = spg.gen_rmat(20, 18, 8 << 20, seed=4)
    B = spg.gen_rmat(18, 18, 8 << 18, seed=5)
    ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=1024, xcd_swizzle=1, timing=1,
                   expand_mode=spg.EXPAND_KBINS, heavy_factor=50, rank_min_count=128)
    ctx.set_operands_from_coo(A, B)
    P = predictor(spg, ctx)
    try:
        st = ctx.multiply_learned(P)
        Cl = ctx.get_c()
    finally:
        P.close()
    assert st["unrecovered"] == 0
    check_structure(Cl, ctx.upper_bound(A.num_rows), B.num_cols, nprod_of(A, B))
    check_linear(Cl, csr_host(A, (A.num_rows, A.num_cols)), csr_host(B, (B.num_rows, B.num_cols)),
                 B.num_cols, 3)
    ctx.multiply()
    same_c(ctx.get_c(), Cl)


def test_batch256_graph_products_vs_oracle(spg, orc, gpu_ctx):
    """Products of the 256-product batch from a learned graph replay with
    the row-chunk expansion, whole, against the oracle's restated method:
    the three smallest, three mid-sized and the largest below 2^24
    products."""
    print("case full batch256, graph replay products vs oracle", flush=True)
    import bench
    from helpers import csc_of, csr_of
    ctx = gpu_ctx
    pairs = bench.make_batch(256, 5)
    ctx.set_config(expand_mode=spg.EXPAND_ROWS, **BENCH_CFG)
    ctx.batch_set_operands(pairs)
    np_b = np.array([nprod_of(A, B) for A, B in pairs])
    order = np.argsort(np_b, kind="stable")
    small = [int(b) for b in order if np_b[b] < (1 << 24)]
    pick = sorted(set(small[:3] + small[len(small) // 2 - 1:len(small) // 2 + 2] + small[-1:]))
    P = predictor(spg, ctx)
    try:
        ctx.batch_multiply(predictor=P, graph=True)  # records and captures
        st = ctx.batch_multiply(predictor=P, graph=True)  # replays
        ctx.synchronize()
    finally:
        P.close()
    assert st["unrecovered"] == 0
    for b in pick:
        A, B = pairs[b]
        m, K, n = A.num_rows, A.num_cols, B.num_cols
        Cg = ctx.batch_get_c(b)
        ref, nprod = orc.spgemm_restated(m, n, K, csc_of((A.rows, A.cols, A.vals), (m, K)),
                                         csr_of((B.rows, B.cols, B.vals), (K, n)))
        assert nprod == np_b[b]
        assert np.array_equal(Cg[0], ref[0]) and np.array_equal(Cg[1], ref[1]), f"product {b}"
        err = np.abs(Cg[2].astype(np.float64) - ref[2]) / np.maximum(np.abs(ref[2]), 1e-30)
        err[(np.abs(Cg[2]) <= 1e-30) & (np.abs(ref[2]) <= 1e-30)] = 0.0
        assert err.max(initial=0.0) <= VAL_RTOL, (b, float(err.max()))
    print(f"  products {pick} match the oracle")


def test_batch256_full(spg, gpu_ctx):
    """The 256-product batch: learned + batched (graph) against plain
    batched, every product checked by structure and linearity."""
    print("case full batch256", flush=True)
    import bench
    ctx = gpu_ctx
    pairs = bench.make_batch(256, 5)
    ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=1024, xcd_swizzle=1, timing=1,
                   expand_mode=spg.EXPAND_KBINS, heavy_factor=50, rank_min_count=128)
    ctx.batch_set_operands(pairs)
    ctx.batch_multiply()
    plain = [ctx.batch_get_c(b) for b in range(len(pairs))]
    P = predictor(spg, ctx)
    try:
        st = ctx.batch_multiply(predictor=P, graph=True)
        ctx.synchronize()
    finally:
        P.close()
    assert st["unrecovered"] == 0
    for b, (A, B) in enumerate(pairs):
        Cl = ctx.batch_get_c(b)
        same_c(Cl, plain[b])
        # upper bound of every row with duplicates kept, as the library counts
        blen = np.bincount(B.rows, minlength=B.num_rows)
        ub = np.bincount(A.rows, weights=blen[A.cols], minlength=A.num_rows).astype(np.int64)
        check_structure(Cl, ub, B.num_cols, nprod_of(A, B))
        check_linear(Cl, csr_host(A, (A.num_rows, A.num_cols)),
                     csr_host(B, (B.num_rows, B.num_cols)), B.num_cols, 10 + b)
