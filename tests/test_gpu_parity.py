"""GPU parity of every phase against the CPU oracle, through the C-ABI.

Bar: upper bounds, k-bins, symbolic counts and C's structure (row_ptr,
col_idx) bit-exact; the intermediate products equal as a multiset per C row
(their order inside a row is not part of the contract); C values within
1e-5 relative of the oracle's restated method (FP32 products, FP64 sums,
one rounding; tolerance VAL_RTOL below), error against the FP64 oracle
reported.
"""
import numpy as np
import pytest

from helpers import PROD_FIXTURES, csc_of, csr_of, load_prod, rel_err

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-5      # value tolerance: 1e-5 relative to the FP32 restated oracle
ZERO_ATOL = 1e-30    # exact cancellations: both sides must be ~0


def run_gpu(spg, ctx, A, Ash, B, Bsh, simple=False, via_coo=True, cfg=None):
    base_cfg = dict(tile_log2=13, item_products=4096, long_row_ub=4096, xcd_swizzle=1, timing=1,
                    expand_mode=spg.EXPAND_ROWS)
    base_cfg.update(cfg or {})
    ctx.set_config(**base_cfg)
    m, K = Ash
    n = Bsh[1]
    if via_coo:
        ctx.coo2cm(spg.HostCOO(*A, m, K).sorted(spg.COL_MAJOR))
        ctx.coo2cm(spg.HostCOO(*B, K, n).sorted(spg.ROW_MAJOR))
    else:
        ctx.set_operand(spg.OPERAND_A, *csc_of(A, Ash), m, K)
        ctx.set_operand(spg.OPERAND_B, *csr_of(B, Bsh), K, n)
    out = {}
    out["nprod"] = ctx.init_gemm()
    out["ub"] = ctx.upper_bound(m)
    if simple:
        ctx.simple_gemm()
    else:
        out["bin_counts"] = ctx.categorize_gemm()
        out["kbins"] = ctx.kbins(K)
        ctx.bin_gemm()
    out["inter"] = ctx.intermediate_by_row(m, out["nprod"])
    out["nnz"] = ctx.symbolic_gemm()
    out["cnt"] = ctx.row_counts(m)
    ctx.merge_gemm()
    out["C"] = ctx.get_c()
    return out


def oracle_all(orc, A, Ash, B, Bsh):
    m, K = Ash
    n = Bsh[1]
    Acsc, Bcsr = csc_of(A, Ash), csr_of(B, Bsh)
    ub, base, nprod = orc.upper_bound(m, K, Acsc[0], Acsc[1], Bcsr[0])
    bins = orc.kbins(K, Acsc[0], Bcsr[0], nprod)
    icol, ival = orc.expand(m, K, Acsc, Bcsr, base)
    cnt, _ = orc.symbolic_from_slots(m, n, base, icol)
    C, _ = orc.spgemm_restated(m, n, K, Acsc, Bcsr)
    G = orc.gustavson(m, n, csr_of(A, Ash), Bcsr)
    return dict(ub=ub, base=base, nprod=nprod, kbins=bins, inter=(icol, ival), cnt=cnt, C=C, G=G)


def row_multiset(rowptr, cols, vals):
    m = len(rowptr) - 1
    rows = np.repeat(np.arange(m), np.diff(rowptr))
    vb = vals.view(np.int32)
    o = np.lexsort((vb, cols, rows))
    return cols[o], vb[o]


def check_values(Cg, Co, G):
    _, _, gv = Cg
    _, _, ov = Co
    gv64 = gv.astype(np.float64)
    ov64 = ov.astype(np.float64)
    both_zero = (np.abs(gv64) <= ZERO_ATOL) & (np.abs(ov64) <= ZERO_ATOL)
    err = rel_err(gv64, ov64)
    err[both_zero] = 0.0
    assert err.max(initial=0.0) <= VAL_RTOL, f"max rel err {err.max()}"
    # report-only: exact matches and error vs FP64
    _, _, v64, sab = G
    e64 = np.abs(gv64 - v64) / np.maximum(sab, 1e-300)
    return float(np.mean(gv == ov)) if len(gv) else 1.0, float(e64.max(initial=0.0))


def compare(out, ref, check_bins=True):
    assert out["nprod"] == ref["nprod"]
    assert np.array_equal(out["ub"], ref["ub"])
    if check_bins:
        gb, gp, go = out["kbins"]
        ob, op, oo = ref["kbins"]
        assert np.array_equal(gb, ob) and np.array_equal(go, oo) and np.array_equal(gp, op)
    rp, ic, iv = out["inter"]
    assert np.array_equal(rp, ref["base"])
    gc, gvb = row_multiset(rp, ic, iv)
    oc, ovb = row_multiset(ref["base"], *ref["inter"])
    assert np.array_equal(gc, oc) and np.array_equal(gvb, ovb)
    assert np.array_equal(out["cnt"], ref["cnt"])
    gptr, gidx, _ = out["C"]
    optr, oidx, _ = ref["C"]
    assert out["nnz"] == optr[-1]
    assert np.array_equal(gptr, optr) and np.array_equal(gidx, oidx)
    return check_values(out["C"], ref["C"], ref["G"])


@pytest.mark.parametrize("name", PROD_FIXTURES)
def test_golden_fixtures(spg, orc, gpu_ctx, name):
    A, Ash, B, Bsh, z = load_prod(name)
    print(f"case golden {name}", flush=True)
    out = run_gpu(spg, gpu_ctx, A, Ash, B, Bsh)
    compare(out, oracle_all(orc, A, Ash, B, Bsh))
    assert np.array_equal(out["C"][0], z["C_ptr"]) and np.array_equal(out["C"][1], z["C_idx"])


def gen(spg, kind, **kw):
    if kind == "uniform":
        U = spg.gen_uniform_rows(kw["n"], kw["n"], kw["k"], kw["seed"])
    elif kind == "stencil":
        U = spg.gen_stencil27(kw["L"], kw["L"], kw["L"], kw["seed"])
    elif kind == "rmat":
        U = spg.gen_rmat(kw["s"], kw.get("cs", kw["s"]), kw["ef"] << kw["s"], seed=kw["seed"])
    return (U.rows, U.cols, U.vals), (U.num_rows, U.num_cols)


CASES = [
    ("cfg1_uniform_4096x8", "uniform", dict(n=4096, k=8, seed=1)),
    ("stencil_16", "stencil", dict(L=16, seed=3)),
    ("rmat_s12_ef16", "rmat", dict(s=12, ef=16, seed=0)),
    ("rmat_s14_ef8", "rmat", dict(s=14, ef=8, seed=4)),
]


@pytest.mark.parametrize("mode", [0, 1], ids=["kbins", "rows"])
@pytest.mark.parametrize("label,kind,kw", CASES, ids=[c[0] for c in CASES])
def test_square_products(spg, orc, gpu_ctx, label, kind, kw, mode):
    print(f"case {label} mode={mode}", flush=True)
    A, sh = gen(spg, kind, **kw)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh, cfg=dict(expand_mode=mode))
    match, e64 = compare(out, oracle_all(orc, A, sh, A, sh))
    print(f"  exact-match fraction {match:.6f}, max |c-c64|/sum|ab| {e64:.3e}")


def test_rectangular(spg, orc, gpu_ctx):
    print("case rectangular rmat", flush=True)
    R1 = spg.gen_rmat(13, 11, 8 << 13, seed=21)
    R2 = spg.gen_rmat(11, 12, 6 << 11, seed=22)
    A, Ash = (R1.rows, R1.cols, R1.vals), (R1.num_rows, R1.num_cols)
    B, Bsh = (R2.rows, R2.cols, R2.vals), (R2.num_rows, R2.num_cols)
    compare(run_gpu(spg, gpu_ctx, A, Ash, B, Bsh, via_coo=False), oracle_all(orc, A, Ash, B, Bsh))


@pytest.mark.parametrize("mode", [0, 1], ids=["kbins", "rows"])
@pytest.mark.parametrize("tile_log2,long_ub", [(6, 16), (8, 64), (5, 1)])
def test_long_row_tiling(spg, orc, gpu_ctx, tile_log2, long_ub, mode):
    """Small tiles force the column-tiled path (bsplit, per-tile slots,
    bitmap symbolic pass, dense merge) on a size the oracle checks, for both
    expansion mappings."""
    print(f"case tiling tile_log2={tile_log2} long_ub={long_ub} mode={mode}", flush=True)
    A, sh = gen(spg, "rmat", s=10, ef=16, seed=7)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh,
                  cfg=dict(tile_log2=tile_log2, long_row_ub=long_ub, expand_mode=mode))
    compare(out, oracle_all(orc, A, sh, A, sh))


@pytest.mark.parametrize("rc_ub", [1 << 12, 1 << 30], ids=["rc_light", "rc_all"])
@pytest.mark.parametrize("tile_log2,long_ub,s", [(6, 16, 10), (5, 1, 13)])
def test_rowchunk_split(spg, orc, gpu_ctx, tile_log2, long_ub, s, rc_ub):
    """The binned expansion with the tiled rows below a bound expanded in
    row chunks (slots reserved per chunk and tile, runs written front to
    back) and the rest by k: same intermediates and C as the oracle, with
    16 (entries per chunk: 64) and 256 tiles (the largest row-chunk row:
    8 entries per chunk)."""
    print(f"case rowchunk tile_log2={tile_log2} long_ub={long_ub} s={s} rc_ub={rc_ub}", flush=True)
    A, sh = gen(spg, "rmat", s=s, ef=8, seed=9)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh,
                  cfg=dict(tile_log2=tile_log2, long_row_ub=long_ub, expand_mode=spg.EXPAND_KBINS,
                           rowchunk_ub=rc_ub))
    compare(out, oracle_all(orc, A, sh, A, sh))


def test_rows_mode_many_tiles(spg, orc, gpu_ctx):
    """ROWS mode with 512 tiles per long row (more than the row-chunk kernel
    takes): those rows go through the per-virtual-row kernels, the untiled
    rows through the entry-chunk kernel."""
    print("case rows mode, 512 tiles", flush=True)
    A, sh = gen(spg, "rmat", s=14, ef=4, seed=13)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh,
                  cfg=dict(tile_log2=5, long_row_ub=16, expand_mode=spg.EXPAND_ROWS))
    compare(out, oracle_all(orc, A, sh, A, sh))


def test_transpose_staged_buckets(spg, orc, gpu_ctx):
    """The counting-sort transpose with buckets of thousands of entries
    (here one row per bucket: 256 rows of 4200 or 20000 entries), which are
    placed in LDS and written front to back, or placed directly when they
    exceed the staging area (the 20000-entry rows)."""
    print("case transpose staged buckets", flush=True)
    rng = np.random.default_rng(77)
    m, K, n = 256, 1 << 16, 1000
    rows, cols = [], []
    for i in range(m):
        r = 20000 if i % 37 == 5 else 4200
        rows.append(np.full(r, i, np.int32))
        cols.append(rng.choice(K, r, replace=False).astype(np.int32))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    vals = rng.uniform(-1, 1, len(rows)).astype(np.float32)
    rb = np.arange(K, dtype=np.int32)
    cb = rng.integers(0, n, K).astype(np.int32)
    vb = rng.uniform(-1, 1, K).astype(np.float32)
    out = run_gpu(spg, gpu_ctx, (rows, cols, vals), (m, K), (rb, cb, vb), (K, n), via_coo=False,
                  cfg=dict(transpose_mode=spg.TRANSPOSE_SORT, expand_mode=spg.EXPAND_ROWS))
    compare(out, oracle_all(orc, (rows, cols, vals), (m, K), (rb, cb, vb), (K, n)))


@pytest.mark.parametrize("tmode", [2, 1], ids=["sort", "atomic"])
def test_transpose_modes(spg, orc, gpu_ctx, tmode):
    """InitGEMM's A-by-rows transpose: the LDS counting sort and the
    one-atomic-per-entry transpose give the same bounds, intermediates
    and C, on a power-law product (hub rows in one bucket) and on rows with
    empty buckets."""
    print(f"case transpose mode={tmode}", flush=True)
    A, sh = gen(spg, "rmat", s=13, ef=16, seed=31)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh,
                  cfg=dict(transpose_mode=tmode, expand_mode=spg.EXPAND_ROWS))
    compare(out, oracle_all(orc, A, sh, A, sh))
    # rows far apart: most row buckets empty
    m, K = 70000, 300
    rng = np.random.default_rng(5)
    rows = np.sort(rng.choice(m, 900, replace=False)).astype(np.int32)
    cols = rng.integers(0, K, 900).astype(np.int32)
    vals = rng.uniform(-1, 1, 900).astype(np.float32)
    Bm = spg.gen_uniform_rows(K, 500, 6, 32)
    B, Bsh = (Bm.rows, Bm.cols, Bm.vals), (Bm.num_rows, Bm.num_cols)
    out = run_gpu(spg, gpu_ctx, (rows, cols, vals), (m, K), B, Bsh,
                  cfg=dict(transpose_mode=tmode, expand_mode=spg.EXPAND_ROWS))
    compare(out, oracle_all(orc, (rows, cols, vals), (m, K), B, Bsh))


def test_simple_equals_binned(spg, orc, gpu_ctx):
    print("case simple vs binned", flush=True)
    A, sh = gen(spg, "rmat", s=12, ef=8, seed=3)
    ref = oracle_all(orc, A, sh, A, sh)
    o1 = run_gpu(spg, gpu_ctx, A, sh, A, sh, simple=True)
    compare(o1, ref, check_bins=False)
    o2 = run_gpu(spg, gpu_ctx, A, sh, A, sh, simple=False)
    compare(o2, ref)
    assert np.array_equal(o1["C"][2], o2["C"][2])


@pytest.mark.parametrize("item_products,swz", [(64, 0), (257, 1), (100000, 1)])
def test_work_item_shapes(spg, orc, gpu_ctx, item_products, swz):
    """Small work items split long B rows across items (2D split of heavy k)."""
    print(f"case items P={item_products} swizzle={swz}", flush=True)
    A, sh = gen(spg, "rmat", s=11, ef=16, seed=8)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh,
                  cfg=dict(item_products=item_products, xcd_swizzle=swz, expand_mode=spg.EXPAND_KBINS))
    compare(out, oracle_all(orc, A, sh, A, sh))


def test_edge_structures(spg, orc, gpu_ctx):
    """Empty rows/cols (b_len = 0), b_len >= 128 (never binned by the
    reference, D6), a heavy k, one dense row, duplicate entries."""
    print("case edge structures", flush=True)
    n = 600
    rng = np.random.default_rng(1)
    rows, cols = [], []
    for j in range(590):        # dense column 3 (rows 590.. stay empty)
        rows.append(j)
        cols.append(3)
    for j in range(n):          # dense row 5: k = 5 has b_len 600 (bin 8)
        rows.append(5)
        cols.append(j)
    for j in range(0, n, 4):    # row 11: b_len 150 (bin 8)
        rows.append(11)
        cols.append(j)
    rows += [100, 100, 100]     # duplicate entries of (100, 200)
    cols += [200, 200, 201]
    for i in range(20, 560):    # background, 12 per row
        for c in rng.choice(n, 12, replace=False):
            rows.append(i)
            cols.append(int(c))
    for c in rng.choice(n, 60, replace=False):  # row 3: k = 3 becomes heavy
        rows.append(3)
        cols.append(int(c))
    rows = np.array(rows, np.int32)
    cols = np.array(cols, np.int32)
    vals = (rng.integers(-64, 64, len(rows)) / 64.0).astype(np.float32)
    A = (rows, cols, vals)
    sh = (n, n)
    out = run_gpu(spg, gpu_ctx, A, sh, A, sh, via_coo=False)
    ref = oracle_all(orc, A, sh, A, sh)
    compare(out, ref)
    obins = ref["kbins"][0]
    assert obins[3] == 0 and obins[5] == 8 and obins[11] == 8 and (obins == -1).any()


def test_tiny_shapes(spg, orc, gpu_ctx):
    print("case tiny shapes", flush=True)
    for (m, K, n, e) in [(1, 1, 1, 1), (1, 5, 1, 3), (3, 1, 4, 2), (2, 2, 2, 0)]:
        rng = np.random.default_rng(m * 100 + K * 10 + n)
        ra = rng.integers(0, m, e).astype(np.int32)
        ca = rng.integers(0, K, e).astype(np.int32)
        rb = rng.integers(0, K, e).astype(np.int32)
        cb = rng.integers(0, n, e).astype(np.int32)
        va = np.ones(e, np.float32)
        A, B = (ra, ca, va), (rb, cb, va)
        compare(run_gpu(spg, gpu_ctx, A, (m, K), B, (K, n)), oracle_all(orc, A, (m, K), B, (K, n)))


def test_repeat_is_deterministic(spg, gpu_ctx):
    print("case determinism", flush=True)
    A, sh = gen(spg, "rmat", s=13, ef=16, seed=2)
    o1 = run_gpu(spg, gpu_ctx, A, sh, A, sh)
    o2 = run_gpu(spg, gpu_ctx, A, sh, A, sh)
    for a, b in zip(o1["C"], o2["C"]):
        assert np.array_equal(a, b)


def test_reference_call_order_without_symbolic(spg, orc, gpu_ctx):
    """src/main.cu order: COO2CM x2 -> Init -> Categorize -> Bin -> Merge."""
    print("case reference call order", flush=True)
    A, sh = gen(spg, "stencil", L=10, seed=3)
    ctx = gpu_ctx
    ctx.set_config(tile_log2=13, item_products=4096, long_row_ub=4096, xcd_swizzle=1, timing=1,
                   expand_mode=spg.EXPAND_ROWS)
    ptrA = ctx.coo2cm(spg.HostCOO(*A, *sh).sorted(spg.COL_MAJOR), return_host=True)
    ptrB = ctx.coo2cm(spg.HostCOO(*A, *sh).sorted(spg.ROW_MAJOR), return_host=True)
    Acsc, Bcsr = csc_of(A, sh), csr_of(A, sh)
    for g, o in zip(ptrA + ptrB, Acsc + Bcsr):
        assert np.array_equal(g, o)
    ctx.init_gemm()
    ctx.categorize_gemm()
    ctx.bin_gemm()
    ctx.merge_gemm()
    C = ctx.get_c()
    ref, _ = orc.spgemm_restated(sh[0], sh[1], sh[0], Acsc, Bcsr)
    assert np.array_equal(C[0], ref[0]) and np.array_equal(C[1], ref[1])
    assert np.all(rel_err(C[2], ref[2]) <= VAL_RTOL)


def test_phase_errors(spg, gpu_ctx):
    ctx = spg.Context(0)
    try:
        with pytest.raises(spg.SpgError):
            ctx.bin_gemm()
        with pytest.raises(spg.SpgError):
            ctx.init_gemm()
        with pytest.raises(spg.SpgError):
            ctx.get_c()
        with pytest.raises(spg.SpgError):
            ctx.set_config(tile_log2=20)
    finally:
        ctx.close()


def test_wide_b_hash_path(spg, orc, gpu_ctx):
    """n > 2^18 columns: rows over the full column range with more than 128
    distinct columns take the LDS hash + sort path instead of the bitmap
    rank path."""
    print("case wide B", flush=True)
    rng = np.random.default_rng(5)
    m, K, n = 1500, 1200, 300000
    ra = np.repeat(np.arange(m), 24).astype(np.int32)
    ca = rng.integers(0, K, len(ra)).astype(np.int32)
    rb = np.repeat(np.arange(K), 110).astype(np.int32)
    cb = rng.integers(0, n, len(rb)).astype(np.int32)
    A = (ra, ca, rng.uniform(-1, 1, len(ra)).astype(np.float32))
    B = (rb, cb, rng.uniform(-1, 1, len(rb)).astype(np.float32))
    compare(run_gpu(spg, gpu_ctx, A, (m, K), B, (K, n), via_coo=False),
            oracle_all(orc, A, (m, K), B, (K, n)))


@pytest.mark.parametrize("rank_min", [128, 1024, 4096])
def test_wide_row_merge_classes(spg, orc, gpu_ctx, rank_min):
    """Untiled rows over a 2^17-column window (wider than a tile) with about
    200 to 4000 distinct columns.  rank_min_count 128 merges them all through
    the bitmap rank kernel; 1024 sends those up to 1024 columns to the LDS
    hashes (512 to 2048 slots) and the rest to rank; 4096 sends every one to
    a hash (up to 8192 slots)."""
    print(f"case wide-row merge classes rank_min_count={rank_min}", flush=True)
    rng = np.random.default_rng(11)
    K, n, per_b = 4000, 1 << 17, 100
    sizes = [2, 3, 5, 8, 12, 20, 30, 40]  # entries per row: 200 .. 4000 products
    m = 8 * 40
    ra, ca = [], []
    for i in range(m):
        r = sizes[i % len(sizes)]
        ra += [i] * r
        ca += list(rng.choice(K, r, replace=False))
    ra, ca = np.array(ra, np.int32), np.array(ca, np.int32)
    rb = np.repeat(np.arange(K), per_b).astype(np.int32)
    cb = np.concatenate([rng.choice(n, per_b, replace=False) for _ in range(K)]).astype(np.int32)
    A = (ra, ca, rng.uniform(-1, 1, len(ra)).astype(np.float32))
    B = (rb, cb, rng.uniform(-1, 1, len(rb)).astype(np.float32))
    out = run_gpu(spg, gpu_ctx, A, (m, K), B, (K, n), via_coo=False,
                  cfg=dict(long_row_ub=4096, rank_min_count=rank_min))
    assert int(np.diff(out["C"][0]).max()) > 3000 and int(np.diff(out["C"][0]).min()) < 250
    compare(out, oracle_all(orc, A, (m, K), B, (K, n)))


@pytest.mark.parametrize("n", [8200, 40001, 65536])
def test_wave_rank_windows(spg, orc, gpu_ctx, n):
    """Untiled rows of a window wider than a tile and at most 2^16 columns,
    129 to 1024 distinct columns: the one-wave bitmap rank merge (and the
    one-wave bitmap symbolic count).  The window widths give bitmaps of 257
    words (one word past a quad), 1251 and 2048 words, so a lane's share
    of words is partly or wholly past the window."""
    print(f"case wave rank windows n={n}", flush=True)
    rng = np.random.default_rng(n)
    K, per_b = 3000, 64
    sizes = [3, 5, 8, 12, 16]  # entries per row: 192 .. 1024 products
    m = 5 * 60
    ra, ca = [], []
    for i in range(m):
        r = sizes[i % len(sizes)]
        ra += [i] * r
        ca += list(rng.choice(K, r, replace=False))
    ra, ca = np.array(ra, np.int32), np.array(ca, np.int32)
    rb = np.repeat(np.arange(K), per_b).astype(np.int32)
    cb = np.concatenate([rng.choice(n, per_b, replace=False) for _ in range(K)]).astype(np.int32)
    A = (ra, ca, rng.uniform(-1, 1, len(ra)).astype(np.float32))
    B = (rb, cb, rng.uniform(-1, 1, len(rb)).astype(np.float32))
    out = run_gpu(spg, gpu_ctx, A, (m, K), B, (K, n), via_coo=False,
                  cfg=dict(long_row_ub=4096, rank_min_count=128))
    cnt = np.diff(out["C"][0])
    assert cnt.max() > 700 and cnt.min() > 128 and cnt.max() <= 1024
    compare(out, oracle_all(orc, A, (m, K), B, (K, n)))
