"""Learned allocation on the GPU: predictor inference and training kernels
against the float64 restatement in oracle/mlp.py, and the learned merge
(predicted capacities, overflow detection, exact fallback, compaction)
against the plain exact-count path and the CPU oracle.

Bar: C's structure identical to the plain path's and the oracle's; values
within VAL_RTOL of the plain path (same FP64 accumulation, one FP32
rounding); zero unrecovered overflows.  Inference within PRED_RTOL of the
float64 forward pass (FP32 features and arithmetic on the GPU); the
gradient within GRAD_RTOL (relative to its norm) of float64
backpropagation.
"""
import numpy as np
import pytest

from helpers import csc_of, csr_of, rel_err

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-5
PRED_RTOL = 2e-3
GRAD_RTOL = 2e-3

BASE_CFG = dict(tile_log2=13, item_products=4096, long_row_ub=4096, xcd_swizzle=1, timing=1,
                expand_mode=0, pred_margin=0.25, pred_slack=8, pred_tight=0)


def gen(spg, kind, seed, **kw):
    if kind == "uniform":
        U = spg.gen_uniform_rows(kw["n"], kw["n"], kw["k"], seed)
    elif kind == "stencil":
        U = spg.gen_stencil27(kw["L"], kw["L"], kw["L"], seed)
    else:
        U = spg.gen_rmat(kw["s"], kw["s"], kw["ef"] << kw["s"], seed=seed)
    return (U.rows, U.cols, U.vals), (U.num_rows, U.num_cols)


def load(spg, ctx, A, Ash, B, Bsh, cfg=None):
    c = dict(BASE_CFG)
    c.update(cfg or {})
    ctx.set_config(**c)
    ctx.set_operand(spg.OPERAND_A, *csc_of(A, Ash), Ash[0], Ash[1])
    ctx.set_operand(spg.OPERAND_B, *csr_of(B, Bsh), Bsh[0], Bsh[1])


def plain(spg, ctx, A, Ash, B, Bsh, cfg=None):
    load(spg, ctx, A, Ash, B, Bsh, cfg)
    ctx.init_gemm()
    ctx.categorize_gemm()
    ctx.bin_gemm()
    ctx.symbolic_gemm()
    ctx.merge_gemm()
    return ctx.get_c()


def scramble(spg, ctx):
    """Fills the context's C buffers with another product's output, so a
    position the learned path fails to write cannot pass by holding the
    plain run's value."""
    U = spg.gen_uniform_rows(20000, 20000, 24, 12345)
    ctx.set_config(**BASE_CFG)
    ctx.set_operands_from_coo(U, U)
    ctx.multiply()


def learned(spg, ctx, A, Ash, B, Bsh, pred=None, predictor=None, cfg=None):
    scramble(spg, ctx)
    load(spg, ctx, A, Ash, B, Bsh, cfg)
    ctx.init_gemm()
    ctx.categorize_gemm()
    ctx.bin_gemm()
    if predictor is not None:
        ctx.predict_rows(predictor, Ash[0])
    else:
        ctx.set_row_predictions(pred)
    st = ctx.learned_merge_gemm()
    return ctx.get_c(), st


def same_c(Cl, Cp):
    assert np.array_equal(Cl[0], Cp[0])
    assert np.array_equal(Cl[1], Cp[1])
    err = rel_err(Cl[2].astype(np.float64), Cp[2].astype(np.float64))
    both_zero = (np.abs(Cl[2]) <= 1e-30) & (np.abs(Cp[2]) <= 1e-30)
    err[both_zero] = 0
    assert err.max(initial=0.0) <= VAL_RTOL


def host_features(A, Ash, B, Bsh):
    from oracle import mlp
    Bptr = csr_of(B, Bsh)[0]
    return mlp.features(Ash[0], A[0], A[1], Bptr)


CASES = [
    ("uniform_4096x8", "uniform", 1, dict(n=4096, k=8)),
    ("stencil_12", "stencil", 3, dict(L=12)),
    ("rmat_s12_ef16", "rmat", 0, dict(s=12, ef=16)),
    ("rmat_s13_ef8", "rmat", 4, dict(s=13, ef=8)),
]


def test_inference_matches_numpy(spg, gpu_ctx):
    from oracle import mlp
    print("case learned inference", flush=True)
    A, sh = gen(spg, "rmat", 2, s=12, ef=16)
    P = spg.Predictor(0, seed=11)
    try:
        w, norm = P.get_weights()
        norm = np.array([1.0, 3.0, 2.0, 1.0, 0.7, 1.9, 1.2, 0.8], np.float32)
        P.set_weights(w, norm)
        load(spg, gpu_ctx, A, sh, A, sh)
        gpu_ctx.init_gemm()
        pred = gpu_ctx.predict_rows(P, sh[0])
        x, ub = host_features(A, sh, A, sh)
        assert np.array_equal(ub, gpu_ctx.upper_bound(sh[0]))
        ref = mlp.predict(w, norm.astype(np.float64), x, ub, sh[1])
        err = np.abs(pred - ref) / np.maximum(np.abs(ref), 1.0)
        print(f"  max rel err {err.max():.2e}; rows {sh[0]}")
        assert err.max() <= PRED_RTOL
        assert np.all(pred[ub == 0] == 0)
        assert np.all(pred >= 0) and np.all(pred <= np.minimum(ub, sh[1]))
    finally:
        P.close()


def _samples(spg, ctx, P, cases):
    """Runs the plain path on each case, adds its rows as samples; returns
    the host features and targets in the order the samples were added."""
    from oracle import mlp
    xs, ts = [], []
    for kind, seed, kw in cases:
        A, sh = gen(spg, kind, seed, **kw)
        C = plain(spg, ctx, A, sh, A, sh)
        P.add_samples(ctx)
        x, ub = host_features(A, sh, A, sh)
        cnt = np.diff(C[0])
        keep = ub > 0
        xs.append(x[keep])
        ts.append(mlp.target(cnt[keep], ub[keep], sh[1]))
    return np.concatenate(xs), np.concatenate(ts)


def test_gradient_matches_backprop(spg, gpu_ctx):
    from oracle import mlp
    print("case learned gradient", flush=True)
    P = spg.Predictor(0, seed=5)
    try:
        x, t = _samples(spg, gpu_ctx, P, [("rmat", 1, dict(s=11, ef=8)),
                                          ("uniform", 2, dict(n=2048, k=6))])
        loss, g = P.gradient()
        w, norm = P.get_weights()
        assert np.allclose(norm, mlp.normalization(x).astype(np.float32), rtol=1e-4, atol=1e-5)
        lref, gref = mlp.loss_and_grad(w, norm.astype(np.float64), x, t)
        print(f"  loss gpu {loss:.6f} ref {lref:.6f}; samples {len(t)}")
        assert abs(loss - lref) <= 1e-3 * max(lref, 1e-3)
        assert np.linalg.norm(g - gref) <= GRAD_RTOL * np.linalg.norm(gref)
        # each parameter block is non-trivially exercised
        for lo, hi in [(0, 256), (256, 320), (320, 4416), (4416, 4480), (4480, 4544)]:
            assert np.linalg.norm(g[lo:hi] - gref[lo:hi]) <= 1e-2 * np.linalg.norm(gref[lo:hi]) + 1e-6
    finally:
        P.close()


def test_training_converges_and_is_deterministic(spg, gpu_ctx, tmp_path):
    print("case learned training", flush=True)
    cases = [("rmat", s, dict(s=10 + s % 3, ef=4 + 4 * (s % 3))) for s in range(6)]
    cases += [("uniform", 20 + s, dict(n=2048, k=4 + 3 * s)) for s in range(3)]
    weights = []
    for rep in range(2):
        P = spg.Predictor(0, seed=7)
        try:
            _samples(spg, gpu_ctx, P, cases)
            l0, _ = P.gradient()
            l1 = P.train(epochs=40, batch=1024, lr=2e-3)
            l2, _ = P.gradient()
            print(f"  loss before {l0:.4f} after {l2:.4f} (last epoch {l1:.4f})")
            assert np.isfinite(l2) and l2 < 0.5 * l0
            weights.append(P.get_weights())
            if rep == 0:
                path = str(tmp_path / "mlp.bin")
                P.save(path)
                Q = spg.Predictor(0, seed=99)
                Q.load(path)
                wq, nq = Q.get_weights()
                Q.close()
                assert np.array_equal(wq, weights[0][0]) and np.array_equal(nq, weights[0][1])
        finally:
            P.close()
    assert np.array_equal(weights[0][0], weights[1][0])


@pytest.mark.parametrize("label,kind,seed,kw", CASES, ids=[c[0] for c in CASES])
def test_learned_equals_plain(spg, orc, gpu_ctx, label, kind, seed, kw):
    """Predictions from a briefly trained network (some rows over- and some
    under-predicted): the learned merge gives the plain path's C."""
    print(f"case learned vs plain {label}", flush=True)
    A, sh = gen(spg, kind, seed, **kw)
    P = spg.Predictor(0, seed=3)
    try:
        _samples(spg, gpu_ctx, P, [("rmat", 31, dict(s=11, ef=8)), ("uniform", 32, dict(n=2048, k=8))])
        P.train(epochs=10, batch=512, lr=2e-3)
        Cp = plain(spg, gpu_ctx, A, sh, A, sh)
        Cl, st = learned(spg, gpu_ctx, A, sh, A, sh, predictor=P)
    finally:
        P.close()
    print(f"  stats {st}")
    same_c(Cl, Cp)
    ref, _ = orc.spgemm_restated(sh[0], sh[1], sh[0], csc_of(A, sh), csr_of(A, sh))
    assert np.array_equal(Cl[0], ref[0]) and np.array_equal(Cl[1], ref[1])
    assert st["unrecovered"] == 0 and st["nnz"] == Cp[0][-1]


@pytest.mark.parametrize("tight", [1, 0], ids=["tight", "classmax"])
@pytest.mark.parametrize("mode", ["zero", "huge", "random", "exact"])
@pytest.mark.parametrize("tile_log2,long_ub", [(13, 4096), (7, 64)])
def test_forced_predictions(spg, gpu_ctx, mode, tile_log2, long_ub, tight):
    """Predictions forced from the host: all zero (with tight capacities
    every row with more than one column overflows and is merged again), far
    too large (capacities at the upper bound), random around the exact
    count, exact."""
    print(f"case forced predictions {mode} tile_log2={tile_log2} tight={tight}", flush=True)
    A, sh = gen(spg, "rmat", 6, s=12, ef=12)
    cfg = dict(tile_log2=tile_log2, long_row_ub=long_ub)
    Cp = plain(spg, gpu_ctx, A, sh, A, sh, cfg)
    cnt = np.diff(Cp[0]).astype(np.float64)
    rng = np.random.default_rng(0)
    pred = {"zero": np.zeros(sh[0]), "huge": np.full(sh[0], 1e12),
            "random": cnt * rng.uniform(0.0, 2.0, sh[0]), "exact": cnt}[mode]
    lcfg = dict(cfg, pred_tight=tight)
    if mode != "huge":
        lcfg.update(pred_margin=0.0, pred_slack=0)
    Cl, st = learned(spg, gpu_ctx, A, sh, A, sh, pred=pred.astype(np.float32), cfg=lcfg)
    print(f"  stats {st}")
    same_c(Cl, Cp)
    assert st["unrecovered"] == 0
    if mode == "zero":
        if tight:  # capacity 1: every row with two or more columns
            assert st["overflow_vrows"] > 0.5 * st["vrows"]
        elif tile_log2 == 13:  # one tile per row: the 256-column wave class
            assert st["overflow_vrows"] == int(np.sum(cnt > 256))
    if mode == "huge" or (mode == "exact" and tile_log2 == 13):
        assert st["overflow_vrows"] == 0
    if mode == "exact" and tile_log2 < 13:
        # a row's exact count is shared among its column tiles by weight,
        # so with tight capacities some tiles still overflow
        assert st["overflow_vrows"] < 0.5 * st["vrows"]
    if mode == "random" and tight:
        assert 0 < st["overflow_vrows"] < st["vrows"]


def test_forced_overflow_wide_hash(spg, gpu_ctx):
    """n > 2^18: rows over the full column range take LDS hash tables sized
    from the prediction; zero predictions fill them up (reported as -1)
    and the exact hash fallback merges them."""
    print("case forced overflow wide hash", flush=True)
    rng = np.random.default_rng(5)
    m, K, n = 1200, 1000, 300000
    ra = np.repeat(np.arange(m), 20).astype(np.int32)
    ca = rng.integers(0, K, len(ra)).astype(np.int32)
    rb = np.repeat(np.arange(K), 90).astype(np.int32)
    cb = rng.integers(0, n, len(rb)).astype(np.int32)
    A = (ra, ca, rng.uniform(-1, 1, len(ra)).astype(np.float32))
    B = (rb, cb, rng.uniform(-1, 1, len(rb)).astype(np.float32))
    Cp = plain(spg, gpu_ctx, A, (m, K), B, (K, n))
    for pred in (np.zeros(m), np.full(m, 200.0), np.diff(Cp[0]).astype(np.float64)):
        Cl, st = learned(spg, gpu_ctx, A, (m, K), B, (K, n), pred=pred.astype(np.float32),
                         cfg=dict(pred_margin=0.0, pred_slack=0, pred_tight=1))
        print(f"  stats {st}")
        same_c(Cl, Cp)
        assert st["unrecovered"] == 0


def test_learned_requires_predictions(spg, gpu_ctx):
    A, sh = gen(spg, "uniform", 1, n=256, k=4)
    load(spg, gpu_ctx, A, sh, A, sh)
    gpu_ctx.init_gemm()
    gpu_ctx.categorize_gemm()
    gpu_ctx.bin_gemm()
    with pytest.raises(spg.SpgError):
        gpu_ctx.learned_merge_gemm()
