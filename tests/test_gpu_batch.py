"""Batched products on the GPU (block-diagonal stacking, per-row column
windows, one run of the phase kernels for the whole batch, graph replay)
against one-product runs and the CPU oracle.

Bar: every product's C structure identical to the oracle's restated method
and to its own one-product GPU run; values within VAL_RTOL of the oracle;
every graph replay identical to the recorded run.
"""
import numpy as np
import pytest

from helpers import csc_of, csr_of, rel_err

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-5


def make_batch(spg, seed, count=10, big=False):
    """Uneven products: uniform rows, clipped R-MAT, a stencil, rectangular
    pairs, an empty one and a 1x1 one; some wider than one 8192 tile."""
    rng = np.random.default_rng(seed)
    pairs = []
    for j in range(count):
        kind = j % 5
        if kind == 0:
            n = int(rng.integers(300, 3000))
            U = spg.gen_uniform_rows(n, n, int(rng.integers(4, 20)), seed + j)
            pairs.append((U, U))
        elif kind == 1:
            n = int(rng.integers(2000, 20000 if big else 12000))
            s = int(np.ceil(np.log2(n)))
            R = spg.gen_rmat(s, s, int(rng.integers(4, 12)) << s, seed=seed + j, clip_rows=n,
                             clip_cols=n)
            pairs.append((R, R))
        elif kind == 2:
            S = spg.gen_stencil27(8, 9, 10, seed + j)
            pairs.append((S, S))
        elif kind == 3:
            A = spg.gen_rmat(11, 9, 6 << 11, seed=seed + j)
            B = spg.gen_rmat(9, 14, 5 << 9, seed=seed + 100 + j)
            pairs.append((A, B))
        else:
            m = int(rng.integers(100, 9000))
            A = spg.gen_uniform_rows(m, 700, 5, seed + j)
            B = spg.gen_uniform_rows(700, 9000, 9, seed + 50 + j)
            pairs.append((A, B))
    z = np.zeros(0, np.int32)
    pairs.append((spg.HostCOO(z, z, np.zeros(0, np.float32), 40, 30),
                  spg.HostCOO(z, z, np.zeros(0, np.float32), 30, 50)))
    one = spg.HostCOO(np.zeros(1, np.int32), np.zeros(1, np.int32), np.ones(1, np.float32), 1, 1)
    pairs.append((one, one))
    return pairs


def oracle_c(orc, A, B):
    return orc.spgemm_restated(A.num_rows, B.num_cols, A.num_cols,
                               csc_of((A.rows, A.cols, A.vals), (A.num_rows, A.num_cols)),
                               csr_of((B.rows, B.cols, B.vals), (B.num_rows, B.num_cols)))[0]


def check(C, R):
    assert np.array_equal(C[0], R[0]) and np.array_equal(C[1], R[1])
    if len(R[2]):
        err = rel_err(C[2].astype(np.float64), R[2].astype(np.float64))
        both_zero = (np.abs(C[2]) <= 1e-30) & (np.abs(R[2]) <= 1e-30)
        err[both_zero] = 0
        assert err.max() <= VAL_RTOL


CFGS = {"default": dict(tile_log2=13, long_row_ub=4096),
        "small_tiles": dict(tile_log2=9, long_row_ub=64)}


@pytest.mark.parametrize("mode", [0, 1], ids=["kbins", "rows"])
@pytest.mark.parametrize("cfg", sorted(CFGS))
def test_batch_matches_oracle(spg, orc, gpu_ctx, cfg, mode):
    print(f"case batch {cfg} mode={mode}", flush=True)
    pairs = make_batch(spg, 7)
    ctx = gpu_ctx
    ctx.set_config(item_products=4096, xcd_swizzle=1, timing=1, expand_mode=mode, **CFGS[cfg])
    ctx.batch_set_operands(pairs)
    ro, co = ctx.batch_layout()
    assert ro[-1] == sum(A.num_rows for A, _ in pairs)
    assert np.all(co % 8192 == 0)
    ctx.batch_multiply()
    for b, (A, B) in enumerate(pairs):
        check(ctx.batch_get_c(b), oracle_c(orc, A, B))


@pytest.mark.parametrize("mode", [0, 1], ids=["kbins", "rows"])
def test_batch_graph_replays(spg, orc, gpu_ctx, mode):
    """Record + capture, then replays with eager work and synchronizations
    in between; every replay gives the recorded run's C (both expansion
    mappings: the row-wise one adds its own readbacks, buffers and fills)."""
    print(f"case batch graph mode={mode}", flush=True)
    pairs = make_batch(spg, 11, count=8, big=True)
    ctx = gpu_ctx
    ctx.set_config(tile_log2=13, item_products=4096, long_row_ub=1024, xcd_swizzle=1, timing=1,
                   expand_mode=mode)
    ctx.batch_set_operands(pairs)
    ctx.batch_multiply(graph=True)
    first = [ctx.batch_get_c(b) for b in range(len(pairs))]
    for b, (A, B) in enumerate(pairs):
        check(first[b], oracle_c(orc, A, B))
    other = spg.Context(0)
    try:
        for rep in range(3):
            # unrelated eager work on another context between replays
            U = spg.gen_uniform_rows(500, 500, 6, 40 + rep)
            other.set_operands_from_coo(U, U)
            other.multiply()
            other.synchronize()
            ctx.batch_multiply(graph=True)
            ctx.synchronize()
            for b in range(len(pairs)):
                C = ctx.batch_get_c(b)
                for x, y in zip(C, first[b]):
                    assert np.array_equal(x, y), (rep, b)
    finally:
        other.close()


def test_batch_learned(spg, orc, gpu_ctx):
    print("case batch learned", flush=True)
    pairs = make_batch(spg, 13)
    ctx = gpu_ctx
    ctx.set_config(tile_log2=13, item_products=4096, long_row_ub=4096, xcd_swizzle=1, timing=1,
                   expand_mode=0, pred_margin=0.25, pred_slack=8, pred_tight=0)
    P = spg.Predictor(0, seed=2)
    try:
        for s in range(3):
            R = spg.gen_rmat(11, 11, 8 << 11, seed=60 + s)
            ctx.set_operands_from_coo(R, R)
            ctx.multiply()
            P.add_samples(ctx)
        P.train(epochs=10, batch=512, lr=2e-3)
        ctx.batch_set_operands(pairs)
        # eager; recorded + captured; replayed
        for graph, scramble in ((False, True), (True, True), (True, False)):
            if scramble:  # other output in the C buffers (unwritten positions would show)
                ctx.batch_set_operands(make_batch(spg, 99, count=3))
                ctx.batch_multiply()
                ctx.batch_set_operands(pairs)
            st = ctx.batch_multiply(predictor=P, graph=graph)
            ctx.synchronize()
            print(f"  graph={graph} stats {st}")
            assert st["unrecovered"] == 0
            for b, (A, B) in enumerate(pairs):
                check(ctx.batch_get_c(b), oracle_c(orc, A, B))
    finally:
        P.close()


def test_batch_then_single_product(spg, orc, gpu_ctx):
    """Setting ordinary operands after a batch leaves batch mode."""
    print("case batch then single", flush=True)
    ctx = gpu_ctx
    ctx.set_config(tile_log2=13, item_products=4096, long_row_ub=4096, xcd_swizzle=1, timing=1,
                   expand_mode=0)
    ctx.batch_set_operands(make_batch(spg, 3, count=4))
    ctx.batch_multiply()
    R = spg.gen_rmat(12, 12, 8 << 12, seed=9)
    ctx.set_operands_from_coo(R, R)
    ctx.multiply()
    check(ctx.get_c(), oracle_c(orc, R, R))
    with pytest.raises(spg.SpgError):
        ctx.batch_multiply()


@pytest.mark.parametrize("learned", [False, True], ids=["plain", "learned"])
def test_batch_completion_times(spg, orc, gpu_ctx, learned):
    """SPG_BATCH_LATENCY: every product's completion time from the device
    wall clock, eager and in a replayed graph.  The times lie between 0 and
    the host-timed length of the run, a product without entries reports 0,
    C is the same as without recording, and a later run without the flag
    leaves no times to read."""
    import time
    from spgemm_amd import train_predictor as tp
    print(f"case batch completion times learned={learned}", flush=True)
    ctx = gpu_ctx
    pairs = make_batch(spg, 71)
    ctx.batch_set_operands(pairs)
    P = tp.load_or_train(ctx)[0] if learned else None
    try:
        ctx.batch_multiply(predictor=P)
        ref = [ctx.batch_get_c(b) for b in range(len(pairs))]
        for graph in (False, True, True):  # eager, record + capture, replay
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.batch_multiply(predictor=P, graph=graph, latency=True)
            ctx.synchronize()
            wall = (time.perf_counter() - t0) * 1e3
            ms = ctx.batch_latency(len(pairs))
            nnz = np.array([ref[b][0][-1] for b in range(len(pairs))])
            assert np.all(ms[nnz > 0] > 0) and np.all(ms[nnz == 0] == 0)
            assert ms.max() <= wall * 1.05 + 0.05, (ms.max(), wall)
            for b in range(len(pairs)):
                C = ctx.batch_get_c(b)
                assert np.array_equal(C[0], ref[b][0]) and np.array_equal(C[1], ref[b][1])
                assert np.array_equal(C[2], ref[b][2])
            print(f"  graph={graph}: p50 {np.percentile(ms[nnz > 0], 50):.3f} ms, "
                  f"max {ms.max():.3f} ms, run {wall:.3f} ms")
        ctx.batch_multiply(predictor=P, graph=True)  # no recording: a new graph
        ctx.synchronize()
        with pytest.raises(spg.SpgError):
            ctx.batch_latency(len(pairs))
    finally:
        if P is not None:
            P.close()
