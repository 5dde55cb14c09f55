"""Autotuner on the GPU: the chosen settings come from the candidate
lists, persist in the table, are found again by spgTuneLookup, and the
product computed with them still matches the oracle exactly (structure)."""
import numpy as np
import pytest

from helpers import csc_of, csr_of, rel_err

pytestmark = pytest.mark.gpu

CANDIDATES = {"tile_log2": {11, 12, 13}, "long_row_ub": {512, 1024, 2048, 4096},
              "item_products": {2048, 4096, 8192, 16384, 32768}, "expand_mode": {0, 1},
              "heavy_factor": {10, 25, 50, 100, 200},
              "rank_min_count": {64, 128, 256, 512, 1024}}


def test_tune_table_roundtrip(spg, orc, gpu_ctx, tmp_path):
    print("case tune single product", flush=True)
    R = spg.gen_rmat(12, 12, 8 << 12, seed=17)
    table = str(tmp_path / "tuned.jsonl")
    ctx = gpu_ctx
    ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=4096, xcd_swizzle=1, timing=1,
                   expand_mode=0, heavy_factor=50)
    ctx.set_operands_from_coo(R, R)
    cfg, ms = ctx.tune(reps=1, seed=3, table=table)
    print(f"  chose {cfg} at {ms:.3f} ms")
    assert ms > 0
    for k, vals in CANDIDATES.items():
        assert cfg[k] in vals
    assert cfg["timing"] == 1  # the caller's setting survives the tuning runs
    lines = open(table).read().splitlines()
    assert len(lines) == 1 and '"key":"m=4096 k=4096 n=4096' in lines[0]
    other = spg.Context(0)
    try:
        other.set_operands_from_coo(R, R)
        assert other.tune_lookup(table)
        got = other.get_config()
        for k in CANDIDATES:
            assert got[k] == cfg[k]
        assert not other.tune_lookup(table, learned=True)  # no learned entry
        other.multiply()
        C = other.get_c()
    finally:
        other.close()
    ref, _ = orc.spgemm_restated(4096, 4096, 4096, csc_of((R.rows, R.cols, R.vals), (4096, 4096)),
                                 csr_of((R.rows, R.cols, R.vals), (4096, 4096)))
    assert np.array_equal(C[0], ref[0]) and np.array_equal(C[1], ref[1])
    assert rel_err(C[2], ref[2]).max(initial=0) <= 1e-5


def test_tune_batch_learned(spg, gpu_ctx, tmp_path):
    print("case tune batch learned", flush=True)
    ctx = gpu_ctx
    ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=4096, xcd_swizzle=1, timing=1,
                   expand_mode=0, heavy_factor=50)
    P = spg.Predictor(0, seed=1)
    try:
        pairs = []
        for j in range(6):
            U = spg.gen_uniform_rows(1500 + 400 * j, 1500 + 400 * j, 6 + j, 70 + j)
            pairs.append((U, U))
        ctx.batch_set_operands(pairs)
        cfg, ms = ctx.tune(predictor=P, reps=1, seed=0, table=str(tmp_path / "t.jsonl"))
        for k, vals in CANDIDATES.items():
            assert cfg[k] in vals
        ref = [None] * len(pairs)
        ctx.batch_multiply(predictor=P)
        got = [ctx.batch_get_c(b) for b in range(len(pairs))]
        ctx.set_config(**{k: v for k, v in cfg.items() if k != "timing"}, timing=1)
        ctx.set_config(tile_log2=13, item_products=8192, long_row_ub=4096, expand_mode=0,
                       heavy_factor=50)
        ctx.batch_multiply()
        for b in range(len(pairs)):
            ref[b] = ctx.batch_get_c(b)
            assert np.array_equal(got[b][0], ref[b][0]) and np.array_equal(got[b][1], ref[b][1])
    finally:
        P.close()
