"""Benchmark of the outer-product SpGEMM hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat_s18_ef16]
                [--path both|learned|plain]

A step is one pass of the hot path over operands already resident in HBM.
Before each path is timed, the autotuner (spgTune, untimed) picks its
tunables on the same operands.  Two paths are timed on the same product:

  plain:   InitGEMM (upper bound, tiling), CategorizeGEMM (k-bins), BinGEMM
           (binned expansion), SymbolicGEMM (exact count), MergeGEMM
  learned: InitGEMM, CategorizeGEMM, BinGEMM, PredictRows (MLP row-count
           prediction), LearnedMergeGEMM (merge into predicted capacities,
           exact fallback for overflowed rows, compaction into CSR)

The predictor (spgemm_amd/models/mix32.mlp, or trained first with
--retrain, not timed) was trained on the exact counts of 32 seeded smaller
products (R-MAT, uniform rows, stencils).  Prints one JSON line; value = the
GFLOP/s of the path the workload is defined on (the R-MAT s18, stencil and
uniform products: the plain exact-count path; the rectangular product: the
learned path; batch256: learned + batched), with the other path beside it.
The roofline object prices the expansion + merge kernels of one step
against MI355X's 8 TB/s HBM peak with the algorithmic byte count
    16*nnz(A) + 8*nnz(B) + 16*nprod + 8*nnz(C) + 8*(m+1)
divided by their HIP-event time on the library's stream.  cpu_baseline
times the oracle's restatement of the reference method on a bounded row
sample of the same product on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import spgemm_amd as spg  # noqa: E402

METRIC = "SpGEMM GFLOP/s & %HBM peak, learned+batched vs plain, exact structure"
TUNED_FIELDS = ("tile_log2", "long_row_ub", "item_products", "expand_mode", "heavy_factor",
                "rank_min_count")
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # headline: the path the workload is defined on (the other is reported beside)
    # C = A*A, A = R-MAT scale 18, edge factor 16, (0.57, 0.19, 0.19, 0.05)
    "rmat_s18_ef16": dict(kind="rmat", row_bits=18, col_bits=18, ef=16, seed=0,
                          headline="plain"),
    # C = A*A, A = 27-point stencil on 64^3
    "stencil27_64": dict(kind="stencil", L=64, seed=3, headline="plain"),
    # C = A*A, n = 4096, 8 distinct columns per row
    "uniform_4096x8": dict(kind="uniform", n=4096, k=8, seed=1, headline="plain"),
    # C = A*B, A = 2^20 x 2^18 R-MAT (8 * 2^20 draws), B = R-MAT scale 18, ef 8;
    # learned allocation (mispredict, overflow and fallback rates reported)
    "rect_rmat_20x18_ef8": dict(kind="rect", seed=4, headline="learned"),
    # 256 independent C = A*A, n log-uniform in [1024, 65536]; even: uniform
    # rows with 4..32 nnz per row, odd: R-MAT scale ceil(log2 n) clipped to
    # n, edge factor 4..16
    "batch256": dict(kind="batch", count=256, seed=5),
}


def make_batch(count, seed):
    rng = np.random.default_rng(seed)
    pairs = []
    for j in range(count):
        n = int(np.floor(np.exp(rng.uniform(np.log(1024), np.log(65536)))))
        if j % 2 == 0:
            A = spg.gen_uniform_rows(n, n, int(rng.integers(4, 33)), seed * 1000 + j)
        else:
            s = int(np.ceil(np.log2(n)))
            A = spg.gen_rmat(s, s, int(rng.integers(4, 17)) * n, seed=seed * 1000 + j,
                             clip_rows=n, clip_cols=n)
        pairs.append((A, A))
    return pairs


def make_workload(name):
    w = WORKLOADS[name]
    if w["kind"] == "rmat":
        A = spg.gen_rmat(w["row_bits"], w["col_bits"], w["ef"] << w["row_bits"], seed=w["seed"])
    elif w["kind"] == "stencil":
        A = spg.gen_stencil27(w["L"], w["L"], w["L"], w["seed"])
    elif w["kind"] == "rect":
        A = spg.gen_rmat(20, 18, 8 << 20, seed=w["seed"])
        B = spg.gen_rmat(18, 18, 8 << 18, seed=w["seed"] + 1)
        return A, B
    else:
        A = spg.gen_uniform_rows(w["n"], w["n"], w["k"], w["seed"])
    return A, A


def run_step(ctx):
    nprod = ctx.init_gemm()
    ctx.categorize_gemm()
    ctx.bin_gemm()
    nnz = ctx.symbolic_gemm()
    ctx.merge_gemm()
    return nprod, nnz


def run_step_learned(ctx, P, m):
    nprod = ctx.init_gemm()
    ctx.categorize_gemm()
    ctx.bin_gemm()
    ctx.predict_rows(P)
    st = ctx.learned_merge_gemm()
    return nprod, st["nnz"], st


def timed(ctx, step, steps, warmup):
    for _ in range(warmup):
        step()
    ctx.synchronize()
    phase_sum = {p: 0.0 for p in spg.PHASES}
    t0 = time.perf_counter()
    for _ in range(steps):
        r = step()
        for k, v in ctx.phase_times().items():
            phase_sum[k] += v
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return r, ms, {k: v / steps for k, v in phase_sum.items() if v > 0}


def roofline_of(bytes_alg, phase_ms, workload, path):
    t_em = (phase_ms.get("expand", 0.0) + phase_ms.get("merge", 0.0)) * 1e-3
    achieved = bytes_alg / t_em / 1e9 if t_em > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(workload, path),
            "kernels": "expansion + merge kernels of one step (k_expand*, k_merge_*)",
            "bytes_alg_per_step": int(bytes_alg), "kernel_ms_per_step": round(t_em * 1e3, 4)}


def load_traffic(workload, path):
    """HBM bytes per step of the expansion+merge kernels of the path measured
    with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), if a summary exists."""
    p = os.path.join(ROOT, "profiles", f"pmc_{workload}_{path}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return float(json.load(f)["hbm_bytes_per_step"])
    except Exception:
        return None


def cpu_threads_available():
    """Host threads the CPU baseline may use: the CPUs this process may run
    on, at most 16 (one GPU's share of the host)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _median_runs(run, runs=3):
    """1 untimed warm-up, then the median wall time (s) of runs timed runs."""
    run()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _row_blocks(base, m, products):
    """Three disjoint row blocks (start, middle, 90 %) of about products/3
    intermediate products each."""
    per = max(1, int(products) // 3)
    starts = (0, m // 2, (9 * m) // 10)
    blocks = []
    for q, start in enumerate(starts):
        limit = starts[q + 1] if q + 1 < len(starts) else m
        e = int(np.searchsorted(base, base[start] + per, side="left"))
        blocks.append((start, min(limit, max(start + 1, e))))
    return blocks


def cpu_baseline(A, B, threads, seconds_per_run=3.0):
    """The oracle's restatement of the reference method (upper bound, k-bins,
    expansion, merge, compaction) on row blocks of the same product, timed
    on the host: single-threaded and with `threads` threads, each as 1
    warm-up + the median of 3 runs of a sample sized to about
    seconds_per_run (a calibration run on a small block fixes the size)."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import csc_of, csr_of
    m, K = A.num_rows, A.num_cols
    n = B.num_cols
    Acsc = csc_of((A.rows, A.cols, A.vals), (m, K))
    Bcsr = csr_of((B.rows, B.cols, B.vals), (K, n))
    ub, base, nprod = O.upper_bound(m, K, Acsc[0], Acsc[1], Bcsr[0])
    res = {}
    for nt in sorted({1, threads}):
        calib = _row_blocks(base, m, min(nprod, 2e7 * nt))
        t0 = time.perf_counter()
        got = sum(O.cpu_baseline(m, n, K, Acsc, Bcsr, a, b, nt)["nprod"] for a, b in calib)
        rate = max(got, 1) / max(time.perf_counter() - t0, 1e-6)
        blocks = _row_blocks(base, m, min(nprod, rate * seconds_per_run))
        prods = [0]

        def run():
            prods[0] = sum(O.cpu_baseline(m, n, K, Acsc, Bcsr, a, b, nt)["nprod"] for a, b in blocks)

        sec = _median_runs(run)
        res[nt] = {"value": round(2.0 * prods[0] / sec / 1e9, 4), "sec": round(sec, 3),
                   "products": int(prods[0]), "blocks": blocks}
    def desc(r):
        return ("rows " + ", ".join(f"[{a},{b})" for a, b in r["blocks"]) +
                f" of the same product ({r['products']} of its {nprod} intermediate products), "
                f"restated method (upper bound, k-bins, expansion, merge, compaction), "
                f"1 warm-up + median of 3 runs ({r['sec']} s each)")

    top = res[threads]
    return {"value": top["value"], "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": desc(top), "cores_visible": os.cpu_count(),
            "single_thread": {"value": res[1]["value"], "unit": "GFLOP/s", "cores": 1,
                              "sample": desc(res[1])}}


def cpu_baseline_batch(pairs, threads, seconds_per_run=2.0):
    """The restated method on whole products of the batch (seeded random
    order, as many as take about seconds_per_run), 1 warm-up + median of 3,
    with `threads` threads."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import csc_of, csr_of
    order = np.random.default_rng(1).permutation(len(pairs))
    ops = []
    spent, total = 0.0, 0
    for b in order:
        A, B = pairs[b]
        m, K, n = A.num_rows, A.num_cols, B.num_cols
        op = (m, n, K, csc_of((A.rows, A.cols, A.vals), (m, K)), csr_of((B.rows, B.cols, B.vals), (K, n)))
        t0 = time.perf_counter()
        total += O.cpu_baseline(*op, 0, m, threads)["nprod"]
        spent += time.perf_counter() - t0
        ops.append((int(b), op))
        if spent >= seconds_per_run:
            break

    def run():
        for _, op in ops:
            O.cpu_baseline(*op, 0, op[0], threads)

    sec = _median_runs(run)
    return {"value": round(2.0 * total / sec / 1e9, 4), "unit": "GFLOP/s", "cores": threads,
            "kind": "port", "cores_visible": os.cpu_count(),
            "sample": f"{len(ops)} of {len(pairs)} products of the batch (seeded order), "
                      f"{total} intermediate products, restated method (upper bound, k-bins, "
                      f"expansion, merge, compaction), 1 warm-up + median of 3 runs ({sec:.2f} s)"}


def bench_batch(args):
    """--workload batch256: the batch's own line."""
    w = WORKLOADS[args.workload]
    pairs = make_batch(w["count"], w["seed"])
    if args.batch_only:
        return bench_batch_only(args, pairs)
    out = batch_result(args, pairs, "batch256")
    print(json.dumps(out), flush=True)


def batch_result(args, pairs, workload):
    """256 uneven products: one product per launch sequence (plain path,
    every product's operands resident on its own context) against the whole
    batch in one replayed graph (plain, and with learned allocation)."""
    from spgemm_amd import train_predictor as tp
    nprods = []
    ctxs = []
    cfg = dict(expand_mode=spg.EXPAND_KBINS)
    # the batched context (and its merge streams) is created before the 256
    # one-product contexts, so its streams hold hardware queues of their own
    # (the process has few; streams created later share them)
    ctx = spg.Context(0)
    ctx.set_config(**cfg)
    P, train_info = (tp.train(ctx) if args.retrain else tp.load_or_train(ctx))
    ctx.batch_set_operands(pairs)
    for A, B in pairs:  # operands resident before timing; buffers sized by a warm-up
        c = spg.Context(0)
        c.set_config(**cfg)
        c.set_operands_from_coo(A, B)
        ctxs.append(c)
    for _ in range(max(1, args.warmup)):
        for c in ctxs:
            c.multiply()
    ctxs[0].synchronize()
    nnz_c = []
    for c in ctxs:
        nnz_c.append(int(c.get_c_device()[3]))
    lat = np.zeros((args.steps, len(ctxs)))
    for s in range(args.steps):
        for b, c in enumerate(ctxs):
            t0 = time.perf_counter()
            nprods_b = c.init_gemm()
            c.categorize_gemm()
            c.bin_gemm()
            c.symbolic_gemm()
            c.merge_gemm()
            c.synchronize()
            lat[s, b] = (time.perf_counter() - t0) * 1e3
            if s == 0:
                nprods.append(nprods_b)
    for c in ctxs:
        c.close()
    nprod = int(sum(nprods))
    flops = 2.0 * nprod
    plain_ms = float(lat.sum(axis=1).mean())
    per = lat.mean(axis=0)

    res = {}
    base_cfg = ctx.get_config()
    for name, pred in (("batched_plain", None), ("batched_learned", P)):
        ctx.set_config(**base_cfg)
        tuned = apply_tuning(ctx, args, pred, pred is not None)
        ctx.batch_multiply(predictor=pred)          # eager run: phase times
        ctx.synchronize()
        phase = {k: v for k, v in ctx.phase_times().items() if v > 0}
        st = ctx.batch_multiply(predictor=pred, graph=True)  # record + capture
        ctx.synchronize()
        for _ in range(args.warmup):
            ctx.batch_multiply(predictor=pred, graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.batch_multiply(predictor=pred, graph=True)
            ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        # per-product completion times from the device clock (separate,
        # untimed replays: recording adds one atomic per written row)
        ctx.batch_multiply(predictor=pred, graph=True, latency=True)  # record + capture
        ctx.batch_multiply(predictor=pred, graph=True, latency=True)  # replay
        lat = ctx.batch_latency(len(pairs))
        res[name] = dict(ms=ms, phase_ms=phase, stats=st, tuned=tuned, lat=lat[lat > 0])
    nnz_total = int(ctx.get_c_device()[3])
    assert nnz_total == sum(nnz_c), "batched and per-product nnz(C) differ"
    P.close()
    ctx.close()

    nnzA = sum(A.nnz for A, _ in pairs)
    nnzB = sum(B.nnz for _, B in pairs)
    m_tot = sum(A.num_rows for A, _ in pairs)
    bytes_alg = 16 * nnzA + 8 * nnzB + 16 * nprod + 8 * nnz_total + 8 * (m_tot + len(pairs))
    head = res["batched_learned"]
    gfl = {k: flops / (r["ms"] * 1e-3) / 1e9 for k, r in res.items()}
    plain_gfl = flops / (plain_ms * 1e-3) / 1e9
    out = {"metric": METRIC, "value": round(gfl["batched_learned"], 2), "unit": "GFLOP/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(head["ms"], 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batch256: 256 independent C=A*A, n log-uniform in "
                                  "[1024, 65536], half uniform rows (4-32 nnz/row), half R-MAT "
                                  "(ef 4-16), seed 5",
                      "path": "learned allocation + batched (block-diagonal stacking, one "
                              "replayed HIP graph for the batch)",
                      "products": len(pairs), "nnz_A": int(nnzA), "nprod": nprod,
                      "nnz_C": nnz_total, "parallelism": "single GPU"},
           "phase_ms": {k: round(v, 4) for k, v in head["phase_ms"].items()},
           "roofline": roofline_of(bytes_alg, head["phase_ms"], workload, "learned"),
           "plain_one_per_launch": {
               "value": round(plain_gfl, 2), "ms_per_batch": round(plain_ms, 4),
               "latency_ms_p50": round(float(np.percentile(per, 50)), 4),
               "latency_ms_p99": round(float(np.percentile(per, 99)), 4),
               "nprod_p50": int(np.percentile(nprods, 50)), "nprod_p99": int(np.percentile(nprods, 99))},
           "learned_batched_vs_plain": round(gfl["batched_learned"] / plain_gfl, 4),
           "learned_batched_vs_batched_plain": round(gfl["batched_learned"] / gfl["batched_plain"], 4)}
    for k, r in res.items():
        out[k] = {"value": round(gfl[k], 2), "ms_per_batch": round(r["ms"], 4),
                  "latency_ms_p50": round(float(np.percentile(r["lat"], 50)), 4),
                  "latency_ms_p99": round(float(np.percentile(r["lat"], 99)), 4),
                  "latency_source": "device wall clock when each product's last C row is "
                                    "written, from the start of a graph replay",
                  "eager_phase_ms": {p: round(v, 4) for p, v in r["phase_ms"].items()},
                  "roofline_frac": roofline_of(bytes_alg, r["phase_ms"], workload,
                                               "learned")["frac"]}
        if r["tuned"]:
            out[k]["tuned_config"] = {f: r["tuned"][f] for f in TUNED_FIELDS}
    st = res["batched_learned"]["stats"]
    out["batched_learned"]["rates"] = {
        "overflow_vrow_rate": round(st["overflow_vrows"] / max(st["vrows"], 1), 6),
        "fallback_slot_fraction": round(st["fallback_slots"] / max(st["nnz"], 1), 6),
        "padded_slots_per_nnz": round(st["padded_slots"] / max(st["nnz"], 1), 4),
        "unrecovered_overflows": int(st["unrecovered"])}
    out["batched_learned"]["predictor_training"] = train_info
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(pairs, args.cpu_threads or cpu_threads_available())
    return out


def bench_batch_only(args, pairs):
    """The batched runs alone (for profiling): warmup + steps eager passes
    of the --path paths (plain, learned or both), then, unless --no-graph,
    steps graph calls of the learned pass (the first records and captures),
    with the learned pass's settings (--set, --tune-table entry, else tuned
    here first)."""
    from spgemm_amd import train_predictor as tp
    ctx = spg.Context(0)
    P, _ = tp.load_or_train(ctx)
    ctx.batch_set_operands(pairs)
    apply_tuning(ctx, args, P, True)
    preds = {"both": (None, P), "learned": (P,), "plain": (None,)}[args.path]
    for _ in range(args.warmup + args.steps):
        for pred in preds:
            ctx.batch_multiply(predictor=pred)
    if not args.no_graph and args.path != "plain":
        for _ in range(args.steps):
            ctx.batch_multiply(predictor=P, graph=True)
    ctx.synchronize()
    print(json.dumps({"batch_only": True, "phase_ms": ctx.phase_times()}), flush=True)
    P.close()
    ctx.close()


def bench_mtx_list(args):
    """The reference's benchmark sweep (bin/launch.sh: ./outer X.mtx X.mtx for
    every matrix of a list): C = A*A for each Matrix Market file named in
    the list file (one path per line; blank lines and # comments skipped),
    read with readMTX (values kept; --compat-values forces 1.1 as the
    reference does).  Prints one JSON line per matrix: nprod, nnz(C) and the
    GFLOP/s of the plain and the learned path (operands resident, median of
    --steps timed steps after --warmup)."""
    from spgemm_amd import train_predictor as tp
    with open(args.mtx_list) as f:
        paths = [ln.strip() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    base = os.path.dirname(os.path.abspath(args.mtx_list))
    ctx = spg.Context(0)
    P, _ = tp.load_or_train(ctx)
    n_ok = 0
    for p in paths:
        full = p if os.path.isabs(p) else os.path.join(base, p)
        rec = {"sweep": "C=A*A", "matrix": p}
        try:
            A = spg.read_mtx(full, compat_values=args.compat_values)
        except spg.SpgError as e:
            rec["error"] = f"read failed: {e}"
            print(json.dumps(rec), flush=True)
            continue
        rec.update(m=A.num_rows, n=A.num_cols, nnz_A=int(A.nnz))
        if A.num_rows != A.num_cols:
            rec["error"] = "not square: A*A undefined"
            print(json.dumps(rec), flush=True)
            continue
        ctx.set_operands_from_coo(A, A)
        times = {"plain": [], "learned": []}
        for _ in range(args.warmup + args.steps):
            for name in ("plain", "learned"):
                t0 = time.perf_counter()
                if name == "plain":
                    nprod, nnz = run_step(ctx)
                else:
                    nprod, nnz_l, _ = run_step_learned(ctx, P, A.num_rows)
                    assert nnz_l == nnz, "learned and plain nnz(C) differ"
                ctx.synchronize()
                times[name].append((time.perf_counter() - t0) * 1e3)
        C = ctx.get_c()
        rec.update(nprod=int(nprod), nnz_C=int(nnz))
        rec["nnz_C_above_1e-4"] = int(np.count_nonzero(C[2] > 1e-4))  # src/mm.cu:438-449
        for name, t in times.items():
            ms = float(np.median(t[args.warmup:]))
            rec[name] = {"ms": round(ms, 4), "gflops": float("%.4g" % (2.0 * nprod / (ms * 1e-3) / 1e9))}
        print(json.dumps(rec), flush=True)
        n_ok += 1
    P.close()
    ctx.close()
    return 0 if n_ok == len(paths) else 1


def copy_probe(mib):
    """One device-to-device copy of mib MiB (the HIP runtime's copy kernel):
    the known-size transfer that scales the FETCH_SIZE / WRITE_SIZE counters
    of a rocprofv3 --pmc run into bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    for _ in range(2):  # 3 = hipMemcpyDeviceToDevice
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n, "copies": 2}), flush=True)


def parse_set(text):
    """--set tile_log2=13,long_row_ub=512: fixed settings (spg_config fields)."""
    out = {}
    for item in filter(None, (text or "").split(",")):
        k, v = item.split("=")
        out[k.strip()] = float(v) if k.strip() == "pred_margin" else int(v)
    return out


def apply_tuning(ctx, args, predictor, learned):
    """Tunes the context for the path (untimed), or applies the entry of
    --tune-table for this workload when there is one, or the fixed --set
    settings; returns the settings."""
    if args.set:
        ctx.set_config(**parse_set(args.set))
        return ctx.get_config()
    if args.no_tune:
        return None
    if args.tune_table and ctx.tune_lookup(args.tune_table, learned=learned):
        return ctx.get_config()
    return ctx.tune(predictor=predictor, reps=3, seed=0, table=args.tune_table)[0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat_s18_ef16", choices=sorted(WORKLOADS))
    ap.add_argument("--path", default="both", choices=("both", "learned", "plain"))
    ap.add_argument("--retrain", action="store_true",
                    help="train the predictor instead of loading spgemm_amd/models/mix32.mlp")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (untimed) that picks each path's tunables first")
    ap.add_argument("--set", default=None,
                    help="fixed settings instead of tuning, e.g. tile_log2=13,long_row_ub=512")
    ap.add_argument("--tune-table", default=None,
                    help="JSON-lines table: apply its entry for the workload, else tune and append")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (counter calibration)")
    ap.add_argument("--no-graph", action="store_true",
                    help="--batch-only: eager passes only (no graph replays)")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch workloads: skip the one-product-per-launch baseline (profiling)")
    ap.add_argument("--mtx-list", default=None,
                    help="sweep: C=A*A for every Matrix Market file named in this list file")
    ap.add_argument("--compat-values", action="store_true",
                    help="--mtx-list: every value read as 1.1, as the reference's readMTX")
    ap.add_argument("--no-batch", action="store_true",
                    help="default workload: skip the attached 256-product batch measurement")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.copy_probe_mib:
        return copy_probe(args.copy_probe_mib)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "C row blocks across GPUs but this library runs on a single MI355X.", file=sys.stderr)
        sys.exit(2)
    if args.mtx_list:
        sys.exit(bench_mtx_list(args))

    if WORKLOADS[args.workload]["kind"] == "batch":
        return bench_batch(args)
    ctx = spg.Context(0)
    P = train_info = None
    if args.path != "plain":
        from spgemm_amd import train_predictor as tp
        if args.retrain:
            P, train_info = tp.train(ctx)
        else:
            P, train_info = tp.load_or_train(ctx)
    A, B = make_workload(args.workload)
    ctx.set_operands_from_coo(A, B)
    m, n = A.num_rows, B.num_cols

    res = {}
    tuned = {}
    base_cfg = ctx.get_config()
    if args.path != "learned":
        tuned["plain"] = apply_tuning(ctx, args, None, False)
        (nprod, nnz), ms, ph = timed(ctx, lambda: run_step(ctx), args.steps, args.warmup)
        exact = np.diff(ctx.get_c()[0]) if args.path == "both" else None
        res["plain"] = dict(nprod=nprod, nnz=nnz, ms=ms, phase_ms=ph)
    if args.path != "plain":
        ctx.set_config(**base_cfg)
        tuned["learned"] = apply_tuning(ctx, args, P, True)
        (nprod, nnz, st), ms, ph = timed(ctx, lambda: run_step_learned(ctx, P, m), args.steps,
                                         args.warmup)
        res["learned"] = dict(nprod=nprod, nnz=nnz, ms=ms, phase_ms=ph, stats=st)
        pred = ctx.predict_rows(P, m)
        cnt = np.diff(ctx.get_c()[0])
        if "plain" in res:
            assert np.array_equal(cnt, exact), "learned and plain row counts differ"
        live = ctx.upper_bound(m) > 0
        relerr = np.abs(pred[live] - cnt[live]) / np.maximum(cnt[live], 1)
        res["learned"]["rates"] = {
            "rows": int(live.sum()),
            "under_predicted_rows": round(float(np.mean(pred[live] < cnt[live])), 6),
            "mean_abs_rel_err": round(float(relerr.mean()), 6),
            "median_abs_rel_err": round(float(np.median(relerr)), 6),
            "overflow_vrow_rate": round(st["overflow_vrows"] / max(st["vrows"], 1), 6),
            "fallback_slot_fraction": round(st["fallback_slots"] / max(st["nnz"], 1), 6),
            "padded_slots_per_nnz": round(st["padded_slots"] / max(st["nnz"], 1), 4),
            "unrecovered_overflows": int(st["unrecovered"])}

    path_name = WORKLOADS[args.workload]["headline"]
    if path_name not in res:
        path_name = "learned" if "learned" in res else "plain"
    head = res[path_name]
    nprod, nnz = head["nprod"], head["nnz"]
    flops = 2.0 * nprod
    gflops = {k: flops / (r["ms"] * 1e-3) / 1e9 for k, r in res.items()}
    bytes_alg = 16 * A.nnz + 8 * B.nnz + 16 * nprod + 8 * nnz + 8 * (m + 1)
    out = {"metric": METRIC, "value": round(gflops[path_name], 2), "unit": "GFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms"], 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"{args.workload} C=A*B" if B is not A else f"{args.workload} C=A*A",
                      "path": {"learned": "learned allocation: MLP row-count prediction, merge into "
                                          "predicted capacities, exact fallback, compaction",
                               "plain": "plain exact-count: symbolic pass then merge"}[path_name],
                      "m": m, "k": A.num_cols, "n": n, "nnz_A": int(A.nnz), "nnz_B": int(B.nnz),
                      "nprod": int(nprod), "nnz_C": int(nnz), "parallelism": "single GPU"},
           "phase_ms": {k: round(v, 4) for k, v in head["phase_ms"].items()},
           "roofline": roofline_of(bytes_alg, head["phase_ms"], args.workload, path_name)}
    for k, r in res.items():
        out[k] = {"value": round(gflops[k], 2), "ms_per_step": round(r["ms"], 4),
                  "phase_ms": {p: round(v, 4) for p, v in r["phase_ms"].items()},
                  "roofline_frac": roofline_of(bytes_alg, r["phase_ms"], args.workload, k)["frac"]}
    if "learned" in res:
        out["learned"]["rates"] = res["learned"]["rates"]
        out["learned"]["predictor_training"] = train_info
    if len(res) == 2:
        out["learned_vs_plain"] = round(gflops["learned"] / gflops["plain"], 4)
    for k, cfg in tuned.items():
        if cfg:
            out[k]["tuned_config"] = {f: cfg[f] for f in TUNED_FIELDS}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(A, B, args.cpu_threads or cpu_threads_available())
    if P is not None:
        P.close()
    ctx.close()
    if args.workload == "rmat_s18_ef16" and not args.no_batch:
        # the batched half of the headline ratio: the 256-product batch
        # (learned + batched graph vs batched plain vs one product per launch)
        w = WORKLOADS["batch256"]
        b = batch_result(args, make_batch(w["count"], w["seed"]), "batch256")
        out["batch256"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "config", "phase_ms",
                                             "roofline", "plain_one_per_launch", "batched_plain",
                                             "batched_learned", "learned_batched_vs_plain",
                                             "learned_batched_vs_batched_plain", "cpu_baseline")
                           if k in b}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
