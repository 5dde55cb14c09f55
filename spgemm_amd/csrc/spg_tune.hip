// spg_tune.hip - autotuner of the runtime tunables (no reference
// counterpart: the reference fixes its bin factor, bin count and block
// shapes at compile time, src/mm.cu:17,205,262).
//
// Coordinate descent over tile_log2, long_row_ub, item_products,
// expand_mode, heavy_factor and rank_min_count, from the fastest of the
// context's current settings and a few starting points (both expansion
// mappings at several long-row bounds): every value of
// one tunable is timed (wall clock of a full product or batch, median of
// reps runs after one warm-up) with the others held, the fastest kept if it
// beats the incumbent by more than 1 %, and the tunables are visited in an
// order drawn from the seed; then every pair of long_row_ub and
// rank_min_count (they interact) is timed, and a last sweep follows.  The winner is set on the context and
// can be appended to a table (one JSON object per line) keyed by the
// workload's shape, which spgTuneLookup applies later.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spg_internal.h"
#include "spg_rng.h"

using namespace spg;

namespace {

struct Dim {
    const char* name;
    std::vector<long long> values;
};

long long get(const spg_config& c, int d) {
    switch (d) {
        case 0: return c.tile_log2;
        case 1: return c.long_row_ub;
        case 2: return c.item_products;
        case 3: return c.expand_mode;
        case 4: return c.heavy_factor;
        default: return c.rank_min_count;
    }
}

void set(spg_config& c, int d, long long v) {
    switch (d) {
        case 0: c.tile_log2 = (int32_t)v; break;
        case 1: c.long_row_ub = v; break;
        case 2: c.item_products = (int32_t)v; break;
        case 3: c.expand_mode = (int32_t)v; break;
        case 4: c.heavy_factor = (int32_t)v; break;
        default: c.rank_min_count = (int32_t)v; break;
    }
}

std::string workload_key(spg_context* c, bool learned) {
    char buf[256];
    snprintf(buf, sizeof buf, "m=%d k=%d n=%d nnza=%d nnzb=%d batch=%d learned=%d",
             c->A.num_rows, c->A.num_cols, c->B.num_cols, c->A.nnz, c->B.nnz, c->batch_count,
             learned ? 1 : 0);
    return buf;
}

int run_once(spg_context* c, const spg_predictor* p) {
    if (c->batch_count > 0) return spgBatchMultiply(c, p, 0, nullptr);
    if (p) return spgMultiplyLearned(c, p, nullptr);
    return spgMultiply(c);
}

// Median wall time (ms) of reps runs after one warm-up, or a negative value
// when the configuration fails (e.g. a tile width the shape does not allow).
double time_config(spg_context* c, const spg_predictor* p, const spg_config& cfg, int reps) {
    if (spgSetConfig(c, &cfg) != SPG_OK) return -1.0;
    if (run_once(c, p) != SPG_OK || hipStreamSynchronize(c->stream) != hipSuccess) return -1.0;
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        if (run_once(c, p) != SPG_OK || hipStreamSynchronize(c->stream) != hipSuccess) return -1.0;
        t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                        .count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

bool field(const std::string& line, const char* name, long long* out) {
    const std::string pat = std::string("\"") + name + "\":";
    const size_t q = line.find(pat);
    if (q == std::string::npos) return false;
    return sscanf(line.c_str() + q + pat.size(), "%lld", out) == 1;
}

}  // namespace

extern "C" {

int spgTune(spg_context* c, const spg_predictor* p, int reps, uint64_t seed,
            const char* table_path, spg_config* best, double* best_ms) {
    if (!c) return SPG_EINVAL;
    if (!c->A.set || !c->B.set) return fail(c, SPG_EORDER, "operands must be set before tuning");
    if (reps < 1 || reps > 100) return fail(c, SPG_EINVAL, "reps must be in [1, 100]");
    const std::vector<Dim> dims = {
        {"tile_log2", {11, 12, 13}},
        {"long_row_ub", {512, 1024, 2048, 4096}},
        {"item_products", {2048, 4096, 8192, 16384, 32768}},
        {"expand_mode", {SPG_EXPAND_KBINS, SPG_EXPAND_ROWS}},
        {"heavy_factor", {10, 25, 50, 100, 200}},
        {"rank_min_count", {64, 128, 256, 512, 1024}},
    };
    const int32_t timing = c->cfg.timing;  // the caller's setting, restored on the winner
    spg_config inc = c->cfg;
    inc.timing = 0;
    double inc_ms = time_config(c, p, inc, reps);
    if (inc_ms < 0) return fail(c, SPG_EINVAL, "the starting configuration does not run");
    // a few starting points of both expansion mappings and several long-row
    // bounds (their best settings of the other tunables differ, so a descent
    // from one of them alone can stop in the other's basin)
    {
        const spg_config base = inc;
        const long long starts[][2] = {{SPG_EXPAND_ROWS, 512},  {SPG_EXPAND_ROWS, 1024},
                                       {SPG_EXPAND_ROWS, 2048}, {SPG_EXPAND_KBINS, 512},
                                       {SPG_EXPAND_KBINS, 1024}};
        for (const auto& st : starts) {
            if (st[0] == base.expand_mode && st[1] == base.long_row_ub) continue;
            spg_config cand = base;
            set(cand, 3, st[0]);
            set(cand, 1, st[1]);
            const double ms = time_config(c, p, cand, reps);
            if (ms >= 0 && ms < inc_ms) {
                inc = cand;
                inc_ms = ms;
            }
        }
    }
    std::vector<int> order = {0, 1, 2, 3, 4, 5};
    Xoshiro rng(seed);
    for (int q = 5; q > 0; --q) std::swap(order[(size_t)q], order[(size_t)rng.below((uint64_t)q + 1)]);
    auto sweep_once = [&]() {
        for (int d : order) {
            const long long cur = get(inc, d);
            for (long long v : dims[(size_t)d].values) {
                if (v == cur) continue;
                spg_config cand = inc;
                set(cand, d, v);
                const double ms = time_config(c, p, cand, reps);
                if (ms >= 0 && ms < inc_ms * 0.99) {
                    inc = cand;
                    inc_ms = ms;
                }
            }
        }
    };
    sweep_once();
    // long_row_ub and rank_min_count interact (rows split into tiles never
    // reach the wide-row classes), so one coordinate at a time can stall:
    // every pair of the two is timed once, then a last sweep
    {
        const spg_config base = inc;
        for (long long lu : dims[1].values)
            for (long long rm : dims[5].values) {
                if (lu == base.long_row_ub && rm == base.rank_min_count) continue;
                spg_config cand = base;
                set(cand, 1, lu);
                set(cand, 5, rm);
                const double ms = time_config(c, p, cand, reps);
                if (ms >= 0 && ms < inc_ms * 0.99) {
                    inc = cand;
                    inc_ms = ms;
                }
            }
    }
    sweep_once();
    inc.timing = timing;
    SPG_TRY(spgSetConfig(c, &inc));
    if (best) *best = inc;
    if (best_ms) *best_ms = inc_ms;
    if (table_path) {
        FILE* f = fopen(table_path, "a");
        if (!f) return fail(c, SPG_EIO, "cannot open the tuning table for appending");
        fprintf(f,
                "{\"key\":\"%s\",\"tile_log2\":%d,\"long_row_ub\":%lld,\"item_products\":%d,"
                "\"expand_mode\":%d,\"heavy_factor\":%d,\"rank_min_count\":%d,\"ms\":%.4f,"
                "\"seed\":%llu}\n",
                workload_key(c, p != nullptr).c_str(), inc.tile_log2, (long long)inc.long_row_ub,
                inc.item_products, inc.expand_mode, inc.heavy_factor, inc.rank_min_count, inc_ms,
                (unsigned long long)seed);
        if (fclose(f) != 0) return fail(c, SPG_EIO, "tuning table write failed");
    }
    return SPG_OK;
}

int spgTuneLookup(spg_context* c, const char* table_path, int learned, int* found) {
    if (!c || !table_path) return SPG_EINVAL;
    if (found) *found = 0;
    if (!c->A.set || !c->B.set) return fail(c, SPG_EORDER, "operands must be set first");
    FILE* f = fopen(table_path, "r");
    if (!f) return SPG_OK;  // no table yet: keep the current settings
    const std::string key = "\"key\":\"" + workload_key(c, learned != 0) + "\"";
    std::string hit;
    char buf[1024];
    while (fgets(buf, sizeof buf, f))
        if (strstr(buf, key.c_str())) hit = buf;  // the last entry wins
    fclose(f);
    if (hit.empty()) return SPG_OK;
    spg_config cfg = c->cfg;
    long long v;
    if (field(hit, "tile_log2", &v)) cfg.tile_log2 = (int32_t)v;
    if (field(hit, "long_row_ub", &v)) cfg.long_row_ub = v;
    if (field(hit, "item_products", &v)) cfg.item_products = (int32_t)v;
    if (field(hit, "expand_mode", &v)) cfg.expand_mode = (int32_t)v;
    if (field(hit, "heavy_factor", &v)) cfg.heavy_factor = (int32_t)v;
    if (field(hit, "rank_min_count", &v)) cfg.rank_min_count = (int32_t)v;
    SPG_TRY(spgSetConfig(c, &cfg));
    if (found) *found = 1;
    return SPG_OK;
}

}  // extern "C"
