// autotune.hip -- picks the kernel-variant configuration of a class of products
// by timing every candidate.  No reference counterpart (the reference has one
// fixed kernel per step).
//
// A configuration = which row-staged variants may take rows (the rest fall to
// the heavy-row kernels) and the A tile density at or below which heavy rows
// run row-by-row instead of through the C-tile kernels.  Every configuration
// computes identical bits (each C entry sums its products in the same order
// in every kernel), so tuning changes speed only.  The winner is stored per
// operand class -- log2 buckets of the sizes that decide which kernel suits a
// row -- and used by later products of that class.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <vector>

#include "rowstaged.h"

namespace tsg {

struct Tuner {
    std::map<uint64_t, TuneCfg> best;
    bool forcing = false;
    TuneCfg forced;
};

static Tuner *tuner(Context *ctx) {
    if (!ctx->tuner) ctx->tuner = new Tuner();
    return reinterpret_cast<Tuner *>(ctx->tuner);
}

void tuner_release(Context *ctx) {
    delete reinterpret_cast<Tuner *>(ctx->tuner);
    ctx->tuner = nullptr;
}

constexpr int32_t TUNE_SAMPLE_MIN = 16;  // batches this large are sampled first
constexpr int32_t TUNE_STRIDE = 7;
constexpr int TUNE_FINALISTS = 3;

static uint64_t lb(double x) { return x <= 1.0 ? 0 : (uint64_t)std::min(63.0, std::floor(std::log2(x))); }

uint64_t tune_signature(int64_t tilemA, int64_t numtileA, int64_t nnzA, int64_t tilemB, int64_t numtileB,
                        int32_t dtype) {
    const double ta = tilemA > 0 ? (double)numtileA / tilemA : 0.0;   // A tiles per tile-row
    const double da = numtileA > 0 ? (double)nnzA / numtileA : 0.0;   // A entries per tile
    const double tb = tilemB > 0 ? (double)numtileB / tilemB : 0.0;   // B tiles per tile-row
    return (lb((double)tilemA) << 32) | (lb(ta * 4) << 24) | (lb(da * 4) << 16) | (lb(tb * 4) << 8) |
           (uint64_t)(dtype & 0xFF);
}

bool tuned_config(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts,
                  TuneCfg *cfg) {
    *cfg = TuneCfg{};
    Tuner *t = reinterpret_cast<Tuner *>(ctx->tuner);
    if (!t) return false;
    if (t->forcing) {
        *cfg = t->forced;
        return true;
    }
    if (opts && opts->tuning == 1) return false;
    auto it = t->best.find(tune_signature(A->tilem, A->numtile, A->nnz, B->tilem, B->numtile, A->dtype));
    if (it == t->best.end()) return false;
    *cfg = it->second;
    return true;
}


// Signature of the product a call with these operands runs (for a batch:
// the block-diagonal product's totals).
static uint64_t call_signature(int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B) {
    int64_t tm = 0, nt = 0, nz = 0, tmb = 0, ntb = 0;
    for (int32_t i = 0; i < count; i++) {
        tm += A[i].tilem; nt += A[i].numtile; nz += A[i].nnz; tmb += B[i].tilem; ntb += B[i].numtile;
    }
    return tune_signature(tm, nt, nz, tmb, ntb, A[0].dtype);
}

// One timed run: the product (or batch) and the freeing of its outputs.
static int run_once(Context *ctx, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B, double *ms) {
    std::vector<tsg_tile_matrix> C((size_t)count);
    tsg_options o{};
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    int rc = count == 1 ? spgemm_dispatch(ctx, A, B, C.data(), &o, nullptr)
                        : spgemm_batched(ctx, count, A, B, C.data(), &o, nullptr);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "sync");
    const auto t1 = std::chrono::steady_clock::now();
    for (auto &c : C) free_c(ctx, &c);
    *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return rc;
}

}  // namespace tsg

using namespace tsg;

extern "C" {

int tsg_autotune(tsg_context *c, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B, int32_t reps,
                 tsg_tune_entry *report, int32_t cap, int32_t *nreport, tsg_tune_entry *best_out) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || count < 1 || !A || !B || reps < 1) return fail(TSG_ERR_ARG, "bad autotune arguments");
    (void)hipSetDevice(ctx->device);
    for (int32_t i = 0; i < count; i++) TSG_CHECK(check_operands(&A[i], &B[i]));
    bool staged = true, rg = true;
    for (int32_t i = 0; i < count; i++) {
        staged = staged && staged_ok(&A[i], &B[i], nullptr);
        rg = rg && rg_ok(&A[i], &B[i], nullptr);
    }
    bool dups = false;
    for (int32_t i = 0; i < count; i++) dups = dups || ((A[i].flags | B[i].flags) & 1);
    std::vector<int> masks = staged ? std::vector<int>{15, 7, 13, 5, 1, 4, 0} : std::vector<int>{0};
    std::vector<int> dens = rg ? std::vector<int>{0, 8, 64, 1024} : std::vector<int>{0};
    std::vector<TuneCfg> cands;
    for (int mk : masks)
        for (int d : dens) {
            TuneCfg c;
            c.var_mask = mk;
            c.rg_density = d;
            cands.push_back(c);
        }
    if (!dups)
        for (int mp : {64, 512}) {  // matrix-core C-tile kernels
            TuneCfg c;
            c.var_mask = 0;
            c.rg_density = 0;
            c.mfma = mp;
            cands.push_back(c);
        }
    Tuner *t = tuner(ctx);
    std::vector<tsg_tune_entry> rep;
    int rc = TSG_OK;
    // A batch of TUNE_SAMPLE_MIN or more products is first timed on a sample
    // (every TUNE_STRIDE-th product: an odd stride, so alternating problem
    // kinds both appear; one warm-up and one timed run); only the
    // TUNE_FINALISTS fastest there are timed on the whole batch.
    std::vector<tsg_tile_matrix> SA, SB;
    if (count >= TUNE_SAMPLE_MIN)
        for (int32_t i = 0; i < count; i += TUNE_STRIDE) {
            SA.push_back(A[i]);
            SB.push_back(B[i]);
        }
    std::vector<char> full(cands.size(), 1);
    std::vector<double> sample_ms(cands.size(), 0.0);
    if (!SA.empty()) {
        for (size_t c = 0; c < cands.size() && !rc; c++) {
            t->forcing = true;
            t->forced = cands[c];
            double w = 0;
            rc = run_once(ctx, (int32_t)SA.size(), SA.data(), SB.data(), &w);
            if (!rc) rc = run_once(ctx, (int32_t)SA.size(), SA.data(), SB.data(), &w);
            t->forcing = false;
            sample_ms[c] = w;
        }
        if (rc) return rc;
        std::vector<size_t> order(cands.size());
        for (size_t c = 0; c < order.size(); c++) order[c] = c;
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return sample_ms[x] < sample_ms[y]; });
        for (size_t r = 0; r < order.size(); r++) full[order[r]] = r < (size_t)TUNE_FINALISTS;
    }
    for (size_t c = 0; c < cands.size(); c++) {
        const TuneCfg &cand = cands[c];
        if (!full[c]) {
            rep.push_back(tsg_tune_entry{cand.var_mask, cand.rg_density, cand.mfma, 1, sample_ms[c]});
            continue;
        }
        t->forcing = true;
        t->forced = cand;
        double w = 0;
        rc = run_once(ctx, count, A, B, &w);  // warm-up (pool sizes, caches)
        std::vector<double> ms;
        for (int r = 0; r < reps && !rc; r++) {
            rc = run_once(ctx, count, A, B, &w);
            ms.push_back(w);
        }
        t->forcing = false;
        if (rc) return rc;
        std::sort(ms.begin(), ms.end());
        rep.push_back(tsg_tune_entry{cand.var_mask, cand.rg_density, cand.mfma, 0, ms[ms.size() / 2]});
    }
    const auto it = std::min_element(rep.begin(), rep.end(), [](const tsg_tune_entry &x, const tsg_tune_entry &y) {
        return x.sampled != y.sampled ? x.sampled < y.sampled : x.ms < y.ms;  // fully timed first
    });
    TuneCfg cfg;
    cfg.var_mask = it->var_mask;
    cfg.rg_density = it->rg_density;
    cfg.mfma = it->mfma_min_products;
    t->best[call_signature(count, A, B)] = cfg;
    if (best_out) *best_out = *it;
    if (nreport) *nreport = (int32_t)rep.size();
    if (report)
        for (int32_t i = 0; i < std::min<int32_t>(cap, (int32_t)rep.size()); i++) report[i] = rep[i];
    return TSG_OK;
}

int tsg_autotune_query(tsg_context *c, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                       tsg_tune_entry *cfg) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || count < 1 || !A || !B) return 0;
    Tuner *t = reinterpret_cast<Tuner *>(ctx->tuner);
    if (!t) return 0;
    auto it = t->best.find(call_signature(count, A, B));
    if (it == t->best.end()) return 0;
    if (cfg) *cfg = tsg_tune_entry{it->second.var_mask, it->second.rg_density, it->second.mfma, 0, 0.0};
    return 1;
}

int tsg_autotune_set(tsg_context *c, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                     const tsg_tune_entry *cfg) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || count < 1 || !A || !B || !cfg) return fail(TSG_ERR_ARG, "bad arguments");
    if (cfg->var_mask < 0 || cfg->var_mask > 31 || cfg->rg_density < 0 || cfg->mfma_min_products < 0)
        return fail(TSG_ERR_ARG, "var_mask must lie in [0, 31], rg_density and mfma_min_products be >= 0");
    TuneCfg t;
    t.var_mask = cfg->var_mask;
    t.rg_density = cfg->rg_density;
    t.mfma = cfg->mfma_min_products;
    tuner(ctx)->best[call_signature(count, A, B)] = t;
    return TSG_OK;
}

int tsg_autotune_clear(tsg_context *c) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    if (ctx->tuner) reinterpret_cast<Tuner *>(ctx->tuner)->best.clear();
    return TSG_OK;
}

}  // extern "C"
