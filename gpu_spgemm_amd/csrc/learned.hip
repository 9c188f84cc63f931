// learned.hip -- learned scheduler: a small MLP predicts, per C tile-row, its
// tile count and nonzero count, and the predicted nonzeros size C's value
// arrays so the product needs no host read between its symbolic and numeric
// steps.
//
// No reference counterpart: the reference sizes C with exact counts and two
// blocking reads, numtileC after Step 1 and nnzC after Step 2
// (src/spgemm_raw_v2.cu:119-141, :191-211).  Here:
//   1. the exact Step-1 count (k_row_stats) also emits each tile-row's
//      features; k_learned_cap evaluates the MLP (8 -> 32 -> 32 -> 2, ReLU,
//      FP32, weights in LDS) per tile-row and sums the predicted nonzeros,
//      each clamped to the row's exact bound (256 a C tile);
//   2. one host read returns numtileC, the bins and the predicted sum; C's
//      tile arrays are allocated exactly, its value arrays at the calibrated
//      prediction;
//   3. symbolic and numeric steps run back to back; the numeric kernels write
//      no value past the capacity and raise the overflow flag instead;
//   4. one read at the end returns nnz(C) and the flag; when the prediction
//      was short, only the numeric step runs again, into exactly sized arrays,
//      so the result never depends on the model.
// The network is trained on the GPU by the kernels below (forward, backward
// with per-block partial gradients reduced in a fixed order, Adam), on
// (features -> exact tile count, exact nonzeros) pairs from the exact pipeline.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "rowstaged.h"

namespace tsg {

constexpr int NF = 8, NH = 32, NO = 2;
constexpr int P_W1 = 0, P_B1 = P_W1 + NF * NH, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH, P_W3 = P_B2 + NH,
              P_B3 = P_W3 + NH * NO, NP = P_B3 + NO;
// targets: log1p(C tiles of the tile-row), log1p(C nonzeros of the tile-row)
constexpr int GRAD_BLOCK = 128;
// capacity = calibrated predicted nonzeros + this many
constexpr int64_t CAP_EXTRA = 1024;

struct Learned {
    float *w = nullptr, *m = nullptr, *v = nullptr;
    int64_t step = 0;
    bool init = false;
    // log-space offsets applied to the summed predictions (tiles, nonzeros):
    // calibrated after training so every training product's totals are covered
    float off_nc = 0.0198026f, off_nnz = 0.0198026f;  // log(1.02)
    float *X = nullptr, *T = nullptr;  // training set (device)
    int64_t n = 0, cap = 0;
    std::vector<int64_t> mat_end;      // end row of each collected product in the training set
    // The summed predictions of the last few products, keyed by their
    // operands: a repeat of the same operands reuses them instead of
    // evaluating the network again (a chain of products repeats several).
    // Only C's value capacity depends on them and the kernels check that
    // capacity, so a stale entry can cost a numeric re-run, never a wrong
    // result.
    struct Reuse {
        const void *a = nullptr, *b = nullptr;
        int64_t key[8] = {};
        int64_t pred_tiles = 0, pred_nnz = 0;
        // capacities for the next call of this pair (0 tiles: none yet): the
        // larger of the predictions and the sizes the last call found
        DeviceCaps caps;
        int nbins = 0;
        uint64_t used = 0;  // 0 = empty
    };
    static constexpr int NREUSE = 4;
    Reuse reuse[NREUSE];
    uint64_t tick = 0;
    void forget() {
        for (auto &r : reuse) r.used = 0;
    }
};

static Learned *state(Context *ctx) {
    if (!ctx->learned) ctx->learned = new Learned();
    return reinterpret_cast<Learned *>(ctx->learned);
}

void learned_release(Context *ctx) {
    Learned *L = reinterpret_cast<Learned *>(ctx->learned);
    if (!L) return;
    void *ps[] = {L->w, L->m, L->v, L->X, L->T};
    for (void *p : ps) pool_free(ctx, p);
    delete L;
    ctx->learned = nullptr;
}

// ---------------------------------------------------------------- MLP
// Every unit sums bias + inputs in ascending input order, one FMA each.  The
// loops run input-major so each input's weight row (32 consecutive floats in
// LDS) is read as vectors and every unit's partial sum stays in a register.
__device__ __forceinline__ void mlp_eval(const float *w, const float *x, float *h1, float *h2, float *y) {
#pragma unroll
    for (int j = 0; j < NH; j++) h1[j] = w[P_B1 + j];
#pragma unroll
    for (int i = 0; i < NF; i++) {
        const float xi = x[i];
#pragma unroll
        for (int j = 0; j < NH; j++) h1[j] = __builtin_fmaf(xi, w[P_W1 + i * NH + j], h1[j]);
    }
#pragma unroll
    for (int j = 0; j < NH; j++) {
        h1[j] = h1[j] > 0.0f ? h1[j] : 0.0f;
        h2[j] = w[P_B2 + j];
    }
#pragma unroll 4
    for (int i = 0; i < NH; i++) {
        const float hi = h1[i];
#pragma unroll
        for (int j = 0; j < NH; j++) h2[j] = __builtin_fmaf(hi, w[P_W2 + i * NH + j], h2[j]);
    }
#pragma unroll
    for (int j = 0; j < NH; j++) h2[j] = h2[j] > 0.0f ? h2[j] : 0.0f;
#pragma unroll
    for (int o = 0; o < NO; o++) y[o] = w[P_B3 + o];
#pragma unroll
    for (int i = 0; i < NH; i++)
#pragma unroll
        for (int o = 0; o < NO; o++) y[o] = __builtin_fmaf(h2[i], w[P_W3 + i * NO + o], y[o]);
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ wg, const float *__restrict__ X,
                                                     int64_t n, float *__restrict__ Y) {
    __shared__ float w[NP];
    for (int i = threadIdx.x; i < NP; i += 256) w[i] = wg[i];
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float x[NF], h1[NH], h2[NH], y[NO];
#pragma unroll
    for (int i = 0; i < NF; i++) x[i] = X[s * NF + i];
    mlp_eval(w, x, h1, h2, y);
#pragma unroll
    for (int o = 0; o < NO; o++) Y[s * NO + o] = y[o];
}

// Gradient of L = mean over the batch of sum_o (y_o - t_o)^2.  Each block
// holds 128 samples' activations and deltas in LDS and sums every parameter's
// gradient over them in sample order; partial[block][p] is reduced by
// k_mlp_adam / k_grad_reduce in block order, so results are reproducible.
__global__ void __launch_bounds__(GRAD_BLOCK)
k_mlp_grad(const float *__restrict__ wg, const float *__restrict__ X, const float *__restrict__ T,
           const int32_t *__restrict__ perm, int64_t off, int64_t nb, float *__restrict__ partial,
           float *__restrict__ loss_partial) {
    __shared__ float w[NP];
    __shared__ float sx[GRAD_BLOCK][NF];
    __shared__ float sh1[GRAD_BLOCK][NH + 1];
    __shared__ float sh2[GRAD_BLOCK][NH + 1];
    __shared__ float sd1[GRAD_BLOCK][NH + 1];
    __shared__ float sd2[GRAD_BLOCK][NH + 1];
    __shared__ float sd3[GRAD_BLOCK][NO];
    __shared__ float sloss[GRAD_BLOCK];
    const int tid = threadIdx.x;
    for (int i = tid; i < NP; i += GRAD_BLOCK) w[i] = wg[i];
    __syncthreads();
    const int64_t k = (int64_t)blockIdx.x * GRAD_BLOCK + tid;
    const bool live = k < nb;
    float x[NF], h1[NH], h2[NH], y[NO], t[NO];
    const int64_t src = live ? (perm ? perm[off + k] : off + k) : 0;
#pragma unroll
    for (int i = 0; i < NF; i++) x[i] = live ? X[src * NF + i] : 0.0f;
#pragma unroll
    for (int o = 0; o < NO; o++) t[o] = live ? T[src * NO + o] : 0.0f;
    mlp_eval(w, x, h1, h2, y);
    float d3[NO], loss = 0.0f;
    const float scale = 2.0f / (float)nb;
#pragma unroll
    for (int o = 0; o < NO; o++) {
        const float e = live ? y[o] - t[o] : 0.0f;
        d3[o] = e * scale;
        loss = __builtin_fmaf(e, e, loss);
    }
    float d2[NH];
    for (int i = 0; i < NH; i++) {
        float s = 0.0f;
#pragma unroll
        for (int o = 0; o < NO; o++) s = __builtin_fmaf(w[P_W3 + i * NO + o], d3[o], s);
        d2[i] = h2[i] > 0.0f ? s : 0.0f;
    }
    for (int i = 0; i < NH; i++) {
        float s = 0.0f;
        for (int j = 0; j < NH; j++) s = __builtin_fmaf(w[P_W2 + i * NH + j], d2[j], s);
        sd1[tid][i] = h1[i] > 0.0f ? s : 0.0f;
        sh1[tid][i] = h1[i];
        sh2[tid][i] = h2[i];
        sd2[tid][i] = d2[i];
    }
#pragma unroll
    for (int i = 0; i < NF; i++) sx[tid][i] = x[i];
#pragma unroll
    for (int o = 0; o < NO; o++) sd3[tid][o] = d3[o];
    sloss[tid] = loss;
    __syncthreads();
    for (int p = tid; p < NP; p += GRAD_BLOCK) {
        float g = 0.0f;
        if (p < P_B1) {
            const int i = (p - P_W1) / NH, j = (p - P_W1) % NH;
            for (int s = 0; s < GRAD_BLOCK; s++) g = __builtin_fmaf(sx[s][i], sd1[s][j], g);
        } else if (p < P_W2) {
            const int j = p - P_B1;
            for (int s = 0; s < GRAD_BLOCK; s++) g += sd1[s][j];
        } else if (p < P_B2) {
            const int i = (p - P_W2) / NH, j = (p - P_W2) % NH;
            for (int s = 0; s < GRAD_BLOCK; s++) g = __builtin_fmaf(sh1[s][i], sd2[s][j], g);
        } else if (p < P_W3) {
            const int j = p - P_B2;
            for (int s = 0; s < GRAD_BLOCK; s++) g += sd2[s][j];
        } else if (p < P_B3) {
            const int i = (p - P_W3) / NO, o = (p - P_W3) % NO;
            for (int s = 0; s < GRAD_BLOCK; s++) g = __builtin_fmaf(sh2[s][i], sd3[s][o], g);
        } else {
            const int o = p - P_B3;
            for (int s = 0; s < GRAD_BLOCK; s++) g += sd3[s][o];
        }
        partial[(int64_t)blockIdx.x * NP + p] = g;
    }
    if (tid == 0) {
        float l = 0.0f;
        for (int s = 0; s < GRAD_BLOCK; s++) l += sloss[s];
        loss_partial[blockIdx.x] = l;
    }
}

__global__ void __launch_bounds__(256) k_grad_reduce(const float *__restrict__ partial, int nblocks,
                                                     float *__restrict__ g) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; b++) s += partial[(int64_t)b * NP + p];
    g[p] = s;
}

__global__ void __launch_bounds__(256) k_mlp_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ partial, int nblocks, float lr,
                                                  float bc1, float bc2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP) return;
    float g = 0.0f;
    for (int b = 0; b < nblocks; b++) g += partial[(int64_t)b * NP + p];
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    w[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// ---------------------------------------------------------------- features and predictions
// Features of tile-row bi from the Step-1 statistics: log1p of (A tiles, B
// tiles reached, A nonzeros, B nonzeros reached, span of the reached B tile
// columns, longest reached B tile-row, summed spans of the reached B
// tile-rows, reached B tile-rows holding tiles).
__device__ __forceinline__ void row_features(const int32_t *__restrict__ st, float *x) {
    x[0] = log1pf((float)st[RS_NA]);
    x[1] = log1pf((float)st[RS_NB]);
    x[2] = log1pf((float)st[RS_PA]);
    x[3] = log1pf((float)st[RS_PB]);
    x[4] = log1pf((float)st[RS_SPAN]);
    x[5] = log1pf((float)st[RS_MAXLEN]);
    x[6] = log1pf((float)st[RS_SUMSPAN]);
    x[7] = log1pf((float)st[RS_NONEMPTY]);
}

// One thread per tile-row: features (Xout), and with weights the network's
// outputs (Yout: expm1, i.e. predicted tiles and nonzeros) and the sums of the
// per-row predictions, each rounded up and clamped to the row's exact bounds
// (tiles: B tiles reached and their column span; nonzeros: 256 per C tile).
__global__ void __launch_bounds__(256)
k_learned_cap(int32_t tilem, const int32_t *__restrict__ stats, const float *__restrict__ wg,
              float *__restrict__ Xout, float *__restrict__ Yout, unsigned long long *__restrict__ sum_tiles,
              unsigned long long *__restrict__ sum_nnz) {
    __shared__ float w[NP];
    __shared__ unsigned long long s_t[256 / WAVE], s_n[256 / WAVE];
    if (wg) {
        for (int i = threadIdx.x; i < NP; i += 256) w[i] = wg[i];
        __syncthreads();
    }
    const int32_t bi = blockIdx.x * 256 + threadIdx.x;
    unsigned long long pt = 0, pn = 0;
    if (bi < tilem) {
        const int32_t *st = stats + (int64_t)bi * RS_NSTAT;
        float x[NF];
        row_features(st, x);
        if (Xout)
#pragma unroll
            for (int i = 0; i < NF; i++) Xout[(int64_t)bi * NF + i] = x[i];
        if (wg) {
            float h1[NH], h2[NH], y[NO];
            mlp_eval(w, x, h1, h2, y);
            const float et = expm1f(fminf(y[0], 40.0f)), en = expm1f(fminf(y[1], 40.0f));
            if (Yout) {
                Yout[(int64_t)bi * NO + 0] = et;
                Yout[(int64_t)bi * NO + 1] = en;
            }
            const double tb = (double)min(st[RS_NB], st[RS_SPAN]);
            const double nbnd = 256.0 * (double)st[RS_COUNT];
            pt = (unsigned long long)fmin(ceil(fmax((double)et, 0.0)), tb);
            pn = (unsigned long long)fmin(ceil(fmax((double)en, 0.0)), nbnd);
        }
    }
    if (!wg) return;
    pt = wave_sum(pt);
    pn = wave_sum(pn);
    if (lane_id() == 0) {
        s_t[threadIdx.x / WAVE] = pt;
        s_n[threadIdx.x / WAVE] = pn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int q = 0; q < 256 / WAVE; q++) {
            a += s_t[q];
            b += s_n[q];
        }
        atomicAdd(sum_tiles, a);
        atomicAdd(sum_nnz, b);
    }
}

// Training targets from an exact C: tiles and nonzeros of every tile-row.
__global__ void __launch_bounds__(256) k_targets(int32_t tilem, const int64_t *__restrict__ tile_ptrC,
                                                 const int64_t *__restrict__ tile_nnzC, float *__restrict__ T) {
    const int32_t bi = blockIdx.x * 256 + threadIdx.x;
    if (bi >= tilem) return;
    const int64_t s = tile_ptrC[bi], e = tile_ptrC[bi + 1];
    T[(int64_t)bi * NO + 0] = log1pf((float)(e - s));
    T[(int64_t)bi * NO + 1] = log1pf((float)(tile_nnzC[e] - tile_nnzC[s]));
}

static unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// ---------------------------------------------------------------- the learned product
// Step-1 count with the per-row statistics, then the network's predicted
// nonzeros summed on the device (into p.flags).
static int learned_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                         const tsg_options *opts, Plan &p, int32_t *stats, const float *w, float *X, float *Y) {
    TSG_CHECK(plan_count(ctx, A, B, C, opts, p, stats));
    const int32_t tilem = A->tilem;
    if (tilem > 0) {
        k_learned_cap<<<nblk(tilem, 256), 256, 0, ctx->stream>>>(tilem, stats, w, X, Y, &p.flags->pred_tiles,
                                                                  &p.flags->pred_nnz);
        TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

// Capacities of a repeated product from a finished call: the larger of the
// predicted and the exact sizes, 2 % more, and each bin's rows (a bin that was
// empty stays closed: a row in it sends the call back to the exact pipeline).
static void remember_caps(Learned::Reuse &ru, const Plan &p, int64_t numtile, int64_t nnz, int64_t pred_tiles,
                          int64_t cap_nnz) {
    auto grow = [](int64_t v, int64_t extra) { return v + v / 50 + extra; };
    ru.caps.tiles = grow(std::max(numtile, pred_tiles), 64);
    ru.caps.nnz = grow(std::max(nnz, cap_nnz), 4096);
    ru.nbins = p.nbins;
    for (int b = 0; b < MAX_BINS; b++) {
        const int64_t rows = b < p.nbins ? p.bin_off[b + 1] - p.bin_off[b] : 0;
        ru.caps.bins[b] = rows > 0 ? grow(rows, 64) : 0;
    }
}

// A repeated product of a cached pair: C's arrays sized up front from the
// remembered capacities, every step queued with device-side sizes
// (spgemm_binned_device) and one host read at the end -- where the first call
// (and the exact pipeline) read numtileC and nnz(C) between the steps.  Past
// any capacity (a device flag) the result is discarded: *redo is set and the
// caller runs the exact pipeline.
static int learned_repeat(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                          const tsg_options *opts, const Learned::Reuse &ru, double scale, hipEvent_t *ev,
                          bool *redo, int *syncs) {
    *redo = false;
    DeviceCaps caps = ru.caps;  // scale < 1: the test hook that forces the exact re-run
    if (scale < 1.0) {
        caps.tiles = std::max<int64_t>(1, (int64_t)((double)caps.tiles * scale));
        caps.nnz = std::max<int64_t>(1, (int64_t)((double)caps.nnz * scale));
        for (auto &b : caps.bins) b = (int64_t)((double)b * scale);
    }
    Plan p;
    int64_t seq = 0;
    ctx->device_sizes = true;
    int rc = spgemm_binned_device(ctx, A, B, C, opts, caps, p, ev ? ev + 1 : nullptr, &seq);
    ctx->device_sizes = false;
    PlanFlags fl{};
    if (!rc) {  // the flags, published before the numeric step (which runs meanwhile)
        rc = publish_wait(ctx, seq, PLAN_FLAG_WORDS, ctx->pinned);
        (*syncs)++;
        memcpy(&fl, ctx->pinned, sizeof fl);
    }
    if (!rc && p.nbins != ru.nbins) rc = fail(TSG_ERR_ARG, "bin table changed between calls");
    free_plan(ctx, p);  // stream-ordered: the queued numeric kernels run before any reuse
    if (rc) {
        free_c(ctx, C);
        return rc;
    }
    if (fl.cap || fl.nnz > caps.nnz) {  // some capacity was short
        free_c(ctx, C);
        *redo = true;
        return TSG_OK;
    }
    C->numtile = (int64_t)fl.pred_tiles;
    C->nnz = fl.nnz;
    return TSG_OK;
}

int spgemm_learned(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *st) {
    Learned *L = state(ctx);
    if (!L->init) return fail(TSG_ERR_ARG, "learned scheduler has no model: train it or set its weights first");
    const bool timing = opts && opts->timing;
    // pred_scale < 1 shrinks the capacity (without the constant): a test hook
    // that forces the numeric re-run
    const double pscale = (opts && opts->pred_scale > 0) ? opts->pred_scale : 1.0;
    hipStream_t s = ctx->stream;
    memset(C, 0, sizeof *C);
    C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
    C->dtype = A->dtype; C->location = TSG_DEVICE;
    const int32_t tilem = A->tilem;
    hipEvent_t ev[4] = {};
    int rc = TSG_OK;
    if (timing)
        for (auto &e : ev)
            if (!rc && hipEventCreate(&e) != hipSuccess) rc = fail(TSG_ERR_HIP, "hipEventCreate failed");
    Plan p;
    int32_t *stats = nullptr;
    int syncs = 0, fallback = 0;
    int64_t cap = 0, pred_nnz = 0;
    PlanFlags fl{};
    const int64_t key[8] = {A->nnz, A->numtile, B->nnz, B->numtile, A->tilem, B->tilen, A->dtype, pscale == 1.0};
    Learned::Reuse *hit = nullptr, *slot = &L->reuse[0];
    for (auto &r : L->reuse) {
        if (r.used && r.a == A->tile_ptr && r.b == B->tile_ptr && !memcmp(r.key, key, sizeof key)) hit = &r;
        if (r.used < slot->used) slot = &r;  // empty or least recently used
    }
    const bool reuse = hit != nullptr;
    Learned::Reuse &ru = reuse ? *hit : *slot;
    ru.used = ++L->tick;
    if (timing && !rc) (void)hipEventRecord(ev[0], s);
    if (!rc && reuse && ru.caps.tiles > 0 && tilem > 0 && (!opts || opts->numeric_variant == 0) &&
        staged_ok(A, B, opts) && ru.caps.tiles <= INT32_MAX) {
        bool redo = false;
        rc = learned_repeat(ctx, A, B, C, opts, ru, pscale, timing ? ev : nullptr, &redo, &syncs);
        if (!rc && !redo) {
            if (timing) (void)hipEventRecord(ev[3], s);
            if (st) {
                st->numtileC = C->numtile;
                st->nnzC = C->nnz;
                st->host_syncs = syncs;
                st->fallback_used = 0;
                st->pred_numtileC = ru.caps.tiles;
                st->pred_nnzC = ru.caps.nnz;
                st->pred_reused = 1;
                st->numeric_kernels = 0;
            }
            if (timing) {
                if (st) {
                    float a = 0, b = 0, c3 = 0;
                    (void)hipEventSynchronize(ev[3]);
                    (void)hipEventElapsedTime(&a, ev[0], ev[1]);
                    (void)hipEventElapsedTime(&b, ev[1], ev[2]);
                    (void)hipEventElapsedTime(&c3, ev[2], ev[3]);
                    st->ms_step1 = a; st->ms_step2 = b; st->ms_step3 = c3; st->ms_total = a + b + c3;
                }
                for (auto &e : ev) (void)hipEventDestroy(e);
            }
            return TSG_OK;
        }
        if (!rc) {  // capacities were short: the exact pipeline, and its sizes remembered
            fallback = 1;
            ru.caps = DeviceCaps{};
            C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
            C->dtype = A->dtype; C->location = TSG_DEVICE;
        }
    }
    if (rc) {
        if (timing)
            for (auto &e : ev) if (e) (void)hipEventDestroy(e);
        return rc;
    }
    if (!rc) rc = palloc(ctx, &C->tile_ptr, (size_t)tilem + 1);
    if (reuse) {  // the plain Step-1 count; the predictions are this operand pair's last ones
        if (!rc) rc = plan_count(ctx, A, B, C, opts, p, nullptr, !plan_tail_fused(tilem));
        if (!rc) { rc = plan_tail(ctx, p, tilem, C->tile_ptr, nullptr, &C->numtile, nullptr); syncs++; }
        pred_nnz = ru.pred_nnz;
    } else {
        if (!rc) rc = palloc(ctx, &stats, (size_t)std::max(tilem, 1) * RS_NSTAT);
        if (!rc) rc = learned_count(ctx, A, B, C, opts, p, stats, L->w, nullptr, nullptr);
        if (!rc)
            rc = plan_bin(ctx, p, tilem, C->tile_ptr + tilem, reinterpret_cast<const int64_t *>(&p.flags->pred_nnz));
        if (!rc) { rc = plan_read(ctx, p, &C->numtile, &pred_nnz); syncs++; }
    }
    if (!rc && C->numtile > INT32_MAX) rc = fail(TSG_ERR_OVERFLOW, "numtileC exceeds int32 tile ids");
    if (!rc) rc = palloc(ctx, &C->tile_rowidx, (size_t)C->numtile);
    if (!rc) rc = palloc(ctx, &C->tile_columnidx, (size_t)C->numtile);
    if (!rc) rc = palloc(ctx, &C->mask, (size_t)C->numtile * 16);
    if (!rc) rc = palloc(ctx, &C->tile_csr_Ptr, (size_t)C->numtile * 16);
    if (!rc) rc = palloc(ctx, &C->tile_nnz, (size_t)C->numtile + 1);
    if (!rc) {
        // the calibrated prediction, never above the exact bound of 256 a tile
        const double want = std::ceil((double)pred_nnz * std::exp((double)L->off_nnz) * pscale) +
                            (pscale >= 1.0 ? (double)CAP_EXTRA : 0.0);
        cap = (int64_t)std::min(want, 256.0 * (double)C->numtile);
        if (pool_alloc(ctx, &C->tile_csr_Val, (size_t)cap * vsize(C->dtype)) != TSG_OK ||
            palloc(ctx, &C->tile_csr_Col, (size_t)cap) != TSG_OK) {
            // the prediction does not fit next to C's structure: values are
            // allocated once nnz(C) is known (the exact pipeline's order)
            pool_free(ctx, C->tile_csr_Val);
            C->tile_csr_Val = nullptr;
            cap = 0;
            fallback = 2;
        }
    }
    if (timing && !rc) (void)hipEventRecord(ev[1], s);
    SlotLayout sl;
    sl.overflow = p.overflow();
    sl.val_cap = cap;
    if (!rc) rc = run_symbolic(ctx, A, B, C, p, sl);
    free_heavy_bitmaps(ctx, p);  // the fill has used them
    if (!rc) rc = scan_inclusive_i64(ctx, C->tile_nnz, C->numtile + 1, true, &p.flags->nnz);
    if (!rc) rc = prepare_numeric(ctx, A, C, p, sl);
    if (timing && !rc) (void)hipEventRecord(ev[2], s);
    if (!rc && cap > 0) rc = run_numeric(ctx, A, B, C, p, sl);
    // nnz(C), the capacity flag and the predicted sums in one read
    if (!rc) {
        rc = read_words(ctx, p.flags, (int)(sizeof(PlanFlags) / sizeof(int32_t)), ctx->pinned);
        syncs++;
        memcpy(&fl, ctx->pinned, sizeof fl);
        C->nnz = fl.nnz;
        if (reuse) {
            fl.pred_tiles = (unsigned long long)ru.pred_tiles;
        } else if (!rc) {
            ru.a = A->tile_ptr;
            ru.b = B->tile_ptr;
            memcpy(ru.key, key, sizeof key);
            ru.pred_tiles = (int64_t)fl.pred_tiles;
            ru.pred_nnz = pred_nnz;
        }
    }
    if (!rc && fl.nnz <= cap && fl.cap) rc = fail(TSG_ERR_OVERFLOW, "a staged row exceeded its variant's capacity");
    if (!rc && fl.nnz > cap) {
        // the prediction was short: exact arrays, the values of every tile-row
        // (tile) that fit below the capacity copied over, the rest computed
        if (!fallback) fallback = 1;
        void *old_val = C->tile_csr_Val, *old_col = C->tile_csr_Col;
        C->tile_csr_Val = nullptr;
        C->tile_csr_Col = nullptr;
        rc = pool_alloc(ctx, &C->tile_csr_Val, (size_t)C->nnz * vsize(C->dtype));
        if (!rc) rc = palloc(ctx, &C->tile_csr_Col, (size_t)C->nnz);
        if (!rc && cap > 0 &&
            (hipMemcpyAsync(C->tile_csr_Val, old_val, (size_t)cap * vsize(C->dtype), hipMemcpyDeviceToDevice, s) !=
                 hipSuccess ||
             hipMemcpyAsync(C->tile_csr_Col, old_col, (size_t)cap, hipMemcpyDeviceToDevice, s) != hipSuccess))
            rc = fail(TSG_ERR_HIP, "copying the computed values failed");
        pool_free(ctx, old_val);  // stream-ordered reuse: the copies are queued first
        pool_free(ctx, old_col);
        sl.val_cap = INT64_MAX;
        sl.val_done = cap;
        if (!rc && C->nnz > 0) rc = run_numeric(ctx, A, B, C, p, sl);
    }
    if (!rc && fallback != 2)
        remember_caps(ru, p, C->numtile, C->nnz,
                      (int64_t)std::ceil((double)fl.pred_tiles * std::exp((double)L->off_nc)), cap);
    if (timing && !rc) (void)hipEventRecord(ev[3], s);
    if (st) {
        st->numtileC = C->numtile;
        st->nnzC = C->nnz;
        st->host_syncs = syncs;
        st->fallback_used = fallback;
        st->pred_numtileC = (int64_t)std::ceil((double)fl.pred_tiles * std::exp((double)L->off_nc));
        st->pred_nnzC = cap;
        st->pred_reused = reuse ? 1 : 0;
        st->numeric_kernels = C->nnz > 0 ? numeric_kernel_mask(p) : 0;
    }
    if (timing) {
        if (!rc && st) {
            float a = 0, b = 0, c3 = 0;
            (void)hipEventSynchronize(ev[3]);
            (void)hipEventElapsedTime(&a, ev[0], ev[1]);
            (void)hipEventElapsedTime(&b, ev[1], ev[2]);
            (void)hipEventElapsedTime(&c3, ev[2], ev[3]);
            st->ms_step1 = a; st->ms_step2 = b; st->ms_step3 = c3; st->ms_total = a + b + c3;
        }
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    pool_free(ctx, stats);
    free_plan(ctx, p);
    if (rc) free_c(ctx, C);
    return rc;
}
// ---------------------------------------------------------------- training
static int ensure_weights(Context *ctx, uint64_t seed) {
    Learned *L = state(ctx);
    if (L->w) return TSG_OK;
    TSG_CHECK(palloc(ctx, &L->w, NP));
    TSG_CHECK(palloc(ctx, &L->m, NP));
    TSG_CHECK(palloc(ctx, &L->v, NP));
    std::vector<float> h(NP, 0.0f);
    std::mt19937_64 rng(seed);
    auto init = [&](int off, int fan_in, int fan_out) {  // Glorot uniform
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = 0; i < fan_in * fan_out; i++) h[off + i] = u(rng);
    };
    init(P_W1, NF, NH);
    init(P_W2, NH, NH);
    init(P_W3, NH, NO);
    TSG_HIP(hipMemcpyAsync(L->w, h.data(), NP * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    TSG_HIP(hipMemsetAsync(L->m, 0, NP * sizeof(float), ctx->stream));
    TSG_HIP(hipMemsetAsync(L->v, 0, NP * sizeof(float), ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    L->step = 0;
    return TSG_OK;
}

// Log-space offsets that make the summed prediction of every collected
// product cover that product's exact totals (C tiles, C nonzeros): off = max
// over products of log(actual / predicted) + log 1.02, clamped to
// [log 1.02, log 16].  Each row's prediction is rounded up and clamped to its
// exact bounds as k_learned_cap does, so the offsets apply to the device sums.
static int calibrate(Context *ctx) {
    Learned *L = state(ctx);
    const int64_t n = L->n;
    float *Y = nullptr;
    TSG_CHECK(palloc(ctx, &Y, (size_t)n * NO));
    std::vector<float> hy((size_t)n * NO), ht((size_t)n * NO), hx((size_t)n * NF);
    k_mlp_forward<<<nblk(n, 256), 256, 0, ctx->stream>>>(L->w, L->X, n, Y);
    int rc = TSG_OK;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hy.data(), Y, sizeof(float) * n * NO, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(ht.data(), L->T, sizeof(float) * n * NO, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(hx.data(), L->X, sizeof(float) * n * NF, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "calibration failed");
    pool_free(ctx, Y);
    if (rc) return rc;
    double off[NO] = {-1e30, -1e30};
    int64_t r0 = 0;
    for (const int64_t r1 : L->mat_end) {
        double pred[NO] = {0, 0}, act[NO] = {0, 0};
        for (int64_t i = r0; i < r1; i++) {
            const double cnt = std::rint(std::expm1((double)ht[i * NO + 0]));
            const double nnz = std::rint(std::expm1((double)ht[i * NO + 1]));
            const double nb = std::rint(std::expm1((double)hx[i * NF + 1]));
            const double span = std::rint(std::expm1((double)hx[i * NF + 4]));
            const double et = std::expm1((double)std::min(hy[i * NO + 0], 40.0f));
            const double en = std::expm1((double)std::min(hy[i * NO + 1], 40.0f));
            pred[0] += std::min(std::ceil(std::max(et, 0.0)), std::min(nb, span));
            pred[1] += std::min(std::ceil(std::max(en, 0.0)), 256.0 * cnt);
            act[0] += cnt;
            act[1] += nnz;
        }
        for (int o = 0; o < NO; o++)
            if (act[o] > 0) off[o] = std::max(off[o], std::log(act[o] / std::max(pred[o], 1.0)));
        r0 = r1;
    }
    const double margin = std::log(1.02), hi = std::log(16.0);
    L->off_nc = (float)std::min(std::max(off[0] + margin, margin), hi);
    L->off_nnz = (float)std::min(std::max(off[1] + margin, margin), hi);
    return TSG_OK;
}

static int append_rows(Context *ctx, const float *X, const float *T, int64_t n) {
    Learned *L = state(ctx);
    if (L->n + n > L->cap) {
        int64_t nc = std::max<int64_t>(L->cap * 2, L->n + n);
        float *nx = nullptr, *nt = nullptr;
        TSG_CHECK(palloc(ctx, &nx, (size_t)nc * NF));
        TSG_CHECK(palloc(ctx, &nt, (size_t)nc * NO));
        if (L->n) {
            TSG_HIP(hipMemcpyAsync(nx, L->X, sizeof(float) * L->n * NF, hipMemcpyDeviceToDevice, ctx->stream));
            TSG_HIP(hipMemcpyAsync(nt, L->T, sizeof(float) * L->n * NO, hipMemcpyDeviceToDevice, ctx->stream));
        }
        pool_free(ctx, L->X);
        pool_free(ctx, L->T);
        L->X = nx;
        L->T = nt;
        L->cap = nc;
    }
    TSG_HIP(hipMemcpyAsync(L->X + L->n * NF, X, sizeof(float) * n * NF, hipMemcpyDeviceToDevice, ctx->stream));
    TSG_HIP(hipMemcpyAsync(L->T + L->n * NO, T, sizeof(float) * n * NO, hipMemcpyDeviceToDevice, ctx->stream));
    L->n += n;
    return TSG_OK;
}

}  // namespace tsg

using namespace tsg;

extern "C" {

int tsg_learned_num_params(void) { return NP; }

// Features of every tile-row of A*B, or the MLP's predictions (pred != NULL
// needs a model): host arrays of tilem*8 / tilem*2 floats.
static int features_or_predict(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B, float *feats,
                               float *pred) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !A || !B) return fail(TSG_ERR_ARG, "null argument");
    TSG_CHECK(check_operands(A, B));
    (void)hipSetDevice(ctx->device);
    Learned *L = state(ctx);
    if (pred && !L->init) return fail(TSG_ERR_ARG, "learned scheduler has no model");
    const int32_t tilem = A->tilem;
    Plan p;
    tsg_tile_matrix Ct{};
    float *X = nullptr, *Y = nullptr;
    int32_t *stats = nullptr;
    const size_t rows = (size_t)std::max(tilem, 1);
    int rc = palloc(ctx, &Ct.tile_ptr, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &stats, rows * RS_NSTAT);
    if (!rc) rc = palloc(ctx, &X, rows * NF);
    if (!rc) rc = palloc(ctx, &Y, rows * NO);
    if (!rc) rc = learned_count(ctx, A, B, &Ct, nullptr, p, stats, pred ? L->w : nullptr, X, pred ? Y : nullptr);
    if (!rc && feats && tilem > 0 &&
        hipMemcpyAsync(feats, X, sizeof(float) * tilem * NF, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc && pred && tilem > 0 &&
        hipMemcpyAsync(pred, Y, sizeof(float) * tilem * NO, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "sync");
    void *ps[] = {Ct.tile_ptr, stats, X, Y};
    for (void *q : ps) pool_free(ctx, q);
    free_plan(ctx, p);
    return rc;
}

int tsg_learned_features(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B, float *feats) {
    return features_or_predict(c, A, B, feats, nullptr);
}

int tsg_learned_predict(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B, float *pred) {
    return features_or_predict(c, A, B, nullptr, pred);
}

int tsg_learned_collect(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !A || !B) return fail(TSG_ERR_ARG, "null argument");
    TSG_CHECK(check_operands(A, B));
    (void)hipSetDevice(ctx->device);
    const int32_t tilem = A->tilem;
    if (tilem == 0) return TSG_OK;
    tsg_tile_matrix C;
    tsg_options o{};
    TSG_CHECK(spgemm_binned(ctx, A, B, &C, &o, nullptr));
    Plan p;
    tsg_tile_matrix Ct{};
    float *X = nullptr, *T = nullptr;
    int32_t *stats = nullptr;
    int rc = palloc(ctx, &Ct.tile_ptr, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &stats, (size_t)tilem * RS_NSTAT);
    if (!rc) rc = palloc(ctx, &X, (size_t)tilem * NF);
    if (!rc) rc = palloc(ctx, &T, (size_t)tilem * NO);
    if (!rc) rc = learned_count(ctx, A, B, &Ct, &o, p, stats, nullptr, X, nullptr);
    if (!rc) {
        k_targets<<<nblk(tilem, 256), 256, 0, ctx->stream>>>(tilem, C.tile_ptr, C.tile_nnz, T);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_targets launch");
    }
    Learned *L = state(ctx);
    if (!rc) rc = append_rows(ctx, X, T, tilem);
    if (!rc) L->mat_end.push_back(L->n);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "sync");
    void *ps[] = {Ct.tile_ptr, stats, X, T};
    for (void *q : ps) pool_free(ctx, q);
    free_plan(ctx, p);
    free_c(ctx, &C);
    return rc;
}

int tsg_learned_clear(tsg_context *c) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    Learned *L = state(ctx);
    L->n = 0;
    L->mat_end.clear();
    L->forget();
    return TSG_OK;
}

int tsg_learned_train(tsg_context *c, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *final_loss) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || epochs < 0 || batch <= 0 || !(lr > 0)) return fail(TSG_ERR_ARG, "bad training arguments");
    (void)hipSetDevice(ctx->device);
    Learned *L = state(ctx);
    if (L->n == 0) return fail(TSG_ERR_ARG, "empty training set: call tsg_learned_collect first");
    TSG_CHECK(ensure_weights(ctx, seed));
    const int64_t n = L->n;
    const int nblocks_max = (int)((batch + GRAD_BLOCK - 1) / GRAD_BLOCK);
    float *partial = nullptr, *lossp = nullptr;
    int32_t *perm = nullptr;
    TSG_CHECK(palloc(ctx, &partial, (size_t)nblocks_max * NP));
    TSG_CHECK(palloc(ctx, &lossp, (size_t)nblocks_max));
    TSG_CHECK(palloc(ctx, &perm, (size_t)n));
    std::vector<int32_t> order((size_t)n);
    std::mt19937_64 rng(seed ^ 0x5DEECE66Dull);
    double epoch_loss = 0;
    std::vector<float> hl((size_t)nblocks_max);
    int rc = TSG_OK;
    for (int e = 0; e < epochs && !rc; e++) {
        for (int64_t i = 0; i < n; i++) order[i] = (int32_t)i;
        std::shuffle(order.begin(), order.end(), rng);
        // stream-ordered after the previous epoch's kernels; synchronous so
        // `order` can be reshuffled next epoch
        if (hipMemcpyAsync(perm, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = fail(TSG_ERR_HIP, "perm upload");
            break;
        }
        epoch_loss = 0;
        for (int64_t off = 0; off < n && !rc; off += batch) {
            const int64_t nb = std::min<int64_t>(batch, n - off);
            const int nblocks = (int)((nb + GRAD_BLOCK - 1) / GRAD_BLOCK);
            k_mlp_grad<<<nblocks, GRAD_BLOCK, 0, ctx->stream>>>(L->w, L->X, L->T, perm, off, nb, partial, lossp);
            L->step++;
            const float bc1 = 1.0f - std::pow(0.9f, (float)L->step), bc2 = 1.0f - std::pow(0.999f, (float)L->step);
            k_mlp_adam<<<nblk(NP, 256), 256, 0, ctx->stream>>>(L->w, L->m, L->v, partial, nblocks, lr, bc1, bc2);
            if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "training launch");
            if (!rc && e == epochs - 1) {
                if (hipMemcpyAsync(hl.data(), lossp, sizeof(float) * nblocks, hipMemcpyDeviceToHost, ctx->stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(ctx->stream) != hipSuccess) {
                    rc = fail(TSG_ERR_HIP, "loss read");
                    break;
                }
                for (int b = 0; b < nblocks; b++) epoch_loss += hl[b];
            }
        }
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "sync");
    pool_free(ctx, partial);
    pool_free(ctx, lossp);
    pool_free(ctx, perm);
    if (rc) return rc;
    TSG_CHECK(calibrate(ctx));
    L->init = true;
    L->forget();
    if (final_loss) *final_loss = (float)(epoch_loss / (double)n);
    return TSG_OK;
}

int tsg_learned_calibration(tsg_context *c, float *get, const float *set) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    Learned *L = state(ctx);
    if (set) {
        if (!(set[0] >= 0.0f && set[0] <= 20.0f && set[1] >= 0.0f && set[1] <= 20.0f))
            return fail(TSG_ERR_ARG, "calibration offsets must lie in [0, 20]");
        L->off_nc = set[0];
        L->off_nnz = set[1];
        L->forget();
        }
    if (get) {
        get[0] = L->off_nc;
        get[1] = L->off_nnz;
    }
    return TSG_OK;
}

int tsg_learned_get_weights(tsg_context *c, float *w, int32_t n) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !w || n != NP) return fail(TSG_ERR_ARG, "need %d weights", NP);
    Learned *L = state(ctx);
    if (!L->w) return fail(TSG_ERR_ARG, "no weights yet");
    (void)hipSetDevice(ctx->device);
    TSG_HIP(hipMemcpyAsync(w, L->w, sizeof(float) * NP, hipMemcpyDeviceToHost, ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}

int tsg_learned_set_weights(tsg_context *c, const float *w, int32_t n) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !w || n != NP) return fail(TSG_ERR_ARG, "need %d weights", NP);
    (void)hipSetDevice(ctx->device);
    TSG_CHECK(ensure_weights(ctx, 0));
    Learned *L = state(ctx);
    TSG_HIP(hipMemcpyAsync(L->w, w, sizeof(float) * NP, hipMemcpyHostToDevice, ctx->stream));
    TSG_HIP(hipMemsetAsync(L->m, 0, NP * sizeof(float), ctx->stream));
    TSG_HIP(hipMemsetAsync(L->v, 0, NP * sizeof(float), ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    L->step = 0;
    L->init = true;
    L->forget();
    return TSG_OK;
}

// MLP forward and full-batch gradient on host arrays (the GPU kernels the
// scheduler and the trainer use; exposed so tests can check them).
int tsg_learned_forward(tsg_context *c, const float *X, int64_t n, float *Y) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !X || !Y || n < 0) return fail(TSG_ERR_ARG, "bad arguments");
    Learned *L = state(ctx);
    if (!L->w) return fail(TSG_ERR_ARG, "no weights yet");
    (void)hipSetDevice(ctx->device);
    float *dx = nullptr, *dy = nullptr;
    TSG_CHECK(palloc(ctx, &dx, (size_t)std::max<int64_t>(n, 1) * NF));
    TSG_CHECK(palloc(ctx, &dy, (size_t)std::max<int64_t>(n, 1) * NO));
    int rc = TSG_OK;
    if (n > 0) {
        if (hipMemcpyAsync(dx, X, sizeof(float) * n * NF, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = fail(TSG_ERR_HIP, "copy");
        if (!rc) {
            k_mlp_forward<<<nblk(n, 256), 256, 0, ctx->stream>>>(L->w, dx, n, dy);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(Y, dy, sizeof(float) * n * NO, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                rc = fail(TSG_ERR_HIP, "forward failed");
        }
    }
    pool_free(ctx, dx);
    pool_free(ctx, dy);
    return rc;
}

int tsg_learned_gradient(tsg_context *c, const float *X, const float *T, int64_t n, float *grad, float *loss) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !X || !T || !grad || n <= 0) return fail(TSG_ERR_ARG, "bad arguments");
    Learned *L = state(ctx);
    if (!L->w) return fail(TSG_ERR_ARG, "no weights yet");
    (void)hipSetDevice(ctx->device);
    const int nblocks = (int)((n + GRAD_BLOCK - 1) / GRAD_BLOCK);
    float *dx = nullptr, *dt = nullptr, *partial = nullptr, *lossp = nullptr, *g = nullptr;
    int rc = palloc(ctx, &dx, (size_t)n * NF);
    if (!rc) rc = palloc(ctx, &dt, (size_t)n * NO);
    if (!rc) rc = palloc(ctx, &partial, (size_t)nblocks * NP);
    if (!rc) rc = palloc(ctx, &lossp, (size_t)nblocks);
    if (!rc) rc = palloc(ctx, &g, NP);
    if (!rc && (hipMemcpyAsync(dx, X, sizeof(float) * n * NF, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
                hipMemcpyAsync(dt, T, sizeof(float) * n * NO, hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc) {
        k_mlp_grad<<<nblocks, GRAD_BLOCK, 0, ctx->stream>>>(L->w, dx, dt, nullptr, 0, n, partial, lossp);
        k_grad_reduce<<<nblk(NP, 256), 256, 0, ctx->stream>>>(partial, nblocks, g);
        std::vector<float> hl((size_t)nblocks);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(grad, g, sizeof(float) * NP, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(hl.data(), lossp, sizeof(float) * nblocks, hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(TSG_ERR_HIP, "gradient failed");
        if (!rc && loss) {
            double l = 0;
            for (int b = 0; b < nblocks; b++) l += hl[b];
            *loss = (float)(l / (double)n);
        }
    }
    void *ps[] = {dx, dt, partial, lossp, g};
    for (void *q : ps) pool_free(ctx, q);
    return rc;
}


}  // extern "C"
