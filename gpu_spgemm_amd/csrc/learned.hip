// learned.hip -- learned scheduler: a small MLP predicts, per C tile-row, its
// tile count and nonzero count from cheap features, and the product runs
// without the Step-1 count pass.
//
// No reference counterpart: the reference sizes C with an exact count kernel
// and host reads (src/spgemm_raw_v2.cu:119-141, :191-211).  Here:
//   1. one thread per tile-row computes features from A's tile-row and the
//      B tile-rows it reaches (tiles, payloads, column spans), evaluates the
//      MLP (8 -> 32 -> 32 -> 2, ReLU, FP32, weights in LDS) and turns the
//      predicted tile count into a slot capacity (prediction + margin, never
//      above the B tiles reached, an exact upper bound) and a bin;
//   2. C's structure is written into per-row slots (slot layout), each row's
//      actual tile count kept in row_nc; a row that needs more slots than
//      predicted, or values beyond the predicted total, raises an overflow
//      flag instead of writing out of bounds;
//   3. value offsets come from scanning the slot counts (unused slots count
//      0), the exact tile_ptr from scanning row_nc, and a compaction kernel
//      moves the structure to the compact layout;
//   4. one host read at the end returns (overflow, numtileC, nnzC); on
//      overflow the exact binned pipeline runs instead, so the result never
//      depends on the model.
// The network is trained on the GPU by the kernels below (forward, backward
// with per-block partial gradients reduced in a fixed order, Adam), on
// (features -> exact tile count, exact nnz) pairs that the exact pipeline
// emits.
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

struct Learned {
    float *w = nullptr, *m = nullptr, *v = nullptr;
    int64_t step = 0;
    bool init = false;
    // log-space offsets added to the predictions before sizing (calibrated on
    // the training set after each training run)
    float off_nc = 0.0487902f, off_nnz = 0.0487902f;  // log(1.05)
    float *X = nullptr, *T = nullptr;  // training set (device)
    int64_t n = 0, cap = 0;
    struct Cache {
        const void *a = nullptr, *b = nullptr;
        int64_t annz = -1, bnnz = -1, atiles = -1, btiles = -1, tilem = -1, tilen = -1;
        int64_t slots = 0, valcap = 0;
        int64_t bin_off[MAX_BINS + 1] = {};
        int nbins = 0;
        bool valid = false;
    } cache;
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
__device__ __forceinline__ void mlp_eval(const float *w, const float *x, float *h1, float *h2, float *y) {
#pragma unroll 4
    for (int j = 0; j < NH; j++) {
        float s = w[P_B1 + j];
#pragma unroll
        for (int i = 0; i < NF; i++) s = __builtin_fmaf(x[i], w[P_W1 + i * NH + j], s);
        h1[j] = s > 0.0f ? s : 0.0f;
    }
#pragma unroll 4
    for (int j = 0; j < NH; j++) {
        float s = w[P_B2 + j];
        for (int i = 0; i < NH; i++) s = __builtin_fmaf(h1[i], w[P_W2 + i * NH + j], s);
        h2[j] = s > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int o = 0; o < NO; o++) {
        float s = w[P_B3 + o];
        for (int i = 0; i < NH; i++) s = __builtin_fmaf(h2[i], w[P_W3 + i * NO + o], s);
        y[o] = s;
    }
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

// ---------------------------------------------------------------- features, plan, targets
// One thread per tile-row: features; with weights, the MLP's prediction,
// slot capacity, value capacity and bin.
__global__ void __launch_bounds__(256)
k_learned_plan(int32_t tilem, const int64_t *__restrict__ tile_ptrA, const int64_t *__restrict__ tile_nnzA,
               const int4 *__restrict__ link, const int32_t *__restrict__ colidxB, const float *__restrict__ wg,
               BinTable bt, float off_nc, float off_nnz, float scale, int extra, float *__restrict__ Xout,
               float *__restrict__ Yout,
               int64_t *__restrict__ slots, int64_t *__restrict__ valcap, uint8_t *__restrict__ bin) {
    __shared__ float w[NP];
    if (wg) {
        for (int i = threadIdx.x; i < NP; i += 256) w[i] = wg[i];
        __syncthreads();
    }
    const int32_t bi = blockIdx.x * 256 + threadIdx.x;
    if (bi >= tilem) return;
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];
    const int64_t na = ae - as, pa = tile_nnzA[ae] - tile_nnzA[as];
    int64_t nb = 0, pb = 0, sumspan = 0, nonempty = 0, maxlen = 0;
    int32_t cmin = INT_MAX, cmax = -1;
    for (int64_t t = as; t < ae; t++) {
        const int4 l = link[t];
        nb += l.y;
        pb += l.z;
        if (l.y > 0) {
            const int32_t f = colidxB[l.x], e = colidxB[l.x + l.y - 1];
            cmin = min(cmin, f);
            cmax = max(cmax, e);
            sumspan += e - f + 1;
            nonempty++;
            maxlen = max<int64_t>(maxlen, l.y);
        }
    }
    const int64_t span = cmax >= cmin ? (int64_t)cmax - cmin + 1 : 0;
    float x[NF] = {log1pf((float)na), log1pf((float)nb), log1pf((float)pa), log1pf((float)pb),
                   log1pf((float)span), log1pf((float)maxlen), log1pf((float)sumspan), log1pf((float)nonempty)};
    if (Xout)
#pragma unroll
        for (int i = 0; i < NF; i++) Xout[(int64_t)bi * NF + i] = x[i];
    if (!wg) return;
    float h1[NH], h2[NH], y[NO];
    mlp_eval(w, x, h1, h2, y);
    if (Yout) {
        Yout[(int64_t)bi * NO + 0] = expm1f(y[0]);
        Yout[(int64_t)bi * NO + 1] = expm1f(y[1]);
    }
    // capacities: expm1(prediction + calibrated offset) * scale (+ a small
    // constant), never above the exact upper bounds (distinct B tile columns
    // reached; 256 values a tile)
    const float pnc = fmaxf(expm1f(fminf(y[0] + off_nc, 40.0f)), 0.0f) * scale;
    const float pnz = fmaxf(expm1f(fminf(y[1] + off_nnz, 40.0f)), 0.0f) * scale;
    const int64_t bound = min<int64_t>(nb, span);
    int64_t cap = pnc >= (float)bound ? bound : (int64_t)ceilf(pnc) + extra;
    cap = min<int64_t>(cap, bound);
    int64_t vcap = pnz >= (float)(cap * 256) ? cap * 256 : (int64_t)ceilf(pnz) + 8 * extra;
    vcap = min<int64_t>(vcap, cap * 256);
    slots[bi + 1] = cap;
    valcap[bi + 1] = vcap;
    if (bi == 0) {
        slots[0] = 0;
        valcap[0] = 0;
    }
    int b = bt.nvar;
    for (int v = 0; v < bt.nvar; v++) {
        const StageCaps c = bt.v[v];
        if (na <= c.na && cap <= c.nc && nb <= c.nb && pa <= c.pa && pb <= c.pb) { b = v; break; }
    }
    if (b == bt.nvar && bt.rg && pa <= (int64_t)bt.rg_density * na) b = bt.nvar + 1;
    bin[bi] = (uint8_t)b;
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

// ---------------------------------------------------------------- slot layout -> compact layout
__global__ void __launch_bounds__(256) k_rownc_ptr(int32_t tilem, const int32_t *__restrict__ row_nc,
                                                   int64_t *__restrict__ tile_ptr) {
    const int32_t bi = blockIdx.x * 256 + threadIdx.x;
    if (bi >= tilem) return;
    tile_ptr[bi + 1] = row_nc[bi];
    if (bi == 0) tile_ptr[0] = 0;
}

// One wave per tile-row copies its tiles' headers from the slots to the
// compact arrays; the value offsets are already exact.
__global__ void __launch_bounds__(256)
k_compact(int32_t tilem, const int64_t *__restrict__ slot_ptr, const int32_t *__restrict__ row_nc,
          const int64_t *__restrict__ tile_ptr, const int32_t *__restrict__ pcol, const int64_t *__restrict__ pnnz,
          const uint16_t *__restrict__ pmask, const uint8_t *__restrict__ pptr, int32_t *__restrict__ rowidx,
          int32_t *__restrict__ colidx, int64_t *__restrict__ tile_nnz, uint16_t *__restrict__ mask,
          uint8_t *__restrict__ ptr) {
    const int32_t bi = (int32_t)(((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE);
    const int lane = lane_id();
    if (bi >= tilem) return;
    const int64_t s0 = slot_ptr[bi], d0 = tile_ptr[bi];
    const int n = row_nc[bi];
    for (int t = lane; t < n; t += WAVE) {
        const int64_t src = s0 + t, dst = d0 + t;
        rowidx[dst] = bi;
        colidx[dst] = pcol[src];
        tile_nnz[dst] = pnnz[src];
        reinterpret_cast<uint4 *>(ptr)[dst] = reinterpret_cast<const uint4 *>(pptr)[src];
        reinterpret_cast<uint4 *>(mask)[dst * 2] = reinterpret_cast<const uint4 *>(pmask)[src * 2];
        reinterpret_cast<uint4 *>(mask)[dst * 2 + 1] = reinterpret_cast<const uint4 *>(pmask)[src * 2 + 1];
    }
    if (bi == tilem - 1 && lane == 0) tile_nnz[tile_ptr[tilem]] = pnnz[slot_ptr[tilem]];
}

// Launch sizes taken from the capacity cache are checked on the device before
// any kernel uses them: when this call's plan (total slots, value capacity,
// bin offsets) differs from the cached one, every slot capacity becomes 0, so
// no kernel writes C (each row overflows or is empty), and the overflow flag
// sends the call to the exact pipeline.
__global__ void __launch_bounds__(256) k_plan_guard(const int64_t *__restrict__ plan_hdr,
                                                    const int64_t *__restrict__ expect, int nexp,
                                                    int64_t *__restrict__ slotp, int32_t tilem,
                                                    int32_t *__restrict__ overflow) {
    __shared__ int bad;
    if (threadIdx.x == 0) {
        int m = 0;
        for (int i = 0; i < nexp; i++) m |= plan_hdr[i] != expect[i];
        bad = m;
        if (m) *overflow = 1;
    }
    __syncthreads();
    if (!bad) return;
    for (int64_t i = threadIdx.x; i <= tilem; i += 256) slotp[i] = 0;
}

// hdr[0] = overflow, hdr[1] = numtileC, hdr[2] = nnzC
__global__ void k_learned_result(const int32_t *__restrict__ overflow, const int64_t *__restrict__ tile_ptr,
                                 int32_t tilem, const int64_t *__restrict__ pnnz, const int64_t *__restrict__ slotp,
                                 int64_t *__restrict__ hdr) {
    if (threadIdx.x != 0) return;
    hdr[0] = *overflow;
    hdr[1] = tile_ptr[tilem];
    hdr[2] = pnnz[slotp[tilem]];
}

static unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// ---------------------------------------------------------------- the learned product
int spgemm_learned(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *st) {
    Learned *L = state(ctx);
    if (!L->init) return fail(TSG_ERR_ARG, "learned scheduler has no model: train it or set its weights first");
    const bool timing = opts && opts->timing;
    const double pscale = (opts && opts->pred_scale > 0) ? opts->pred_scale : 1.0;
    // capacity = calibrated prediction + 1 slot (+ 8 values); pred_scale < 1
    // shrinks it without the constant (a test hook that forces the fallback)
    const float cscale = (float)pscale;
    const int cextra = pscale >= 1.0 ? 1 : 0;
    hipStream_t s = ctx->stream;
    hipEvent_t ev[4] = {};
    if (timing)
        for (auto &e : ev) TSG_HIP(hipEventCreate(&e));
    if (timing) (void)hipEventRecord(ev[0], s);
    memset(C, 0, sizeof *C);
    C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
    C->dtype = A->dtype; C->location = TSG_DEVICE;
    const int32_t tilem = A->tilem;
    const BinTable bt = bin_table_for(ctx, A, B, opts);
    Plan p;
    int64_t *slotp = nullptr, *valc = nullptr, *expect = nullptr, *res = nullptr;
    int32_t *row_nc = nullptr, *pcol = nullptr, *prow = nullptr;
    int64_t *pnnz = nullptr;
    uint16_t *pmask = nullptr;
    uint8_t *pptr = nullptr;
    int syncs = 0;
    int64_t slots = 0, vcap = 0;
    auto cleanup = [&]() {
        void *ps[] = {slotp, valc, expect, res, row_nc, pcol, prow, pnnz, pmask, pptr};
        for (void *q : ps) pool_free(ctx, q);
        free_plan(ctx, p);
    };
    int rc = plan_setup(ctx, A, B, bt, p);
    if (!rc) rc = palloc(ctx, &slotp, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &valc, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &expect, MAX_BINS + 3);
    if (!rc) rc = palloc(ctx, &res, 4);
    if (!rc && tilem > 0) {
        k_learned_plan<<<nblk(tilem, 256), 256, 0, s>>>(tilem, A->tile_ptr, A->tile_nnz, p.link, B->tile_columnidx,
                                                         L->w, bt, L->off_nc, L->off_nnz, cscale, cextra, nullptr,
                                                         nullptr, slotp, valc, p.bin);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_learned_plan launch");
        if (!rc) rc = scan_inclusive_i64(ctx, slotp, (int64_t)tilem + 1);
        if (!rc) rc = scan_inclusive_i64(ctx, valc, (int64_t)tilem + 1);
    } else if (!rc) {
        if (hipMemsetAsync(slotp, 0, 8, s) != hipSuccess || hipMemsetAsync(valc, 0, 8, s) != hipSuccess)
            rc = fail(TSG_ERR_HIP, "memset");
    }
    if (!rc) rc = plan_bin(ctx, p, tilem, slotp + tilem, valc + tilem);
    // sizes: from the cache when this product was planned before (the final
    // check compares the cached plan with this one), else one host read
    Learned::Cache &cc = L->cache;
    const bool hit = cc.valid && cc.a == A->tile_ptr && cc.b == B->tile_ptr && cc.annz == A->nnz &&
                     cc.tilem == A->tilem && cc.tilen == B->tilen &&
                     cc.bnnz == B->nnz && cc.atiles == A->numtile && cc.btiles == B->numtile && cc.nbins == p.nbins &&
                     pscale == 1.0;
    if (!rc) {
        if (hit) {
            slots = cc.slots;
            vcap = cc.valcap;
            for (int b = 0; b <= p.nbins; b++) p.bin_off[b] = cc.bin_off[b];
        } else {
            rc = plan_read(ctx, p, &slots, &vcap);
            syncs++;
            if (!rc && pscale == 1.0) {
                cc.a = A->tile_ptr; cc.b = B->tile_ptr; cc.annz = A->nnz; cc.bnnz = B->nnz;
                cc.tilem = A->tilem; cc.tilen = B->tilen;
                cc.atiles = A->numtile; cc.btiles = B->numtile; cc.slots = slots; cc.valcap = vcap;
                cc.nbins = p.nbins;
                for (int b = 0; b <= p.nbins; b++) cc.bin_off[b] = p.bin_off[b];
                cc.valid = true;
            }
        }
    }
    if (!rc && hit) {  // check the cached sizes against this call's plan before any launch uses them
        int64_t h[MAX_BINS + 3];
        h[0] = slots;
        h[1] = vcap;
        for (int b = 0; b <= p.nbins; b++) h[2 + b] = p.bin_off[b];
        memcpy(ctx->pinned + 8, h, sizeof(int64_t) * (p.nbins + 3));
        if (hipMemcpyAsync(expect, ctx->pinned + 8, sizeof(int64_t) * (p.nbins + 3), hipMemcpyHostToDevice, s) !=
            hipSuccess)
            rc = fail(TSG_ERR_HIP, "upload of the plan sizes failed");
        if (!rc) {
            k_plan_guard<<<1, 256, 0, s>>>(p.hdr, expect, p.nbins + 3, slotp, tilem, p.overflow());
            if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_plan_guard launch");
        }
    }
    // The slot layout and the compact C both hold `slots` tiles (120 bytes a
    // tile) next to the values: when that exceeds int32 tile ids or the free
    // device memory (with the pool's cached blocks), the exact pipeline runs
    // instead (same result).
    if (!rc) {
        size_t freeb = 0, totalb = 0;
        const double need = (double)slots * 120.0 + (double)vcap * (double)(vsize(A->dtype) + 1);
        if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) rc = fail(TSG_ERR_HIP, "hipMemGetInfo failed");
        if (!rc && (slots > INT32_MAX || need > 0.9 * ((double)freeb + (double)ctx->pool.cached))) {
            cleanup();
            tsg_options o = opts ? *opts : tsg_options{};
            o.scheduler = TSG_SCHED_PLAIN;
            TSG_CHECK(spgemm_binned(ctx, A, B, C, &o, st));
            if (st) {
                st->fallback_used = 2;
                st->host_syncs += syncs;
                st->pred_numtileC = slots;
                st->pred_nnzC = vcap;
            }
            if (timing)
                for (auto &e : ev) (void)hipEventDestroy(e);
            return TSG_OK;
        }
    }
    // slot-layout structure and the compact C
    if (!rc) rc = palloc(ctx, &row_nc, (size_t)tilem);
    if (!rc) rc = palloc(ctx, &prow, (size_t)slots);
    if (!rc) rc = palloc(ctx, &pcol, (size_t)slots);
    if (!rc) rc = palloc(ctx, &pnnz, (size_t)slots + 1);
    if (!rc) rc = palloc(ctx, &pmask, (size_t)slots * 16);
    if (!rc) rc = palloc(ctx, &pptr, (size_t)slots * 16);
    if (!rc) rc = palloc(ctx, &C->tile_ptr, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &C->tile_rowidx, (size_t)slots);
    if (!rc) rc = palloc(ctx, &C->tile_columnidx, (size_t)slots);
    if (!rc) rc = palloc(ctx, &C->tile_nnz, (size_t)slots + 1);
    if (!rc) rc = palloc(ctx, &C->mask, (size_t)slots * 16);
    if (!rc) rc = palloc(ctx, &C->tile_csr_Ptr, (size_t)slots * 16);
    if (!rc) rc = pool_alloc(ctx, &C->tile_csr_Val, (size_t)vcap * vsize(C->dtype));
    if (!rc) rc = palloc(ctx, &C->tile_csr_Col, (size_t)vcap);
    if (!rc && hipMemsetAsync(pnnz, 0, sizeof(int64_t) * ((size_t)slots + 1), s) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "memset");
    if (timing && !rc) (void)hipEventRecord(ev[1], s);
    tsg_tile_matrix Cp = *C;  // slot-layout view: kernels index tiles by slot
    Cp.tile_ptr = slotp;
    Cp.tile_rowidx = prow;
    Cp.tile_columnidx = pcol;
    Cp.tile_nnz = pnnz;
    Cp.mask = pmask;
    Cp.tile_csr_Ptr = pptr;
    Cp.numtile = slots;
    SlotLayout sl;
    sl.row_nc = row_nc;
    sl.overflow = p.overflow();
    sl.val_cap = vcap;
    if (!rc) rc = run_symbolic(ctx, A, B, &Cp, p, sl);
    if (!rc) rc = scan_inclusive_i64(ctx, pnnz, slots + 1);
    if (!rc) rc = prepare_numeric(ctx, A, &Cp, p, sl);
    if (!rc && tilem > 0) {
        k_rownc_ptr<<<nblk(tilem, 256), 256, 0, s>>>(tilem, row_nc, C->tile_ptr);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_rownc_ptr launch");
    } else if (!rc && hipMemsetAsync(C->tile_ptr, 0, 8, s) != hipSuccess) {
        rc = fail(TSG_ERR_HIP, "memset");
    }
    if (!rc) rc = scan_inclusive_i64(ctx, C->tile_ptr, (int64_t)tilem + 1);
    if (timing && !rc) (void)hipEventRecord(ev[2], s);
    if (!rc) rc = run_numeric(ctx, A, B, &Cp, p, sl);
    if (!rc && tilem > 0) {
        k_compact<<<nblk((int64_t)tilem * WAVE, 256), 256, 0, s>>>(tilem, slotp, row_nc, C->tile_ptr, pcol, pnnz,
                                                                     pmask, pptr, C->tile_rowidx, C->tile_columnidx,
                                                                     C->tile_nnz, C->mask, C->tile_csr_Ptr);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_compact launch");
    } else if (!rc && hipMemsetAsync(C->tile_nnz, 0, 8, s) != hipSuccess) {
        rc = fail(TSG_ERR_HIP, "memset");
    }
    if (timing && !rc) (void)hipEventRecord(ev[3], s);
    bool redo = false;
    if (!rc) {
        k_learned_result<<<1, 64, 0, s>>>(p.overflow(), C->tile_ptr, tilem, pnnz, slotp, res);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(ctx->pinned, res, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = fail(TSG_ERR_HIP, "reading the learned result failed");
        syncs++;
        if (!rc) {
            redo = ctx->pinned[0] != 0;
            if (redo) cc.valid = false;
            C->numtile = ctx->pinned[1];
            C->nnz = ctx->pinned[2];
        }
    }
    if (timing) {
        if (!rc && !redo && st) {
            float a = 0, b = 0, c3 = 0;
            (void)hipEventElapsedTime(&a, ev[0], ev[1]);
            (void)hipEventElapsedTime(&b, ev[1], ev[2]);
            (void)hipEventElapsedTime(&c3, ev[2], ev[3]);
            st->ms_step1 = a; st->ms_step2 = b; st->ms_step3 = c3; st->ms_total = a + b + c3;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    cleanup();
    if (rc) {
        free_c(ctx, C);
        return rc;
    }
    if (redo) {  // the model under-predicted somewhere: exact pipeline, same result
        free_c(ctx, C);
        tsg_options o = opts ? *opts : tsg_options{};
        o.scheduler = TSG_SCHED_PLAIN;
        TSG_CHECK(spgemm_binned(ctx, A, B, C, &o, st));
        if (st) {
            st->fallback_used = 1;
            st->host_syncs += syncs;
            st->pred_numtileC = slots;
            st->pred_nnzC = vcap;
        }
        return TSG_OK;
    }
    if (st) {
        st->numtileC = C->numtile;
        st->nnzC = C->nnz;
        st->host_syncs = syncs;
        st->fallback_used = 0;
        st->pred_numtileC = slots;
        st->pred_nnzC = vcap;
    }
    return TSG_OK;
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
    TSG_HIP(hipMemcpy(L->w, h.data(), NP * sizeof(float), hipMemcpyHostToDevice));
    TSG_HIP(hipMemset(L->m, 0, NP * sizeof(float)));
    TSG_HIP(hipMemset(L->v, 0, NP * sizeof(float)));
    L->step = 0;
    return TSG_OK;
}

// Log-space offsets that make every training row's capacity cover its exact
// size: off = max over rows of (target - prediction), clamped to
// [log 1.05, log 16].
static int calibrate(Context *ctx) {
    Learned *L = state(ctx);
    const int64_t n = L->n;
    float *Y = nullptr;
    TSG_CHECK(palloc(ctx, &Y, (size_t)n * NO));
    std::vector<float> hy((size_t)n * NO), ht((size_t)n * NO);
    k_mlp_forward<<<nblk(n, 256), 256, 0, ctx->stream>>>(L->w, L->X, n, Y);
    int rc = TSG_OK;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(hy.data(), Y, sizeof(float) * n * NO, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ht.data(), L->T, sizeof(float) * n * NO, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "calibration failed");
    pool_free(ctx, Y);
    if (rc) return rc;
    float off[NO] = {0.0f, 0.0f};
    for (int64_t i = 0; i < n; i++)
        for (int o = 0; o < NO; o++) off[o] = std::max(off[o], ht[i * NO + o] - hy[i * NO + o]);
    const float lo = std::log(1.05f), hi = std::log(16.0f);
    L->off_nc = std::min(std::max(off[0], lo), hi);
    L->off_nnz = std::min(std::max(off[1], lo), hi);
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
    float *X = nullptr, *Y = nullptr;
    int64_t *slotp = nullptr, *valc = nullptr;
    const BinTable bt = bin_table_for(ctx, A, B, nullptr);
    int rc = plan_setup(ctx, A, B, bt, p);
    if (!rc) rc = palloc(ctx, &X, (size_t)std::max(tilem, 1) * NF);
    if (!rc) rc = palloc(ctx, &Y, (size_t)std::max(tilem, 1) * NO);
    if (!rc) rc = palloc(ctx, &slotp, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &valc, (size_t)tilem + 1);
    if (!rc && tilem > 0) {
        k_learned_plan<<<nblk(tilem, 256), 256, 0, ctx->stream>>>(
            tilem, A->tile_ptr, A->tile_nnz, p.link, B->tile_columnidx, pred ? L->w : nullptr,
            bt, L->off_nc, L->off_nnz, 1.0f, 1, X, Y, slotp, valc, p.bin);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_learned_plan launch");
    }
    if (!rc && feats && tilem > 0 &&
        hipMemcpyAsync(feats, X, sizeof(float) * tilem * NF, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc && pred && tilem > 0 &&
        hipMemcpyAsync(pred, Y, sizeof(float) * tilem * NO, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "sync");
    void *ps[] = {X, Y, slotp, valc};
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
    float *X = nullptr, *T = nullptr;
    int64_t *slotp = nullptr, *valc = nullptr;
    int rc = plan_setup(ctx, A, B, bin_table_for(ctx, A, B, nullptr), p);
    if (!rc) rc = palloc(ctx, &X, (size_t)tilem * NF);
    if (!rc) rc = palloc(ctx, &T, (size_t)tilem * NO);
    if (!rc) rc = palloc(ctx, &slotp, (size_t)tilem + 1);
    if (!rc) rc = palloc(ctx, &valc, (size_t)tilem + 1);
    if (!rc) {
        k_learned_plan<<<nblk(tilem, 256), 256, 0, ctx->stream>>>(tilem, A->tile_ptr, A->tile_nnz, p.link,
                                                                   B->tile_columnidx, nullptr, BinTable{}, 0.0f, 0.0f,
                                                                   1.0f, 0, X, nullptr, slotp, valc, p.bin);
        k_targets<<<nblk(tilem, 256), 256, 0, ctx->stream>>>(tilem, C.tile_ptr, C.tile_nnz, T);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "feature/target launch");
    }
    if (!rc) rc = append_rows(ctx, X, T, tilem);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "sync");
    void *ps[] = {X, T, slotp, valc};
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
    L->cache.valid = false;
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
        if (hipMemcpy(perm, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
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
                if (hipMemcpy(hl.data(), lossp, sizeof(float) * nblocks, hipMemcpyDeviceToHost) != hipSuccess) {
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
    L->cache.valid = false;
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
        L->cache.valid = false;
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
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    TSG_HIP(hipMemcpy(w, L->w, sizeof(float) * NP, hipMemcpyDeviceToHost));
    return TSG_OK;
}

int tsg_learned_set_weights(tsg_context *c, const float *w, int32_t n) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !w || n != NP) return fail(TSG_ERR_ARG, "need %d weights", NP);
    (void)hipSetDevice(ctx->device);
    TSG_CHECK(ensure_weights(ctx, 0));
    Learned *L = state(ctx);
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    TSG_HIP(hipMemcpy(L->w, w, sizeof(float) * NP, hipMemcpyHostToDevice));
    TSG_HIP(hipMemset(L->m, 0, NP * sizeof(float)));
    TSG_HIP(hipMemset(L->v, 0, NP * sizeof(float)));
    L->step = 0;
    L->init = true;
    L->cache.valid = false;
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
        if (hipMemcpy(dx, X, sizeof(float) * n * NF, hipMemcpyHostToDevice) != hipSuccess) rc = fail(TSG_ERR_HIP, "copy");
        if (!rc) {
            k_mlp_forward<<<nblk(n, 256), 256, 0, ctx->stream>>>(L->w, dx, n, dy);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess ||
                hipMemcpy(Y, dy, sizeof(float) * n * NO, hipMemcpyDeviceToHost) != hipSuccess)
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
    if (!rc && (hipMemcpy(dx, X, sizeof(float) * n * NF, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dt, T, sizeof(float) * n * NO, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc) {
        k_mlp_grad<<<nblocks, GRAD_BLOCK, 0, ctx->stream>>>(L->w, dx, dt, nullptr, 0, n, partial, lossp);
        k_grad_reduce<<<nblk(NP, 256), 256, 0, ctx->stream>>>(partial, nblocks, g);
        std::vector<float> hl((size_t)nblocks);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess ||
            hipMemcpy(grad, g, sizeof(float) * NP, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hl.data(), lossp, sizeof(float) * nblocks, hipMemcpyDeviceToHost) != hipSuccess)
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
