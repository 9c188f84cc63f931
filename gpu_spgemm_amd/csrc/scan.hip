// scan.hip -- device-wide int64 inclusive scan (reduce, apply; partials scanned
// in between only for more than 2048 blocks).
//
// Replaces the reference's cub::DeviceScan::ExclusiveSum calls
// (src/spgemm_raw_v2.cu:131-138 and :202-208).  Callers store 0 in data[0] and
// the per-item counts in data[1..n-1]; the inclusive scan then leaves the
// exclusive offsets plus the total in data[n-1].  64-bit throughout, and the
// temporary partial sums come from the context pool (no hipMalloc/hipFree per
// call as in the reference).
#include <cstring>

#include "tsg_internal.h"

namespace tsg {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// The items a scan covers: n, or with dlast (a captured graph, where the
// length lives on the device) the items up to index *dlast, at most n.
__device__ __forceinline__ int64_t scan_len(int64_t n, const int64_t *dlast) {
    return dlast ? min(n, *dlast + 1) : n;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const int64_t *__restrict__ data,
                                                               int64_t n, const int64_t *__restrict__ dlast,
                                                               int64_t *__restrict__ partial, bool zero_first) {
    __shared__ int64_t s_w[SCAN_THREADS / WAVE];
    n = scan_len(n, dlast);
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        int64_t idx = base + i * SCAN_THREADS + threadIdx.x;
        if (idx < n && !(zero_first && idx == 0)) s += data[idx];
    }
    s = wave_sum(s);
    if (lane_id() == 0) s_w[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < SCAN_THREADS / WAVE; w++) t += s_w[w];
        partial[blockIdx.x] = t;
    }
}

// Inclusive scan of one SCAN_TILE chunk per block, plus the carried-in offset:
// the inclusive partial of the previous block (partial_incl), or the sum of
// the earlier blocks' raw partials added up by this block (partial_raw, few
// blocks: saves scanning the partials in a launch of their own).
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(int64_t *__restrict__ data, int64_t n,
                                                              const int64_t *__restrict__ dlast,
                                                              const int64_t *__restrict__ partial_incl,
                                                              const int64_t *__restrict__ partial_raw,
                                                              bool zero_first, int64_t *__restrict__ total_out,
                                                              const int32_t *__restrict__ pub_word, int32_t *pub,
                                                              int64_t seq) {
    __shared__ int64_t s_d[SCAN_TILE + SCAN_TILE / 32];
    __shared__ int64_t s_tot;
    __shared__ int64_t s_w[SCAN_THREADS / WAVE];
    __shared__ int64_t s_c[SCAN_THREADS / WAVE];
    n = scan_len(n, dlast);
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t carry = 0;
    if (partial_raw) {
        for (int64_t j = threadIdx.x; j < blockIdx.x; j += SCAN_THREADS) carry += partial_raw[j];
        carry = wave_sum(carry);
        if (lane_id() == 0) s_c[threadIdx.x / WAVE] = carry;
    }
    auto pad = [](int i) { return i + (i >> 5); };
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        int li = i * SCAN_THREADS + threadIdx.x;
        int64_t idx = base + li;
        s_d[pad(li)] = (idx < n && !(zero_first && idx == 0)) ? data[idx] : 0;
    }
    __syncthreads();
    int64_t v[SCAN_ITEMS];
    int64_t run = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        run += s_d[pad(threadIdx.x * SCAN_ITEMS + i)];
        v[i] = run;
    }
    int64_t incl = wave_incl_sum(run);
    int w = threadIdx.x / WAVE;
    if (lane_id() == WAVE - 1) s_w[w] = incl;
    __syncthreads();
    int64_t off = (blockIdx.x > 0 && partial_incl) ? partial_incl[blockIdx.x - 1] : 0;
    if (partial_raw)
        for (int q = 0; q < SCAN_THREADS / WAVE; q++) off += s_c[q];
    for (int q = 0; q < w; q++) off += s_w[q];
    off += incl - run;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) s_d[pad(threadIdx.x * SCAN_ITEMS + i)] = v[i] + off;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        int li = i * SCAN_THREADS + threadIdx.x;
        int64_t idx = base + li;
        if (idx < n) {
            data[idx] = s_d[pad(li)];
            if (total_out && idx == n - 1) *total_out = s_d[pad(li)];
            if (idx == n - 1) s_tot = s_d[pad(li)];
        }
    }
    if (pub && blockIdx.x == gridDim.x - 1) {  // the block holding the total publishes it (dlast unused)
        __syncthreads();
        if (threadIdx.x == 0) {
            pub[0] = *pub_word;
            pub[1] = 0;
            pub[2] = (int32_t)(uint32_t)(uint64_t)s_tot;
            pub[3] = (int32_t)(uint32_t)((uint64_t)s_tot >> 32);
            __threadfence_system();
            __hip_atomic_store(reinterpret_cast<int64_t *>(pub + PUB_WORDS), seq, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Zeroes a few small buffers in one launch (a kernel: capturable in a graph,
// and no runtime fill between the pipeline's own kernels).
__global__ void k_zero_small(int32_t *__restrict__ a, int na, int64_t *__restrict__ b, int nb) {
    const int t = threadIdx.x;
    if (t < na) a[t] = 0;
    if (t < nb) b[t] = 0;
}

int zero_small(Context *ctx, int32_t *a, int na, int64_t *b, int nb) {
    if (na > 64 || nb > 64) return fail(TSG_ERR_ARG, "zero_small: at most 64 words each");
    if ((!a || na <= 0) && (!b || nb <= 0)) return TSG_OK;
    k_zero_small<<<1, 64, 0, ctx->stream>>>(a, a ? na : 0, b, b ? nb : 0);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

// pub_word: the block holding the total also publishes (*pub_word, 0, total)
// to the host's words with sequence number seq.
static int scan_impl(Context *ctx, int64_t *data, int64_t n, bool zero_first, int64_t *total_out,
                     const int64_t *dlast, const int32_t *pub_word, int64_t seq) {
    if (n <= 0) return TSG_OK;
    int32_t *const pub = pub_word ? ctx->pub_dev : nullptr;
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb > INT32_MAX) return fail(TSG_ERR_OVERFLOW, "scan of %lld items exceeds grid limit", (long long)n);
    if (nb == 1) {
        k_scan_apply<<<1, SCAN_THREADS, 0, ctx->stream>>>(data, n, dlast, nullptr, nullptr, zero_first, total_out,
                                                          pub_word, pub, seq);
        TSG_HIP(hipGetLastError());
        return TSG_OK;
    }
    int64_t *partial = nullptr;
    TSG_CHECK(palloc(ctx, &partial, (size_t)nb));
    k_scan_reduce<<<(unsigned)nb, SCAN_THREADS, 0, ctx->stream>>>(data, n, dlast, partial, zero_first);
    int rc = TSG_OK;
    if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_scan_reduce launch failed");
    // up to SCAN_TILE blocks each add up the raw partials before them (at most
    // 16 KB of reads, cached); more blocks scan the partials first
    const bool raw = nb <= SCAN_TILE;
    if (rc == TSG_OK && !raw) rc = scan_inclusive_i64(ctx, partial, nb);
    if (rc == TSG_OK) {
        k_scan_apply<<<(unsigned)nb, SCAN_THREADS, 0, ctx->stream>>>(data, n, dlast, raw ? nullptr : partial,
                                                                     raw ? partial : nullptr, zero_first, total_out,
                                                                     pub_word, pub, seq);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_scan_apply launch failed");
    }
    pool_free(ctx, partial);
    return rc;
}

int scan_inclusive_i64(Context *ctx, int64_t *data, int64_t n, bool zero_first, int64_t *total_out,
                       const int64_t *dlast) {
    return scan_impl(ctx, data, n, zero_first, total_out, dlast, nullptr, 0);
}

int scan_publish_i64(Context *ctx, int64_t *data, int64_t n, bool zero_first, int64_t *total_out,
                     const int32_t *word, int64_t *seq) {
    if (n <= 0) return fail(TSG_ERR_ARG, "scan_publish_i64: empty scan");
    *seq = publish_seq(ctx);
    return scan_impl(ctx, data, n, zero_first, total_out, nullptr, word, *seq);
}

// The words, then (after a system-scope fence by every writer and a barrier)
// the sequence number: a host that sees the number sees the words.
__global__ void __launch_bounds__(PUB_WORDS) k_publish(const int32_t *__restrict__ src, int n, int32_t *dst,
                                                        int64_t seq) {
    const int t = threadIdx.x;
    if (t < n) dst[t] = src[t];
    __threadfence_system();
    __syncthreads();
    if (t == 0) __hip_atomic_store(reinterpret_cast<int64_t *>(dst + PUB_WORDS), seq, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
}

int64_t publish_seq(Context *ctx) { return ++ctx->pub_seq; }

int publish_wait(Context *ctx, int64_t seq, int n, void *out) {
    const int64_t *flag = reinterpret_cast<const int64_t *>(ctx->pub + PUB_WORDS);
    for (uint64_t it = 1;; it++) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) break;
        if ((it & 16383) == 0) {  // bounded: a stream that finished without the number is an error
            const hipError_t q = hipStreamQuery(ctx->stream);
            if (q == hipSuccess) {
                if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) break;
                return fail(TSG_ERR_HIP, "a device-to-host read did not arrive");
            }
            if (q != hipErrorNotReady) return fail(TSG_ERR_HIP, "device-to-host read: %s", hipGetErrorString(q));
        }
    }
    memcpy(out, ctx->pub, sizeof(int32_t) * (size_t)n);
    return TSG_OK;
}

int publish_start(Context *ctx, const void *dsrc, int n, int64_t *seq) {
    if (n < 0 || n > PUB_WORDS) return fail(TSG_ERR_ARG, "read_words: at most %d words", PUB_WORDS);
    *seq = publish_seq(ctx);
    k_publish<<<1, PUB_WORDS, 0, ctx->stream>>>(static_cast<const int32_t *>(dsrc), n, ctx->pub_dev, *seq);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int read_words(Context *ctx, const void *dsrc, int n, void *out) {
    int64_t seq = 0;
    TSG_CHECK(publish_start(ctx, dsrc, n, &seq));
    return publish_wait(ctx, seq, n, out);
}

int read_i64(Context *ctx, const int64_t *dptr, int64_t *out) { return read_words(ctx, dptr, 2, out); }

int read_i32(Context *ctx, const int32_t *dptr, int32_t *out) { return read_words(ctx, dptr, 1, out); }

}  // namespace tsg

extern "C" int tsg_scan_i64(tsg_context *c, int64_t *data, int64_t n, int32_t zero_first, int64_t *total) {
    using namespace tsg;
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || (n > 0 && !data) || n < 0) return fail(TSG_ERR_ARG, "bad arguments");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) TSG_HIP(hipSetDevice(ctx->device));
    const int rc = scan_inclusive_i64(ctx, data, n, zero_first != 0, total);
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    return rc;
}
