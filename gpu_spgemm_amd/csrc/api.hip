// api.hip -- C ABI: context, device pool, transfers and the plain SpGEMM
// pipeline (Steps 1-3 of src/spgemm_raw_v2.cu:111-240).
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "rowstaged.h"

namespace tsg {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int pool_alloc(Context *ctx, void **p, size_t bytes) {
    if (ctx->arena) {  // no-alloc call: bump-allocate from the scratch region
        Arena &a = *ctx->arena;
        const size_t r = (bytes + 255) & ~size_t(255);
        if (r > a.size - a.top) {
            *p = nullptr;
            return fail(TSG_ERR_OVERFLOW, "the no-alloc scratch (%zu bytes) cannot hold this product's work arrays "
                        "(%zu more bytes needed)", a.size, r - (a.size - a.top));
        }
        *p = a.base + a.top;
        a.top += r;
        return TSG_OK;
    }
    size_t r = Pool::round(bytes);
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->pool.free_blocks.find(r);
    if (it != ctx->pool.free_blocks.end()) {
        *p = it->second;
        ctx->pool.free_blocks.erase(it);
        ctx->pool.cached -= (int64_t)r;
    } else {
        hipError_t e = hipMalloc(p, r);
        if (e != hipSuccess) {
            // release cached blocks and retry once
            for (auto &kv : ctx->pool.free_blocks) (void)hipFree(kv.second);
            ctx->pool.free_blocks.clear();
            ctx->pool.cached = 0;
            (void)hipGetLastError();
            e = hipMalloc(p, r);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                *p = nullptr;
                return fail(TSG_ERR_OOM, "device allocation of %zu bytes failed", r);
            }
        }
    }
    ctx->pool.live[*p] = r;
    ctx->pool.in_use += (int64_t)r;
    return TSG_OK;
}

int pool_view(Context *ctx, void *block, size_t offset, void **out) {
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->pool.live.find(block);
    if (it == ctx->pool.live.end()) return fail(TSG_ERR_ARG, "view of a pointer %p not owned by this context", block);
    if (offset >= it->second) return fail(TSG_ERR_ARG, "view offset %zu beyond a block of %zu bytes", offset, it->second);
    auto r = ctx->pool.refs.emplace(block, 1);  // the owner's reference
    r.first->second++;
    *out = static_cast<uint8_t *>(block) + offset;
    ctx->pool.views.emplace(*out, block);
    return TSG_OK;
}

int pool_free(Context *ctx, void *p) {
    if (!p || ctx->arena) return TSG_OK;  // arena arrays (and the caller's C) are not the pool's
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->deferred) {
        ctx->deferred->push_back(p);
        return TSG_OK;
    }
    // a view, or the owner of a block with views: drop one reference
    void *blk = nullptr;
    auto v = ctx->pool.views.find(p);
    if (v != ctx->pool.views.end()) {
        blk = v->second;
        ctx->pool.views.erase(v);
    } else if (ctx->pool.refs.count(p)) {
        blk = p;
    }
    if (blk) {
        auto r = ctx->pool.refs.find(blk);
        if (--r->second > 0) return TSG_OK;
        ctx->pool.refs.erase(r);
        p = blk;
    }
    auto it = ctx->pool.live.find(p);
    if (it == ctx->pool.live.end()) return fail(TSG_ERR_ARG, "pointer %p not owned by this context", p);
    size_t r = it->second;
    ctx->pool.live.erase(it);
    ctx->pool.in_use -= (int64_t)r;
    ctx->pool.free_blocks.emplace(r, p);
    ctx->pool.cached += (int64_t)r;
    return TSG_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

template <typename T>
static int up(Context *ctx, const T *src, size_t n, T **dst) {
    *dst = nullptr;
    TSG_CHECK(palloc(ctx, dst, n ? n : 1));
    if (n) TSG_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return TSG_OK;
}

template <typename T>
static int down(Context *ctx, const T *src, size_t n, T **dst) {
    *dst = (T *)malloc(n ? n * sizeof(T) : 1);
    if (!*dst) return fail(TSG_ERR_OOM, "host allocation failed");
    if (n) TSG_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    return TSG_OK;
}

}  // namespace tsg

using namespace tsg;

extern "C" {

const char *tsg_last_error_string(void) { return g_err.c_str(); }

int tsg_device_count(int *count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return TSG_OK;
}

int tsg_create(int device, tsg_context **out) {
    if (!out) return fail(TSG_ERR_ARG, "null context pointer");
    *out = nullptr;
    int cnt = 0;
    tsg_device_count(&cnt);
    if (cnt <= 0) return fail(TSG_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= cnt) return fail(TSG_ERR_ARG, "device %d out of range (%d devices)", device, cnt);
    Context *ctx = new Context();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&ctx->prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->pinned, 64 * sizeof(int64_t), hipHostMallocDefault);
    if (e == hipSuccess)  // PUB_WORDS words + the 64-bit sequence number
        e = hipHostMalloc((void **)&ctx->pub, sizeof(int32_t) * PUB_WORDS + sizeof(int64_t),
                          hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&ctx->pub_dev, ctx->pub, 0);
    if (e == hipSuccess) memset(ctx->pub, 0, sizeof(int32_t) * PUB_WORDS + sizeof(int64_t));
    if (e != hipSuccess) {
        if (ctx->pub) (void)hipHostFree(ctx->pub);
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return fail(TSG_ERR_HIP, "context set-up on device %d failed: %s", device, hipGetErrorString(e));
    }
    ctx->stream = ctx->own_stream;
    const int rc = noalloc_setup(ctx);  // the no-alloc entry point's scratch on this device
    if (rc) {
        (void)hipHostFree(ctx->pub);
        (void)hipHostFree(ctx->pinned);
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return rc;
    }
    *out = reinterpret_cast<tsg_context *>(ctx);
    return TSG_OK;
}

int tsg_destroy(tsg_context *c) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return TSG_OK;
    DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    learned_release(ctx);
    tuner_release(ctx);
    for (auto &kv : ctx->pool.free_blocks) (void)hipFree(kv.second);
    for (auto &kv : ctx->pool.live) (void)hipFree(kv.first);
    (void)hipHostFree(ctx->pub);
    (void)hipHostFree(ctx->pinned);
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return TSG_OK;
}

int tsg_synchronize(tsg_context *c) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    DeviceGuard g(ctx->device);
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}

int tsg_device_name(tsg_context *c, char *buf, size_t len) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !buf || !len) return fail(TSG_ERR_ARG, "bad arguments");
    snprintf(buf, len, "%s (%s, %d CUs)", ctx->prop.name, ctx->prop.gcnArchName, ctx->prop.multiProcessorCount);
    return TSG_OK;
}

void *tsg_stream(tsg_context *c) {
    Context *ctx = reinterpret_cast<Context *>(c);
    return ctx ? (void *)ctx->stream : nullptr;
}

int tsg_set_stream(tsg_context *c, void *stream) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    DeviceGuard g(ctx->device);
    // work already queued on the previous stream finishes before the switch,
    // so nothing of this context runs on two streams at once
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return TSG_OK;
}

int tsg_pool_stats(tsg_context *c, int64_t *in_use, int64_t *cached) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (in_use) *in_use = ctx->pool.in_use;
    if (cached) *cached = ctx->pool.cached;
    return TSG_OK;
}

int tsg_device_malloc(tsg_context *c, size_t bytes, void **ptr) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !ptr) return fail(TSG_ERR_ARG, "null argument");
    *ptr = nullptr;
    DeviceGuard g(ctx->device);
    return pool_alloc(ctx, ptr, bytes ? bytes : 1);
}

int tsg_device_free(tsg_context *c, void *ptr) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    return pool_free(ctx, ptr);
}

int tsg_memcpy(tsg_context *c, void *dst, const void *src, size_t bytes) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || (bytes && (!dst || !src))) return fail(TSG_ERR_ARG, "null argument");
    if (!bytes) return TSG_OK;
    DeviceGuard g(ctx->device);
    TSG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}

int tsg_pool_trim(tsg_context *c) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx) return fail(TSG_ERR_ARG, "null context");
    DeviceGuard g(ctx->device);
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto &kv : ctx->pool.free_blocks) (void)hipFree(kv.second);
    ctx->pool.free_blocks.clear();
    ctx->pool.cached = 0;
    return TSG_OK;
}

// ------------------------------------------------------------------ CSR transfers
int tsg_csr_upload(tsg_context *c, const tsg_csr *h, tsg_csr *d) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !h || !d) return fail(TSG_ERR_ARG, "null argument");
    if (h->location != TSG_HOST) return fail(TSG_ERR_ARG, "source CSR is not on the host");
    DeviceGuard g(ctx->device);
    *d = *h;
    d->location = TSG_DEVICE;
    d->rowptr = nullptr; d->col = nullptr; d->val = nullptr;
    int rc = up(ctx, h->rowptr, (size_t)h->m + 1, &d->rowptr);
    if (!rc) rc = up(ctx, h->col, (size_t)h->nnz, &d->col);
    if (!rc) rc = up(ctx, (const uint8_t *)h->val, (size_t)h->nnz * vsize(h->dtype), (uint8_t **)&d->val);
    if (!rc) { hipError_t e = hipStreamSynchronize(ctx->stream); if (e != hipSuccess) rc = fail(TSG_ERR_HIP, "%s", hipGetErrorString(e)); }
    return rc;
}

int tsg_csr_download(tsg_context *c, const tsg_csr *d, tsg_csr *h) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !h || !d) return fail(TSG_ERR_ARG, "null argument");
    DeviceGuard g(ctx->device);
    *h = *d;
    h->location = TSG_HOST;
    TSG_CHECK(down(ctx, d->rowptr, (size_t)d->m + 1, &h->rowptr));
    TSG_CHECK(down(ctx, d->col, (size_t)d->nnz, &h->col));
    TSG_CHECK(down(ctx, (const uint8_t *)d->val, (size_t)d->nnz * vsize(d->dtype), (uint8_t **)&h->val));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}

int tsg_csr_free(tsg_context *c, tsg_csr *d) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !d) return fail(TSG_ERR_ARG, "null argument");
    if (d->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "CSR is not on the device");
    pool_free(ctx, d->rowptr);
    pool_free(ctx, d->col);
    pool_free(ctx, d->val);
    memset(d, 0, sizeof *d);
    return TSG_OK;
}

// ------------------------------------------------------------------ tile transfers
int tsg_tile_upload(tsg_context *c, const tsg_tile_matrix *h, tsg_tile_matrix *d) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !h || !d) return fail(TSG_ERR_ARG, "null argument");
    if (h->location != TSG_HOST) return fail(TSG_ERR_ARG, "source tiles are not on the host");
    DeviceGuard g(ctx->device);
    tsg_tile_matrix t = *h;
    t.location = TSG_DEVICE;
    size_t nt = (size_t)h->numtile;
    t.tile_rowidx = nullptr; t.csc_tile_ptr = nullptr; t.csc_tile_rowidx = nullptr;
    TSG_CHECK(up(ctx, h->tile_ptr, (size_t)h->tilem + 1, &t.tile_ptr));
    TSG_CHECK(up(ctx, h->tile_columnidx, nt, &t.tile_columnidx));
    if (h->tile_rowidx) TSG_CHECK(up(ctx, h->tile_rowidx, nt, &t.tile_rowidx));
    TSG_CHECK(up(ctx, h->tile_nnz, nt + 1, &t.tile_nnz));
    TSG_CHECK(up(ctx, (const uint8_t *)h->tile_csr_Val, (size_t)h->nnz * vsize(h->dtype), (uint8_t **)&t.tile_csr_Val));
    TSG_CHECK(up(ctx, h->tile_csr_Col, (size_t)h->nnz, &t.tile_csr_Col));
    const size_t rows = h->tile_size ? (size_t)h->tile_size : 16;  // per-tile row entries
    TSG_CHECK(up(ctx, h->tile_csr_Ptr, nt * rows, &t.tile_csr_Ptr));
    TSG_CHECK(up(ctx, h->mask, nt * rows, &t.mask));
    t.csr_tile_nnz = nullptr; t.csr_mask = nullptr; t.csr_col_mask = nullptr; t.csr_col_Ptr = nullptr;
    t.csr_col_Val = nullptr; t.row_nnz_ptr = nullptr;
    t.row_ptr = nullptr; t.row_col = nullptr; t.row_val = nullptr;
    t.val_perm = nullptr; t.val_perm_col = nullptr;
    t.tile_csc_Ptr = nullptr; t.tile_csc_Row = nullptr; t.tile_csc_Val = nullptr;
    if (h->tile_csc_Ptr && h->tile_csc_Row && h->tile_csc_Val) {
        TSG_CHECK(up(ctx, h->tile_csc_Ptr, nt * 16, &t.tile_csc_Ptr));
        TSG_CHECK(up(ctx, h->tile_csc_Row, (size_t)h->nnz, &t.tile_csc_Row));
        TSG_CHECK(up(ctx, (const uint8_t *)h->tile_csc_Val, (size_t)h->nnz * vsize(h->dtype), (uint8_t **)&t.tile_csc_Val));
    }
    if (h->row_ptr && h->row_col && h->row_val) {
        TSG_CHECK(up(ctx, h->row_ptr, (size_t)h->m + 1, &t.row_ptr));
        TSG_CHECK(up(ctx, h->row_col, (size_t)h->nnz, &t.row_col));
        TSG_CHECK(up(ctx, (const uint8_t *)h->row_val, (size_t)h->nnz * vsize(h->dtype), (uint8_t **)&t.row_val));
    }
    if (h->is_csc) {
        if (!h->csc_tile_ptr || !h->csc_tile_rowidx) return fail(TSG_ERR_ARG, "B operand lacks CSC tile arrays");
        TSG_CHECK(up(ctx, h->csc_tile_ptr, (size_t)h->tilen + 1, &t.csc_tile_ptr));
        TSG_CHECK(up(ctx, h->csc_tile_rowidx, nt, &t.csc_tile_rowidx));
        if (h->csr_tile_nnz && h->csr_mask && h->csr_col_mask && h->csr_col_Ptr && h->csr_col_Val && h->row_nnz_ptr) {
            TSG_CHECK(up(ctx, h->csr_tile_nnz, nt + 1, &t.csr_tile_nnz));
            TSG_CHECK(up(ctx, h->csr_mask, nt * 16, &t.csr_mask));
            TSG_CHECK(up(ctx, h->csr_col_mask, nt * 16, &t.csr_col_mask));
            TSG_CHECK(up(ctx, h->csr_col_Ptr, nt * 16, &t.csr_col_Ptr));
            TSG_CHECK(up(ctx, (const uint8_t *)h->csr_col_Val, (size_t)h->nnz * vsize(h->dtype), (uint8_t **)&t.csr_col_Val));
            TSG_CHECK(up(ctx, h->row_nnz_ptr, (size_t)h->tilem + 1, &t.row_nnz_ptr));
        }
    }
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    *d = t;
    return TSG_OK;
}

int tsg_tile_download(tsg_context *c, const tsg_tile_matrix *d, tsg_tile_matrix *h) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !h || !d) return fail(TSG_ERR_ARG, "null argument");
    if (d->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles are not on the device");
    DeviceGuard g(ctx->device);
    tsg_tile_matrix t = *d;
    t.location = TSG_HOST;
    size_t nt = (size_t)d->numtile;
    t.tile_rowidx = nullptr; t.csc_tile_ptr = nullptr; t.csc_tile_rowidx = nullptr;
    TSG_CHECK(down(ctx, d->tile_ptr, (size_t)d->tilem + 1, &t.tile_ptr));
    TSG_CHECK(down(ctx, d->tile_columnidx, nt, &t.tile_columnidx));
    if (d->tile_rowidx) TSG_CHECK(down(ctx, d->tile_rowidx, nt, &t.tile_rowidx));
    TSG_CHECK(down(ctx, d->tile_nnz, nt + 1, &t.tile_nnz));
    TSG_CHECK(down(ctx, (const uint8_t *)d->tile_csr_Val, (size_t)d->nnz * vsize(d->dtype), (uint8_t **)&t.tile_csr_Val));
    TSG_CHECK(down(ctx, d->tile_csr_Col, (size_t)d->nnz, &t.tile_csr_Col));
    const size_t rows = d->tile_size ? (size_t)d->tile_size : 16;  // per-tile row entries
    TSG_CHECK(down(ctx, d->tile_csr_Ptr, nt * rows, &t.tile_csr_Ptr));
    TSG_CHECK(down(ctx, d->mask, nt * rows, &t.mask));
    t.csr_tile_nnz = nullptr; t.csr_mask = nullptr; t.csr_col_mask = nullptr; t.csr_col_Ptr = nullptr;
    t.csr_col_Val = nullptr; t.row_nnz_ptr = nullptr;
    t.row_ptr = nullptr; t.row_col = nullptr; t.row_val = nullptr;
    t.val_perm = nullptr; t.val_perm_col = nullptr;
    t.tile_csc_Ptr = nullptr; t.tile_csc_Row = nullptr; t.tile_csc_Val = nullptr;
    if (d->tile_csc_Ptr && d->tile_csc_Row && d->tile_csc_Val) {
        TSG_CHECK(down(ctx, d->tile_csc_Ptr, nt * 16, &t.tile_csc_Ptr));
        TSG_CHECK(down(ctx, d->tile_csc_Row, (size_t)d->nnz, &t.tile_csc_Row));
        TSG_CHECK(down(ctx, (const uint8_t *)d->tile_csc_Val, (size_t)d->nnz * vsize(d->dtype), (uint8_t **)&t.tile_csc_Val));
    }
    if (d->row_ptr && d->row_col && d->row_val) {
        TSG_CHECK(down(ctx, d->row_ptr, (size_t)d->m + 1, &t.row_ptr));
        TSG_CHECK(down(ctx, d->row_col, (size_t)d->nnz, &t.row_col));
        TSG_CHECK(down(ctx, (const uint8_t *)d->row_val, (size_t)d->nnz * vsize(d->dtype), (uint8_t **)&t.row_val));
    }
    if (d->is_csc) {
        TSG_CHECK(down(ctx, d->csc_tile_ptr, (size_t)d->tilen + 1, &t.csc_tile_ptr));
        TSG_CHECK(down(ctx, d->csc_tile_rowidx, nt, &t.csc_tile_rowidx));
        if (d->csr_tile_nnz && d->csr_mask && d->csr_col_mask && d->csr_col_Ptr && d->csr_col_Val && d->row_nnz_ptr) {
            TSG_CHECK(down(ctx, d->csr_tile_nnz, nt + 1, &t.csr_tile_nnz));
            TSG_CHECK(down(ctx, d->csr_mask, nt * 16, &t.csr_mask));
            TSG_CHECK(down(ctx, d->csr_col_mask, nt * 16, &t.csr_col_mask));
            TSG_CHECK(down(ctx, d->csr_col_Ptr, nt * 16, &t.csr_col_Ptr));
            TSG_CHECK(down(ctx, (const uint8_t *)d->csr_col_Val, (size_t)d->nnz * vsize(d->dtype), (uint8_t **)&t.csr_col_Val));
            TSG_CHECK(down(ctx, d->row_nnz_ptr, (size_t)d->tilem + 1, &t.row_nnz_ptr));
        }
    }
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    *h = t;
    return TSG_OK;
}

int tsg_tile_download_rows(tsg_context *c, const tsg_tile_matrix *d, int32_t tr_lo, int32_t tr_hi,
                           tsg_tile_matrix *h) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !h || !d) return fail(TSG_ERR_ARG, "null argument");
    if (d->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles are not on the device");
    if (d->is_csc) return fail(TSG_ERR_ARG, "a tile-row range of a B operand (CSC-tile data order) is not contiguous");
    if (d->tile_size && d->tile_size != 16) return fail(TSG_ERR_UNSUPPORTED, "16 x 16 tiles only");
    if (tr_lo < 0 || tr_hi < tr_lo || tr_hi > d->tilem)
        return fail(TSG_ERR_ARG, "tile-row range [%d, %d) outside [0, %d)", tr_lo, tr_hi, d->tilem);
    DeviceGuard g(ctx->device);
    // tile range [t0, t1) and value range [v0, v1) of the rows: two small reads
    int64_t t0 = 0, t1 = 0, v0 = 0, v1 = 0;
    TSG_CHECK(read_i64(ctx, d->tile_ptr + tr_lo, &t0));
    TSG_CHECK(read_i64(ctx, d->tile_ptr + tr_hi, &t1));
    TSG_CHECK(read_i64(ctx, d->tile_nnz + t0, &v0));
    TSG_CHECK(read_i64(ctx, d->tile_nnz + t1, &v1));
    tsg_tile_matrix t{};
    t.tilem = tr_hi - tr_lo;
    t.tilen = d->tilen;
    t.m = (int32_t)std::min<int64_t>((int64_t)t.tilem * 16, std::max<int64_t>(0, (int64_t)d->m - (int64_t)tr_lo * 16));
    t.n = d->n;
    t.numtile = t1 - t0;
    t.nnz = v1 - v0;
    t.dtype = d->dtype;
    t.location = TSG_HOST;
    t.flags = d->flags;
    const size_t nt = (size_t)t.numtile, vs = vsize(d->dtype);
    int rc = down(ctx, d->tile_ptr + tr_lo, (size_t)t.tilem + 1, &t.tile_ptr);
    if (!rc) rc = down(ctx, d->tile_columnidx + t0, nt, &t.tile_columnidx);
    if (!rc && d->tile_rowidx) rc = down(ctx, d->tile_rowidx + t0, nt, &t.tile_rowidx);
    if (!rc) rc = down(ctx, d->tile_nnz + t0, nt + 1, &t.tile_nnz);
    if (!rc) rc = down(ctx, (const uint8_t *)d->tile_csr_Val + v0 * vs, (size_t)t.nnz * vs, (uint8_t **)&t.tile_csr_Val);
    if (!rc) rc = down(ctx, d->tile_csr_Col + v0, (size_t)t.nnz, &t.tile_csr_Col);
    if (!rc) rc = down(ctx, d->tile_csr_Ptr + t0 * 16, nt * 16, &t.tile_csr_Ptr);
    if (!rc) rc = down(ctx, d->mask + t0 * 16, nt * 16, &t.mask);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(TSG_ERR_HIP, "%s", hipGetErrorString(e));
    }
    if (rc) {
        tsg_tile_free_host(&t);
        return rc;
    }
    for (int32_t i = 0; i <= t.tilem; i++) t.tile_ptr[i] -= t0;  // rebased to the range
    for (size_t i = 0; i <= nt; i++) t.tile_nnz[i] -= v0;
    if (t.tile_rowidx)
        for (size_t i = 0; i < nt; i++) t.tile_rowidx[i] -= tr_lo;
    *h = t;
    return TSG_OK;
}

int tsg_tile_free(tsg_context *c, tsg_tile_matrix *d) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !d) return fail(TSG_ERR_ARG, "null argument");
    if (d->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles are not on the device");
    void *ps[] = {d->tile_ptr, d->tile_columnidx, d->tile_rowidx, d->tile_nnz, d->tile_csr_Val,
                  d->tile_csr_Col, d->tile_csr_Ptr, d->mask, d->csc_tile_ptr, d->csc_tile_rowidx,
                  d->csr_tile_nnz, d->csr_mask, d->csr_col_mask, d->csr_col_Ptr, d->csr_col_Val, d->row_nnz_ptr,
                  d->row_ptr, d->row_col, d->row_val, d->val_perm, d->val_perm_col,
                  d->tile_csc_Ptr, d->tile_csc_Row, d->tile_csc_Val};
    for (void *p : ps) pool_free(ctx, p);
    memset(d, 0, sizeof *d);
    return TSG_OK;
}

int tsg_tile_free_host(tsg_tile_matrix *h) {
    if (!h) return TSG_OK;
    if (h->location != TSG_HOST) return fail(TSG_ERR_ARG, "tiles are not on the host");
    void *ps[] = {h->tile_ptr, h->tile_columnidx, h->tile_rowidx, h->tile_nnz, h->tile_csr_Val,
                  h->tile_csr_Col, h->tile_csr_Ptr, h->mask, h->csc_tile_ptr, h->csc_tile_rowidx,
                  h->csr_tile_nnz, h->csr_mask, h->csr_col_mask, h->csr_col_Ptr, h->csr_col_Val, h->row_nnz_ptr,
                  h->row_ptr, h->row_col, h->row_val, h->val_perm, h->val_perm_col,
                  h->tile_csc_Ptr, h->tile_csc_Row, h->tile_csc_Val};
    for (void *p : ps) free(p);
    memset(h, 0, sizeof *h);
    return TSG_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ SpGEMM
namespace tsg {

int check_operands(const tsg_tile_matrix *A, const tsg_tile_matrix *B) {
    if (!A || !B) return fail(TSG_ERR_ARG, "null operand");
    if (A->location != TSG_DEVICE || B->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "operands must be on the device");
    if (!B->is_csc) return fail(TSG_ERR_ARG, "B must be laid out by tsg_build_csc_tiles (is_csc = 0)");
    if (A->is_csc) return fail(TSG_ERR_ARG, "A must be in row-tile order (is_csc = 1)");
    if (A->n != B->m) return fail(TSG_ERR_ARG, "inner dimensions differ: A is %dx%d, B is %dx%d", A->m, A->n, B->m, B->n);
    if (A->dtype != B->dtype) return fail(TSG_ERR_ARG, "A and B value types differ");
    if ((A->tile_size && A->tile_size != 16) || (B->tile_size && B->tile_size != 16))
        return fail(TSG_ERR_UNSUPPORTED, "products need 16 x 16 tiles (tile size %d / %d)", A->tile_size, B->tile_size);
    if (!A->mask) return fail(TSG_ERR_ARG, "A lacks row masks");
    // tile masks and row offsets are read as 16-byte words
    const void *words[] = {A->mask, A->tile_csr_Ptr, B->mask, B->tile_csr_Ptr, B->csr_mask, B->csr_col_mask, B->csr_col_Ptr};
    for (const void *w : words)
        if (!aligned16(w)) return fail(TSG_ERR_ARG, "tile masks and row offsets must start on 16-byte boundaries");
    if (A->tilem > MAX_TILE_ROWS || B->tilen > MAX_TILE_ROWS)
        return fail(TSG_ERR_OVERFLOW, "more than %lld tile-rows or tile-columns", (long long)MAX_TILE_ROWS);
    // the kernels address operand tiles (and B tile-row starts) with int32 ids
    if (A->numtile > INT32_MAX || B->numtile > INT32_MAX)
        return fail(TSG_ERR_OVERFLOW, "operand tiles (%lld / %lld) exceed int32 tile ids", (long long)A->numtile,
                    (long long)B->numtile);
    return TSG_OK;
}

// Reference method: every C tile through the C-tile kernels (the reference's
// three steps with exact counts); the host reads
// numtileC and nnzC to size C (the reference's own host syncs,
// src/spgemm_raw_v2.cu:141, :211; its third, the pair-cache count at :192,
// does not exist here because no pair cache is kept).
int spgemm_plain(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                 const tsg_options *opts, tsg_stats *st) {
    const bool timing = opts && opts->timing;
    hipEvent_t ev[4] = {};
    if (timing)
        for (auto &e : ev) TSG_HIP(hipEventCreate(&e));
    memset(C, 0, sizeof *C);
    C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
    C->dtype = A->dtype; C->location = TSG_DEVICE;
    int rc = TSG_OK;
    int syncs = 0;
    if (timing) (void)hipEventRecord(ev[0], ctx->stream);
    // Step 1
    rc = palloc(ctx, &C->tile_ptr, (size_t)C->tilem + 1);
    if (!rc) rc = launch_structure_count(ctx, A, B, C->tile_ptr);
    if (!rc) rc = scan_inclusive_i64(ctx, C->tile_ptr, (int64_t)C->tilem + 1);
    if (!rc) { rc = read_i64(ctx, C->tile_ptr + C->tilem, &C->numtile); syncs++; }
    if (!rc && C->numtile > INT32_MAX) rc = fail(TSG_ERR_OVERFLOW, "numtileC %lld exceeds int32 tile ids", (long long)C->numtile);
    if (!rc) rc = palloc(ctx, &C->tile_rowidx, (size_t)C->numtile);
    if (!rc) rc = palloc(ctx, &C->tile_columnidx, (size_t)C->numtile);
    if (!rc) rc = launch_structure_fill(ctx, A, B, C->tile_ptr, C->tile_rowidx, C->tile_columnidx);
    if (timing && !rc) (void)hipEventRecord(ev[1], ctx->stream);
    // Step 2
    if (!rc) rc = palloc(ctx, &C->mask, (size_t)C->numtile * 16);
    if (!rc) rc = palloc(ctx, &C->tile_csr_Ptr, (size_t)C->numtile * 16);
    if (!rc) rc = palloc(ctx, &C->tile_nnz, (size_t)C->numtile + 1);
    if (!rc) rc = launch_symbolic(ctx, A, B, C->numtile, C->tile_rowidx, C->tile_columnidx, C->mask,
                                  C->tile_csr_Ptr, C->tile_nnz);
    if (!rc) rc = scan_inclusive_i64(ctx, C->tile_nnz, C->numtile + 1);
    if (!rc) { rc = read_i64(ctx, C->tile_nnz + C->numtile, &C->nnz); syncs++; }
    if (timing && !rc) (void)hipEventRecord(ev[2], ctx->stream);
    // Step 3
    if (!rc) rc = pool_alloc(ctx, &C->tile_csr_Val, (size_t)C->nnz * vsize(C->dtype));
    if (!rc) rc = palloc(ctx, &C->tile_csr_Col, (size_t)C->nnz);
    if (!rc) rc = (opts && opts->numeric_variant == 2) ? launch_numeric_mfma(ctx, A, B, C, opts)
                                                       : launch_numeric(ctx, A, B, C, opts);
    if (timing && !rc) (void)hipEventRecord(ev[3], ctx->stream);
    if (st) {
        st->numtileC = C->numtile;
        st->nnzC = C->nnz;
        st->host_syncs = syncs;
        st->fallback_used = 0;
        st->pred_numtileC = st->pred_nnzC = 0;
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
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    if (rc) {
        tsg_tile_matrix tmp = *C;
        void *ps[] = {tmp.tile_ptr, tmp.tile_columnidx, tmp.tile_rowidx, tmp.tile_nnz, tmp.tile_csr_Val,
                      tmp.tile_csr_Col, tmp.tile_csr_Ptr, tmp.mask};
        for (void *p : ps) pool_free(ctx, p);
        memset(C, 0, sizeof *C);
    }
    return rc;
}

int spgemm_learned(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *st);  // learned.hip

}  // namespace tsg

namespace tsg {
int spgemm_dispatch(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                    const tsg_options *opts, tsg_stats *st, NumericReplay *keep) {
    if (st) memset(st, 0, sizeof *st);
    if (opts && (opts->numeric_variant < 0 || opts->numeric_variant > 3))
        return fail(TSG_ERR_ARG, "numeric_variant %d is not 0-3", opts->numeric_variant);
    if (opts && opts->numeric_variant == 3) {
        if (opts->scheduler == TSG_SCHED_LEARNED)
            return fail(TSG_ERR_ARG, "numeric_variant 3 runs under the plain scheduler");
        if (!(A->tile_csc_Ptr && A->tile_csc_Row && A->tile_csc_Val))
            return fail(TSG_ERR_ARG, "numeric_variant 3 needs A's within-tile CSC (tsg_build_tile_csc)");
    }
    // the matrix-core C-tile kernels: when a tuned configuration selects
    // them, or -- without one -- when both operands' tiles hold 64 entries or
    // more on average (dense pairs, where they win: a band of half-width 48
    // runs 4x faster than on the binned kernels)
    TuneCfg cfg;
    const bool eligible = (!opts || (opts->numeric_variant == 0 && (opts->scheduler != TSG_SCHED_LEARNED || keep))) &&
                          !((A->flags | B->flags) & 1);
    const bool tuned = eligible && tuned_config(ctx, A, B, opts, &cfg);
    const bool dense = A->numtile > 0 && B->numtile > 0 && A->nnz >= 64 * A->numtile && B->nnz >= 64 * B->numtile;
    const int32_t gate = !eligible ? 0 : tuned ? cfg.mfma : dense ? 256 : 0;
    if (gate > 0) {
        tsg_options o = opts ? *opts : tsg_options{};
        o.numeric_variant = 2;
        o.mfma_min_products = gate;
        if (keep) {
            keep->path = 2;
            keep->o = o;
        }
        return spgemm_plain(ctx, A, B, C, &o, st);
    }
    // a kept plan needs the exact pipeline (the learned one re-plans every call)
    if (opts && opts->scheduler == TSG_SCHED_LEARNED && !keep) return spgemm_learned(ctx, A, B, C, opts, st);
    if (opts && opts->numeric_variant >= 1 && opts->numeric_variant <= 3) {
        if (keep) {
            keep->path = opts->numeric_variant;
            keep->o = *opts;
        }
        return spgemm_plain(ctx, A, B, C, opts, st);
    }
    if (keep) {
        keep->path = 0;
        if (opts) keep->o = *opts;
        return spgemm_binned(ctx, A, B, C, opts, st, &keep->p);
    }
    return spgemm_binned(ctx, A, B, C, opts, st);
}
}  // namespace tsg

extern "C" int tsg_spgemm(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                          tsg_tile_matrix *C, const tsg_options *opts, tsg_stats *st) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !C) return fail(TSG_ERR_ARG, "null argument");
    TSG_CHECK(check_operands(A, B));
    DeviceGuard g(ctx->device);
    return spgemm_dispatch(ctx, A, B, C, opts, st);
}
