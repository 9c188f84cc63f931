// noalloc.hip -- C = A*B without any device allocation.
//
// No reference counterpart.  For callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): the product runs the exact binned pipeline with every work
// array (A->B links, row bins and lists, scan partials, heavy-row lists) taken
// from `scratch`, a static device array of NOALLOC_SCRATCH_BYTES, and C
// written into the caller's arrays.  Each GPU holds its own copy of the array
// for the life of the process; every call on one GPU shares it, so calls that
// overlap on one GPU race.  Nothing in it needs to start at zero: the
// pipeline clears its own flags and counters by kernels.
#include <mutex>

#include "rowstaged.h"

namespace tsg {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on each device (the array's name in host code is
// only a stand-in): looked up once per device, with that device current.
static std::mutex g_scratch_mu;
static std::map<int, uint8_t *> g_scratch_ptr;

int noalloc_setup(Context *ctx) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(ctx->device)) return TSG_OK;
    int cur = -1;
    TSG_HIP(hipGetDevice(&cur));
    if (cur != ctx->device) return fail(TSG_ERR_ARG, "scratch set-up needs device %d current (is %d)", ctx->device, cur);
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(TSG_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", ctx->device,
                    hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = static_cast<uint8_t *>(ptr);
    return TSG_OK;
}

static uint8_t *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}

// Work-array bytes of the binned pipeline on these operands, without the
// row-by-row bin and the heavy rows' bitmaps (neither runs in this mode); an
// estimate checked before any launch, the arena's own bound check is exact.
static size_t scratch_need(const tsg_tile_matrix *A) {
    const size_t tilem = (size_t)A->tilem, nta = (size_t)A->numtile;
    const size_t blocks = tilem / 1024 + 1;
    size_t b = 0;
    auto add = [&](size_t bytes) { b += (bytes + 255) & ~size_t(255); };
    add(tilem * 4);                      // row lists
    add(tilem);                          // bins
    add((MAX_BINS * blocks + 1) * 8);    // bin counts
    add((MAX_BINS + 3) * 8);             // host-read header
    add(sizeof(PlanFlags));
    add(nta * 16);                       // A->B links
    add(nta * 8);                        // link payload offsets
    add(tilem * 4);                      // heavy-row list
    add(((tilem + 1) / 2048 + 2) * 8);   // scan partials (tile_ptrC)
    return b;
}

}  // namespace tsg

using namespace tsg;

extern "C" int tsg_spgemm_noalloc(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                                  tsg_tile_matrix *C, int64_t cap_tiles, int64_t cap_nnz, const tsg_options *opts,
                                  tsg_stats *st) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !C) return fail(TSG_ERR_ARG, "null argument");
    TSG_CHECK(check_operands(A, B));
    if (opts && (opts->scheduler != TSG_SCHED_PLAIN || opts->numeric_variant != 0))
        return fail(TSG_ERR_ARG, "tsg_spgemm_noalloc runs the plain scheduler's binned pipeline (numeric_variant 0)");
    if (cap_tiles < 0 || cap_nnz < 0) return fail(TSG_ERR_ARG, "negative capacity");
    if (!C->tile_ptr || (cap_tiles > 0 && (!C->tile_columnidx || !C->tile_rowidx || !C->mask || !C->tile_csr_Ptr)) ||
        !C->tile_nnz || (cap_nnz > 0 && (!C->tile_csr_Val || !C->tile_csr_Col)))
        return fail(TSG_ERR_ARG, "C must hold the caller's device arrays: tile_ptr [tilemA+1], tile_rowidx, "
                                 "tile_columnidx [cap_tiles], tile_nnz [cap_tiles+1], mask, tile_csr_Ptr "
                                 "[cap_tiles*16], tile_csr_Val, tile_csr_Col [cap_nnz]");
    const size_t need = scratch_need(A);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(TSG_ERR_OVERFLOW, "A (%lld tiles, %d tile-rows) needs about %zu bytes of work arrays; the "
                    "no-alloc scratch holds %zu", (long long)A->numtile, A->tilem, need, NOALLOC_SCRATCH_BYTES);
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) TSG_HIP(hipSetDevice(ctx->device));
    int rc = noalloc_setup(ctx);  // a no-op after the context's creation did it
    uint8_t *base = rc ? nullptr : scratch_ptr(ctx->device);
    if (!rc && !base) rc = fail(TSG_ERR_HIP, "no scratch address for device %d", ctx->device);
    if (!rc) {
        Arena arena;
        arena.base = base;
        arena.size = NOALLOC_SCRATCH_BYTES;
        FixedOut fo;
        fo.c = *C;
        fo.cap_tiles = cap_tiles;
        fo.cap_nnz = cap_nnz;
        tsg_options o = opts ? *opts : tsg_options{};
        o.timing = 0;  // timing events are runtime allocations
        ctx->arena = &arena;
        ctx->fixed_out = &fo;
        rc = spgemm_binned(ctx, A, B, C, &o, st);
        ctx->arena = nullptr;
        ctx->fixed_out = nullptr;
        if (rc) *C = fo.c;  // on failure C keeps the caller's arrays (spgemm_binned cleared it)
    }
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    return rc;
}
