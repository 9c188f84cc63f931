// tsg_internal.h -- shared internals of libtsg: context, device pool, error
// reporting and wave64 device helpers for gfx950.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "tsg.h"

namespace tsg {

constexpr int TS = 16;
constexpr int WAVE = 64;

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define TSG_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::tsg::fail(TSG_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                               hipGetErrorString(e_), __FILE__, __LINE__);             \
    } while (0)

#define TSG_CHECK(expr)                 \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != TSG_OK) return rc_;  \
    } while (0)

// ---------------------------------------------------------------- device pool
// Caching allocator: freed blocks are kept per exact rounded size and reused,
// so repeated products of one shape make no hipMalloc after the first.
// A view is an address inside a live block handed out as an array of its
// own (a batch's outputs lie inside the one product that computed them);
// the block is recycled when its owner and every view have been freed.
struct Pool {
    std::multimap<size_t, void *> free_blocks;
    std::map<void *, size_t> live;
    std::multimap<void *, void *> views;  // view address -> the block it lies in
    std::map<void *, int64_t> refs;       // block with views -> holders (owner + views)
    int64_t in_use = 0, cached = 0;
    static size_t round(size_t b) {
        if (b < 256) return 256;
        // round to 1/8 of the power of two so reuse across close sizes works
        size_t p = 1;
        while (p < b) p <<= 1;
        size_t step = p >> 3;
        return (b + step - 1) / step * step;
    }
};

// A bump allocator over one fixed device region (the no-alloc entry point's
// static scratch): while a context holds one, pool_alloc takes its arrays from
// it and pool_free gives nothing back; the region is reset per call.
struct Arena {
    uint8_t *base = nullptr;
    size_t size = 0, top = 0;
};
// Caller-owned output arrays of C (the no-alloc entry point): the pipeline
// writes C into them instead of allocating, after checking the capacities.
struct FixedOut {
    tsg_tile_matrix c{};
    int64_t cap_tiles = 0, cap_nnz = 0;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;      // where every call's work is queued (own_stream or the caller's)
    hipStream_t own_stream = nullptr;  // created with the context
    hipDeviceProp_t prop{};
    Pool pool;
    std::mutex mu;
    // small pinned host staging for scalar reads
    int64_t *pinned = nullptr;
    // host reads of a few device words (read_words): coherent pinned slots the
    // device writes with a sequence number the host waits for
    int32_t *pub = nullptr;
    int32_t *pub_dev = nullptr;
    int64_t pub_seq = 0;
    // learned scheduler state lives in learned.hip
    void *learned = nullptr;
    // autotuner state lives in autotune.hip
    void *tuner = nullptr;
    // set only inside tsg_spgemm_noalloc
    Arena *arena = nullptr;
    const FixedOut *fixed_out = nullptr;
    // set while a batch graph is built (the device-sized pipeline may run);
    // with `deferred` set, pool_free defers into it (a captured graph goes on
    // using the arrays it ran with)
    bool device_sizes = false;
    std::vector<void *> *deferred = nullptr;
};

int pool_alloc(Context *ctx, void **p, size_t bytes);
int pool_free(Context *ctx, void *p);
// *out = block + offset (offset < the block's size), freed with pool_free like
// any array; the block stays allocated until its owner and all views are freed.
int pool_view(Context *ctx, void *block, size_t offset, void **out);
template <typename T>
inline int palloc(Context *ctx, T **p, size_t count) {
    return pool_alloc(ctx, reinterpret_cast<void **>(p), count * sizeof(T));
}

inline size_t vsize(int dtype) { return dtype == TSG_F32 ? 4 : 8; }

// A dispatch's work-item count (blocks x threads) is a 32-bit quantity: kernels
// over index spaces that can exceed it run a capped grid and stride.
constexpr int64_t MAX_GRID_BLOCKS = int64_t(1) << 22;
inline unsigned grid_capped(int64_t items, int per_block) {
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    return (unsigned)(b < MAX_GRID_BLOCKS ? b : MAX_GRID_BLOCKS);
}
// Tile-rows at or above this would make one wave per tile-row reach 2^32
// work-items; launches of one workgroup per row go in chunks of ROW_CHUNK.
constexpr int64_t MAX_TILE_ROWS = (int64_t(1) << 26) - 64;
constexpr int64_t ROW_CHUNK = int64_t(1) << 23;

// ---------------------------------------------------------------- scans (scan.hip)
// In-place inclusive scan of n int64 values.  With data[0] = 0 and counts in
// data[1..n-1], this leaves exclusive offsets (the reference's CUB
// ExclusiveSum, src/spgemm_raw_v2.cu:131-138, 202-208, widened to int64).
// zero_first: data[0] is taken as 0 whatever it holds (saves clearing it);
// total_out (device, may be null): also receives data[n-1] after the scan.
// dlast (device, may be null): the scan covers items [0, min(n, *dlast + 1))
// only -- a length known on the device alone (a captured graph); the partial
// sums of the blocks past it are zero.
int scan_inclusive_i64(Context *ctx, int64_t *data, int64_t n, bool zero_first = false,
                       int64_t *total_out = nullptr, const int64_t *dlast = nullptr);
// The same scan (n >= 1), whose last block also publishes the words
// (*word, 0, total as int64) for publish_wait(ctx, *seq, 4, ...): the read of a
// total without a launch of its own.
int scan_publish_i64(Context *ctx, int64_t *data, int64_t n, bool zero_first, int64_t *total_out,
                     const int32_t *word, int64_t *seq);
// Zeroes a[0..na) and b[0..nb) (at most 64 each) with one small kernel on the context's stream.
int zero_small(Context *ctx, int32_t *a, int na, int64_t *b, int nb);
// Reads n 32-bit words (n <= PUB_WORDS) of device memory into out after the
// work queued before it: one small kernel writes them to coherent pinned
// memory with a sequence number, the host spins on that number (blocking;
// counts as a host sync).  Replaces a device-to-host copy and a stream sync.
constexpr int PUB_WORDS = 64;
int read_words(Context *ctx, const void *dsrc, int n, void *out);
// The same protocol for a kernel that publishes by itself: it writes n words to
// ctx->pub_dev, a system-scope fence, then publish_seq()'s number at word
// PUB_WORDS; publish_wait spins for that number and copies the words to out.
int64_t publish_seq(Context *ctx);
int publish_wait(Context *ctx, int64_t seq, int n, void *out);
// read_words in two halves: the publishing kernel queued now (its number in
// *seq), the wait (publish_wait) later, after more work has been queued.
int publish_start(Context *ctx, const void *dsrc, int n, int64_t *seq);
// Reads one int64 from device memory (blocking; counts as a host sync).
int read_i64(Context *ctx, const int64_t *dptr, int64_t *out);
int read_i32(Context *ctx, const int32_t *dptr, int32_t *out);

// ---------------------------------------------------------------- kernels entry (host side)
int launch_structure_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                           int64_t *tile_ptrC);
int launch_structure_fill(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                          const int64_t *tile_ptrC, int32_t *rowidxC, int32_t *colidxC);
int launch_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                    int64_t numtileC, const int32_t *rowidxC, const int32_t *colidxC,
                    uint16_t *maskC, uint8_t *ptrRowC, int64_t *tile_nnzC);
// Slot layout of C used by the learned scheduler (null members: exact layout).
struct SlotLayout {
    int32_t *row_nc = nullptr;     // actual tiles of each tile-row
    int32_t *overflow = nullptr;   // set when a row or C's values outgrow their capacity
    int64_t val_cap = INT64_MAX;   // capacity of C's value arrays
    // values below this already hold their final results (a numeric re-run
    // after a short capacity computes only the rows or tiles ending past it)
    int64_t val_done = -1;
};
// Per-tile-row statistics the Step-1 count can emit (int32 each, clamped):
// A tiles, A nonzeros, B tiles reached, B nonzeros reached, C tiles, span of
// the reached B tile columns, longest reached B tile-row, summed spans of the
// reached B tile-rows, reached B tile-rows that hold tiles.
enum RowStat { RS_NA = 0, RS_PA, RS_NB, RS_PB, RS_COUNT, RS_SPAN, RS_MAXLEN, RS_SUMSPAN, RS_NONEMPTY, RS_NSTAT = 10 };
// Tile-rows reaching more than HEAVY_NB B tiles are counted and filled by a
// workgroup each (structure.hip) instead of one wave.
constexpr int64_t HEAVY_NB = 8192;
// B tile-row lists longer than this are walked by a whole workgroup (or
// OR-ed as bitmaps); tsg_tile_matrix.flags bit 1 marks a matrix with such rows.
constexpr int HEAVY_LIST = 64;
// Words of union bitmaps kept from the heavy rows' count for their fill (256 MB).
constexpr int64_t HEAVY_BITMAP_WORDS = 64ll << 20;
// Union bitmaps the heavy rows' Step-1 count leaves for their fill: the
// bitmap of tile-row bi (bit i = tile column cmin + i, cmin its first
// reachable column) starts at words[off[bi]], or off[bi] = -1 when it did not
// fit in cap words (the fill then rebuilds it).  Null members: none kept.
struct HeavyBitmaps {
    uint32_t *words = nullptr;
    int64_t cap = 0;
    unsigned long long *top = nullptr;  // words handed out (zeroed before the count)
    int64_t *off = nullptr;             // [tilem]
    // Column bitmaps of B's long, dense tile-rows (more than HEAVY_LIST tiles,
    // at least two per word): row k's words are bwords[boff[k] .. boff[k+1]),
    // bit c & 31 of word (c >> 5) - (first column >> 5); boff[k+1] == boff[k]
    // for the other rows.  Null: walk every list.
    uint32_t *bwords = nullptr;
    int64_t *boff = nullptr;  // [tilemB+1]
};
// Build the B tile-row bitmaps of hb (bwords sized numtileB / 2 + 1 words).
int build_brow_bitmaps(Context *ctx, const tsg_tile_matrix *B, HeavyBitmaps &hb);
// C tile counts of the heavy rows in hlist (hc = {row count, pad, 64-bit work
// counter}, zeroed before the list is built), written to tile_ptrC[bi + 1].
int launch_heavy_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, int64_t *tile_ptrC,
                       const int32_t *hlist, int32_t *hc, int32_t *stats, const HeavyBitmaps &hb);
// The rows [begin, begin + count) of a bin's row list; in a captured graph the
// range is read on the device from lo/hi and count is an upper bound.
struct BinRange {
    int64_t begin = 0, count = 0;
    const int64_t *lo = nullptr, *hi = nullptr;
    bool all = false;  // the range's rows are every tile-row in order (device range known to hold them all)
};
inline BinRange host_range(int64_t begin, int64_t count) {
    BinRange r;
    r.begin = begin;
    r.count = count;
    return r;
}
int launch_structure_fill_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                               const tsg_tile_matrix *C, const int32_t *rows, const BinRange &br,
                               const SlotLayout &sl, const HeavyBitmaps &hb = HeavyBitmaps{});
int launch_symbolic_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const int32_t *rows,
                         const BinRange &br, const tsg_tile_matrix *C, const SlotLayout &sl);
int launch_numeric_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                        const int32_t *rows, const BinRange &br, const SlotLayout &sl);
void learned_release(Context *ctx);  // learned.hip
void tuner_release(Context *ctx);    // autotune.hip
// Row-by-row kernels for sparse heavy tile-rows (heavy.hip).  woff holds
// count + 3 int64: the rows' work-item offsets (written by the symbolic
// launch, read by the numeric one) and two work counters.
int launch_rg_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                       const int32_t *rows, const BinRange &br, int64_t *woff, int64_t **gcur, const SlotLayout &sl);
// After C's value offsets are scanned: the rows' numeric work-item offsets
// (count + 1 int64 in woffn).
// *gcur: the numeric kernel's per-workgroup cursor area, allocated on first use.
int prepare_rg_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *C, const int32_t *rows,
                       const BinRange &br, int64_t *woff, int64_t *woffn, int64_t **gcur, const int64_t *nnzC,
                       const SlotLayout &sl);
// any_order: k_rgw_numeric adds the staged products of a batch of A entries with LDS
// atomics, all lanes at once, instead of entry by entry in ascending k (sum_order 1).
int launch_rg_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                      const int32_t *rows, const BinRange &br, int64_t *woff, const int64_t *woffn,
                      int64_t *gcur, const SlotLayout &sl, bool any_order = false);
int check_operands(const tsg_tile_matrix *A, const tsg_tile_matrix *B);  // api.hip
inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }
// Looks up the no-alloc scratch's device address on ctx's device (current).
int noalloc_setup(Context *ctx);  // noalloc.hip
int launch_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                   tsg_tile_matrix *C, const tsg_options *opts);
// Matrix-core variant of launch_numeric (mfma.hip).
int launch_numeric_mfma(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                        const tsg_options *opts);

}  // namespace tsg

// ---------------------------------------------------------------- device helpers
#if defined(__HIPCC__)
namespace tsg {

// Orders LDS traffic between lanes of one wavefront (no hardware barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of a 64-bit ballot below this lane.
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Wave-wide inclusive sum (64 lanes).
template <typename T>
__device__ __forceinline__ T wave_incl_sum(T v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T o = __shfl_up(v, d, WAVE);
        if (lane_id() >= d) v += o;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) { T o = __shfl_xor(v, d, WAVE); v = o < v ? o : v; }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) { T o = __shfl_xor(v, d, WAVE); v = o > v ? o : v; }
    return v;
}

// Wave64 inclusive sum of 32-bit integers with DPP lane moves (no LDS
// traffic): Hillis-Steele over rows of 16 (row_shr 1,2,4,8, zero shifted in),
// then row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.
__device__ __forceinline__ int dpp_incl_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
    return v;
}

// Wave64 inclusive maximum with the same DPP pattern (a lane with no source
// keeps its own value).
__device__ __forceinline__ int dpp_incl_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
    return v;
}

// Value of lane 63 (a wave-uniform scalar).
__device__ __forceinline__ int last_lane(int v) { return __builtin_amdgcn_readlane(v, WAVE - 1); }

// 16-lane group inclusive sum (groups are aligned lane ranges of 16, all
// lanes of a group active).
template <typename T>
__device__ __forceinline__ T group16_incl_sum(T v) {
    int l = lane_id() & 15;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
        T o = __shfl_up(v, d, 16);
        if (l >= d) v += o;
    }
    return v;
}
// int: a 16-lane group is one DPP row, so row_shr 1, 2, 4, 8 (zero shifted in)
// does it without LDS traffic.
template <>
__device__ __forceinline__ int group16_incl_sum<int>(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);
    return v;
}

// Maximum over a 16-lane group (one DPP row), returned to every lane of it.
__device__ __forceinline__ int group16_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xF, 0xF, false));
    return __shfl(v, 15, 16);
}

template <typename T>
__device__ __forceinline__ T group16_sum(T v) {
#pragma unroll
    for (int d = 8; d > 0; d >>= 1) v += __shfl_xor(v, d, 16);
    return v;
}

// Rows [row_begin, row_begin + nrows) of a bin's row list; in a captured graph
// the bin's offsets live on the device only: then [*dlo, *dhi), cut to the
// host's nrows (the rows the grid covers; a longer bin is flagged by the
// pipeline's capacity check).
__device__ __forceinline__ void bin_rows(const int64_t *dlo, const int64_t *dhi, int64_t &row_begin, int64_t &nrows) {
    if (dlo) {
        row_begin = *dlo;
        nrows = min(nrows, *dhi - row_begin);
    }
}
// Item idx of a bin's row list (host range, or the device range dlo/dhi): its
// tile-row, or -1 past the list's end.  rows == nullptr: the list is every
// tile-row in order (item idx is tile-row row_begin + idx).
__device__ __forceinline__ int32_t bin_row(const int32_t *rows, int64_t row_begin, int64_t nrows,
                                           const int64_t *dlo, const int64_t *dhi, int64_t idx) {
    bin_rows(dlo, dhi, row_begin, nrows);
    if (idx >= nrows) return -1;
    return rows ? rows[row_begin + idx] : (int32_t)(row_begin + idx);
}

// Tile range [s,e) of C tile-row bi.  Exact layout: ptr is tile_ptrC.  Slot
// layout (learned scheduler): ptr holds each row's padded slot offsets and
// row_nc the row's actual tile count.
__device__ __forceinline__ void c_row_range(const int64_t *ptr, const int32_t *row_nc, int32_t bi, int64_t &s,
                                            int64_t &e) {
    s = ptr[bi];
    e = row_nc ? s + row_nc[bi] : ptr[bi + 1];
}

// First index in [lo,hi) of sorted arr with arr[i] >= key.
template <typename I, typename T>
__device__ __forceinline__ I lower_bound(const T *__restrict__ arr, I lo, I hi, T key) {
    while (lo < hi) {
        I mid = lo + ((hi - lo) >> 1);
        if (arr[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

}  // namespace tsg
#endif
