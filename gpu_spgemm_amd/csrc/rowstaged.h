// rowstaged.h -- host-side interface of the binned pipeline (rowstaged.hip),
// shared with the learned scheduler (learned.hip).
#pragma once
#include <cstddef>

#include "tsg_internal.h"

namespace tsg {

// Index of the segment holding i: largest q < n with seg[q] <= i (seg ascending).
__device__ __forceinline__ int owner_of(const int32_t *seg, int n, int32_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// 16-byte loads of a tile's 16 row pointers (uint8) and 16 row masks (uint16).
// Host-sized bins [b, e) that hold every tile-row between them: the kernels
// take item idx as tile-row idx (nullptr list), which covers the same rows
// once each and skips a dependent load.
inline const int32_t *row_list(const int32_t *rows, const BinRange &br, int32_t tilem) {
    return (br.all || (!br.lo && br.begin == 0 && br.count == tilem)) ? nullptr : rows;
}

__device__ __forceinline__ uint4 load_ptr16(const uint8_t *p, int64_t tile) {
    return reinterpret_cast<const uint4 *>(p)[tile];
}
__device__ __forceinline__ void load_mask16(const uint16_t *m, int64_t tile, uint4 &lo, uint4 &hi) {
    lo = reinterpret_cast<const uint4 *>(m)[tile * 2];
    hi = reinterpret_cast<const uint4 *>(m)[tile * 2 + 1];
}
__device__ __forceinline__ void store_mask16(uint16_t *dst, const uint4 &lo, const uint4 &hi) {
    reinterpret_cast<uint4 *>(dst)[0] = lo;
    reinterpret_cast<uint4 *>(dst)[1] = hi;
}

constexpr int MAX_VAR = 4;               // staged variants (bins R, S, W, M)
constexpr int MAX_BINS = MAX_VAR + 2;    // + C-tile kernels + row-by-row kernels

// LDS capacities of one staged kernel variant; a tile-row is staged by the
// first variant whose capacities hold it.  Other rows run the row-by-row
// kernels when enabled and their A tiles hold at most rg_density entries on
// average, else the C-tile kernels.  Bins: 0..nvar-1 staged, nvar C-tile,
// nvar+1 row-by-row.
struct StageCaps {
    int32_t na, nc, nb, pa, pb;
    // > 0: the products of a scalar row, estimated as (A entries per scalar
    // row) x (B entries per reached B row) = (pa / 16) (pb / (16 na)), at most this
    int32_t prow = 0;
};
struct BinTable {
    StageCaps v[MAX_VAR];
    int32_t nvar;
    int32_t rg;          // 1: the row-by-row bin exists
    int32_t rg_density;  // A entries per A tile at or below which a heavy row goes row-by-row
    int32_t rl_any;      // 1: bin R runs its own kernels whatever its size (var_mask bit 4)
    int32_t any_order;   // 1: the row-by-row numeric kernel may add products in any order (sum_order 1)
};
constexpr int RG_DENSITY_DEFAULT = 64;

// Device-side flags and counters of one binned product, cleared together by
// one small kernel and read back together with nnz(C).
struct PlanFlags {
    int32_t cap;                    // a staged row exceeded its variant's capacity (or C's value capacity)
    int32_t pad0;
    int64_t nnz;                    // nnz(C), written by the tile_nnz scan
    int32_t heavy_rows;             // length of the heavy-row list built by k_row_stats
    int32_t pad1;
    unsigned long long heavy_work;  // work counter of the heavy rows' Step-1 count
    unsigned long long pred_tiles;  // learned scheduler: summed per-row predictions (C tiles, C nonzeros)
    unsigned long long pred_nnz;
};
static_assert(offsetof(PlanFlags, nnz) == 8 && offsetof(PlanFlags, heavy_rows) == 16 &&
                  offsetof(PlanFlags, heavy_work) == 24 && offsetof(PlanFlags, pred_nnz) == 40 &&
                  sizeof(PlanFlags) == 48,
              "PlanFlags layout: the heavy-row kernels take heavy_rows and heavy_work as {int32, pad, uint64}");
constexpr int PLAN_FLAG_WORDS = (int)(sizeof(PlanFlags) / sizeof(int32_t));

// Per-call working set: A->B links, per-row bins, bin-ordered row lists.
struct Plan {
    int32_t *rows = nullptr;     // row ids grouped by bin, ascending in each bin
    uint8_t *bin = nullptr;      // bin of each tile-row
    int64_t *counts = nullptr;   // (bin, block) counts, scanned
    int64_t *hdr = nullptr;      // [total0, total1, bin offsets...] for one host read
    PlanFlags *flags = nullptr;  // device
    int4 *link = nullptr;        // per A tile: B tile-row start, tiles, payload count
    int64_t *lpay = nullptr;     // per A tile: B tile-row payload start
    int64_t *rgw = nullptr;      // row-by-row bin: symbolic work-item offsets + counters (tilem + 3)
    int64_t *rgn = nullptr;      // row-by-row bin: numeric work-item offsets (tilem + 1)
    int64_t *rgc = nullptr;      // row-by-row bin: numeric cursor area (allocated with rgn's offsets)
    HeavyBitmaps hb;             // heavy rows' union bitmaps, Step-1 count -> fill (freed after the fill)
    int64_t bin_off[MAX_BINS + 1] = {};
    int nbins = 0;
    int nvar = 0;     // staged bins; bin nvar = C-tile kernels
    bool rg = false;  // bin nvar + 1 = row-by-row kernels
    // A captured graph: the bin offsets (bin_off) live only on the device, at
    // d_bins[0..nbins]; the launches for bin b cover bin_cap[b] rows and take
    // the bin's range from there.  Null: host-known bins.
    const int64_t *d_bins = nullptr;
    int64_t bin_cap[MAX_BINS] = {};
    bool rl_any = false;  // BinTable::rl_any
    // a device-sized product whose capacities leave one bin open with room for
    // every tile-row: that bin holds them all (else the capacity check empties
    // every bin), so its row list is the identity
    int all_bin = -1;
    bool any_order = false;  // BinTable::any_order
    // device address of the capacity flag (kernels take it as int32_t *)
    int32_t *overflow() const { return flags ? &flags->cap : nullptr; }
};

bool staged_ok(const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts);
// The row-by-row kernels need the scalar CSR of A and B and no duplicate columns.
bool rg_ok(const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts);
// The row-lane numeric kernel (rownum.hip): needs both operands' scalar CSR and no repeated columns.
bool num_rows_ok(const tsg_tile_matrix *A, const tsg_tile_matrix *B);
// var_mask bits 0-2 admit staged variants S, W, M and bit 3 variant R (a lane
// per scalar row; only when row_lanes); bit 4 gives R's rows their own kernels
// however few they are; rg_density 0 disables the row-by-row bin.
BinTable bin_table(bool ok, bool rg = false, int rg_density = RG_DENSITY_DEFAULT, int var_mask = 15,
                   bool row_lanes = false);
// The bin table of a product under its options and the context's tuned configuration.
BinTable bin_table_for(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts);

// Autotuned configuration (autotune.hip).
struct TuneCfg {
    int32_t var_mask = 15;
    int32_t rg_density = RG_DENSITY_DEFAULT;
    int32_t mfma = 0;  // > 0: matrix-core C-tile kernels with this density gate
};
// Class signature of a product from its operands' sizes.
uint64_t tune_signature(int64_t tilemA, int64_t numtileA, int64_t nnzA, int64_t tilemB, int64_t numtileB,
                        int32_t dtype);
// The configuration to use: a forced one (while tuning), else the stored one
// for the signature unless opts->tuning == 1; false = defaults.
bool tuned_config(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts,
                  TuneCfg *cfg);
void free_plan(Context *ctx, Plan &p);
// Frees the heavy rows' union bitmaps (after the Step-1 fill).
void free_heavy_bitmaps(Context *ctx, Plan &p);
// Allocates the plan's arrays and computes the A->B links.
// zero_flags: clear the capacity flag here (the exact path clears it with its
// heavy-row counters in one launch).
int plan_setup(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const BinTable &bt, Plan &p,
               bool zero_flags = true);
// Partitions rows by p.bin and gathers hdr = [*total0, *total1 (or 0), bin offsets].
int plan_bin(Context *ctx, Plan &p, int32_t tilem, const int64_t *total0, const int64_t *total1);
// Reads hdr to the host: totals and bin offsets (one blocking read).
int plan_read(Context *ctx, Plan &p, int64_t *total0, int64_t *total1);
// A->B links, Step-1 count (with the per-row statistics when stats_out, RS_NSTAT
// int32 a tile-row) and the scan of C->tile_ptr; clears p.flags.
int plan_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
               const tsg_options *opts, Plan &p, int32_t *stats_out, bool scan = true);
// The Step-1 tail: scan of tile_ptrC, binning and the header read (numtileC,
// *total1 when given, bin offsets into p.bin_off).  plan_tail_fused(tilem):
// two launches and plan_count must have left tile_ptrC unscanned (scan =
// false); else the separate scan, binning and read (plan_count scanned).
bool plan_tail_fused(int32_t tilem);
// Capacities of a device-sized product's bins.
struct BinCaps {
    int64_t v[MAX_BINS];
};
// The same without the host read (the header stays in p.hdr on the device);
// with caps, the capacity check of the device-sized pipeline (numtileC against
// cap_tiles, each bin's rows against its cover: past any of them the flag is
// set and every bin emptied) and numtileC also written to *numtile_out.
int plan_tail_device(Context *ctx, Plan &p, int32_t tilem, int64_t *tile_ptrC, const int64_t *total1,
                     const BinCaps *caps = nullptr, int64_t cap_tiles = 0, int64_t *numtile_out = nullptr);
int plan_tail(Context *ctx, Plan &p, int32_t tilem, int64_t *tile_ptrC, const int64_t *total1, int64_t *numtile,
              int64_t *t1);
int run_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                 Plan &p, const SlotLayout &sl);
// Bits of tsg_stats.numeric_kernels for the plan's host-known bins.
int numeric_kernel_mask(const Plan &p);
int run_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                const Plan &p, const SlotLayout &sl);
// Set-up of run_numeric that needs C's scanned value offsets (call once after the scan).
int prepare_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *C, Plan &p, const SlotLayout &sl);
void free_c(Context *ctx, tsg_tile_matrix *C);
int launch_num_rows(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                    const int32_t *rows, const BinRange &br, const SlotLayout &sl);
// Symbolic step of the row-lane bin (rownum.hip): 16 lanes per tile-row.
int launch_sym_rows(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                    const int32_t *rows, const BinRange &br, const int4 *link, const SlotLayout &sl);
BinRange bin_range(const Plan &p, int b, int e);

// Capacities of a device-sized product (a captured graph): C's tile and value
// arrays and the rows each bin's launches cover.
struct DeviceCaps {
    int64_t tiles = 0, nnz = 0;
    int64_t bins[MAX_BINS] = {};
};
// The binned pipeline with no host reads (Context::device_sizes must be set):
// sizes stay on the device, launches cover `caps`, C's arrays are allocated
// at the capacities (C->numtile, C->nnz hold them), and a product past any
// capacity sets p.overflow() and leaves C undefined.  numtileC ends in
// p.hdr[0] and nnzC in p.flags->nnz.  The plan and C stay allocated.
// ev (optional): two events recorded after C's arrays are allocated and after
// the symbolic step.  flags_seq (optional): the plan's flags (capacity flag,
// nnz(C), numtileC in pred_tiles) published before the numeric step is queued;
// wait with publish_wait(ctx, *flags_seq, PLAN_FLAG_WORDS, ...).
int spgemm_binned_device(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                         const tsg_options *opts, const DeviceCaps &caps, Plan &p, hipEvent_t *ev = nullptr,
                         int64_t *flags_seq = nullptr);

// keep != nullptr: the plan (row bins, links) is handed over instead of freed,
// for re-running the numeric step.
int spgemm_binned(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                  const tsg_options *opts, tsg_stats *st, Plan *keep = nullptr);

// What re-running a product's numeric step needs (plan.hip): the kernels the
// product used and, for the binned pipeline, its plan.
struct NumericReplay {
    int path = 0;  // 0 binned pipeline, 1 C-tile kernels, 2 C-tile kernels with matrix cores, 3 outer-product C-tile
    Plan p;
    tsg_options o{};
};
int spgemm_dispatch(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                    const tsg_options *opts, tsg_stats *st, NumericReplay *keep = nullptr);
// One contiguous copy of the batch kernels (batch.hip): count elements of
// `kind` from src to dst, adding `add` to integer kinds.  kind: 0 = bytes,
// 1 = 16-byte units, 2 = int32 + add, 3 = int64 + add, 4 = 8-byte units,
// 5 = 4-byte units, 6 = int64 value `add` (no source).
struct CatJob {
    const void *src;
    void *dst;
    int64_t count;
    int64_t add;
    int32_t kind;
    int32_t pad;
};

// What re-running a batch's numeric step needs: the block-diagonal operands
// and product, and the value copies into and out of them (device job tables).
struct BatchKeep {
    tsg_tile_matrix BA{}, BB{}, BC{};
    NumericReplay nr;
    CatJob *vin = nullptr, *vout = nullptr;
    int64_t nvin = 0, nvout = 0, vin_max = 0, vout_max = 0;
};
int spgemm_batched(Context *ctx, int32_t n, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *stats, BatchKeep *keep = nullptr);
void free_batch_keep(Context *ctx, BatchKeep &k);
// Launches the copies of a job table (kernels only).
int launch_jobs(Context *ctx, const CatJob *dj, int64_t n, int64_t maxc);

// Launches the numeric step again (kernels only: capturable in a graph).
int replay_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                   NumericReplay &r);

}  // namespace tsg
