// heavy.hip -- row-by-row product for heavy tile-rows whose tiles are sparse.
//
// A tile-row that no staged variant holds goes to a workgroup-per-row path.
// When its A tiles hold few entries each (R-MAT, Erdos-Renyi: ~1 entry per
// tile), the C-tile method (src/spgemm_raw_kernels.cuh:105-202,
// src/spgemm_raw_kernels_v2.cuh:10-139), which intersects the A tile-row with
// a B tile-column for every C tile, spends its time on tile pairs that hold no
// product.  These kernels instead walk the products themselves, from the
// scalar CSR copies of A and B: for each row i of the tile-row, each A entry
// (i, k) in ascending k, each B entry (k, j).
//
// Work is split into windows of RGW_COLS consecutive C tile columns, counted
// from the tile-row's first C tile column (its C tile columns are already
// written by the Step-1 fill).  A work item is a run of consecutive windows
// of one tile-row -- about RGW_ITEM_TILES C tiles' worth, at most RGW_GMAX
// windows -- and a persistent grid takes items from a device counter, so the
// few very heavy tile-rows of a power-law matrix spread over many workgroups.
// Inside a window everything lives in LDS:
//
//   k_rgw_symbolic: the window's row masks (RGW_COLS tiles x 16 rows, dense
//     by tile column), set by an LDS atomic OR per product; then the window's
//     C tiles get their masks, row offsets and value counts.  A column is
//     present exactly when some product reaches it, which is the reference's
//     OR over tile pairs.  Long B row segments are walked by the whole
//     workgroup.
//   k_rgw_numeric: a sub-window of C tiles (as many as fit RGN_BYTES of LDS
//     with their values, within RGN_COLS tile columns): the values, a map tile
//     column -> sub-window tile, and each tile's row masks, row offsets and
//     value offset (copied from C's arrays).  The waves
//     take the tile-row's 16 scalar rows in turn; for its row a wave loads 64
//     A entries at a time (a lane per entry: the entry's B segment in the
//     sub-window, its first products' B values and output slots), then adds
//     the entries' products one entry after the other in ascending k -- the
//     owning lane its staged products, the whole wave the rest of a long
//     segment.  The outputs of one A entry are distinct and the entries follow
//     in k order, so each output sums its products in ascending k with one
//     FMA each -- the order of the tiled method (tile pairs ascending, entries
//     ascending within a pair) -- and values are bit-identical to the other
//     kernels.  With sum_order 1 (template flag ANY) the staged products of a
//     batch are instead added by all lanes at once with LDS atomics, in any
//     order (values equal to rounding, not bit for bit).  Numeric work items
//     are ranges of a tile-row's values (about eight sub-windows' worth), so a
//     heavy tile-row spreads over many workgroups; an entry's B position
//     carries from one sub-window to the next within an item (the first
//     RGN_CUR entries of a tile-row in LDS, the next RGN_GCUR in a global area
//     per workgroup), so only an item's first sub-window searches B.
#include <algorithm>
#include <climits>

#include "rowstaged.h"

namespace tsg {

constexpr int RGW_THREADS = 256;       // 16 scalar rows x 16 lanes
constexpr int RGW_COLS = 1024;         // C tile columns per window
constexpr int RGW_ITEM_TILES = 4096;   // C tiles per work item (aim, at most; fewer when C is small)
constexpr int RGW_GMAX = 64;           // windows per work item (cap)
constexpr int RGN_COLS = 4096;         // tile-column span of a numeric sub-window (LDS map)
constexpr int RGN_BYTES = 36864;       // LDS bytes of a numeric sub-window: RGN_TILE_BYTES a tile + its values
constexpr int RGN_CUR = 512;           // A entries of a tile-row whose B positions carry over sub-windows in LDS
constexpr int64_t RGN_GCUR = 262144;   // and up to this many more, in a global area per workgroup
constexpr int64_t RGS_GCUR = 65536;    // A entries of a tile-row whose B positions carry over symbolic windows
constexpr int RGN_TILE_BYTES = 52;     // 16 row masks, 16 row offsets and the tile's value offset
constexpr int RGW_LONG = 64;           // B segments longer than this are walked by the workgroup
constexpr int RGW_ILP = 4;             // A entries per thread per symbolic pass (independent loads)
constexpr int RGW_PF = 16;             // B entries loaded at once by one symbolic thread
constexpr int RGW_PF0 = 4;             // of which the first stage
constexpr int RGN_PF = 8;              // B entries a numeric lane stages for its A entry (VGPRs)
constexpr int RGN_FILL = (RGN_BYTES / RGN_TILE_BYTES + RGW_THREADS - 1) / RGW_THREADS;  // sub-window tiles per thread

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // 16-byte piece

__device__ __forceinline__ uint32_t rgw_below(int c) { return (0xFFFF0000u >> c) & 0xFFFFu; }

template <typename V>
__device__ __forceinline__ V rgw_fma(V a, V b, V c);
template <>
__device__ __forceinline__ double rgw_fma<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float rgw_fma<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Windows of a row with C tiles [cs, ce) and how many go in one work item.
__device__ __forceinline__ void rgw_split(const int32_t *__restrict__ colidxC, int64_t cs, int64_t ce, int item_tiles,
                                          int64_t &nwin, int64_t &per_item) {
    nwin = ce > cs ? (int64_t)(colidxC[ce - 1] - colidxC[cs]) / RGW_COLS + 1 : 0;
    const int64_t items = max<int64_t>(max<int64_t>(1, (ce - cs + item_tiles - 1) / item_tiles),
                                       (nwin + RGW_GMAX - 1) / RGW_GMAX);
    per_item = nwin > 0 ? (nwin + min(items, nwin) - 1) / min(items, nwin) : 1;
}

// Work items of each listed row: woff[q + 1] = item count of row q, woff[0] = 0
// (scanned afterwards into item offsets).
__global__ void __launch_bounds__(256)
k_rgw_count(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nr, const int64_t *__restrict__ dlo,
            const int64_t *__restrict__ dhi, const int64_t *__restrict__ tile_ptrC,
            const int32_t *__restrict__ row_nc, const int32_t *__restrict__ colidxC, int item_tiles,
            int64_t *__restrict__ woff) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t cover = nr;
    bin_rows(dlo, dhi, row_begin, nr);
    if (q == 0) woff[0] = 0;
    if (q >= nr) {
        if (q < cover) woff[q + 1] = 0;  // past a device-sized bin's rows: no items
        return;
    }
    int64_t cs, ce, nwin, per;
    c_row_range(tile_ptrC, row_nc, rows[row_begin + q], cs, ce);
    rgw_split(colidxC, cs, ce, item_tiles, nwin, per);
    woff[q + 1] = (nwin + per - 1) / per;
}

// Clears the two work counters (a kernel, so the numeric launch can be captured in a graph).
__global__ void k_rgw_zero(unsigned long long *ctr) {
    if (threadIdx.x < 2) ctr[threadIdx.x] = 0ull;
}

struct RgwItem {
    int32_t bi, c0, nw;  // tile-row, first tile column of the item's first window, windows
    int64_t cs, ce;      // the tile-row's C tiles
};

// Set-up of one work item, by every thread of the block: the row and windows
// of `item`, the tile-row's scalar row pointers s_rp[0..16] and the first C
// tile of each window s_pb[0..nw] (window w covers tiles [s_pb[w], s_pb[w+1])).
__device__ __forceinline__ void rgw_setup(int64_t item, const int32_t *__restrict__ rows, int64_t row_begin,
                                          int64_t nr, const int64_t *__restrict__ woff, int32_t m,
                                          const int64_t *__restrict__ rpA, const int64_t *__restrict__ tile_ptrC,
                                          const int32_t *__restrict__ row_nc, const int32_t *__restrict__ colidxC,
                                          int item_tiles, RgwItem &s_it, int64_t *s_rp, int64_t *s_pb) {
    if (threadIdx.x == 0) {
        int64_t lo = 0, hi = nr - 1;  // largest q with woff[q] <= item
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (woff[mid] <= item) lo = mid; else hi = mid - 1;
        }
        const int32_t bi = rows[row_begin + lo];
        int64_t cs, ce, nwin, per;
        c_row_range(tile_ptrC, row_nc, bi, cs, ce);
        rgw_split(colidxC, cs, ce, item_tiles, nwin, per);
        const int64_t w0 = (item - woff[lo]) * per;
        s_it.bi = bi;
        s_it.c0 = colidxC[cs] + (int32_t)w0 * RGW_COLS;
        s_it.nw = (int32_t)min<int64_t>(per, nwin - w0);
        s_it.cs = cs;
        s_it.ce = ce;
    }
    __syncthreads();
    if (threadIdx.x <= 16) {
        const int64_t ri = (int64_t)s_it.bi * 16 + (int64_t)threadIdx.x;
        s_rp[threadIdx.x] = rpA[ri < m ? ri : (int64_t)m];
    }
    if ((int)threadIdx.x <= s_it.nw)
        s_pb[threadIdx.x] = lower_bound<int64_t, int32_t>(colidxC, s_it.cs, s_it.ce,
                                                          s_it.c0 + (int32_t)threadIdx.x * RGW_COLS);
    __syncthreads();
}

// First entry of B row segment [b0, b1) at or after scalar column col (b1 when none).
__device__ __forceinline__ int64_t rgw_seek(const int32_t *__restrict__ colB, int64_t b0, int64_t b1, int32_t col) {
    if (b0 >= b1 || colB[b0] >= col) return b0;
    if (colB[b1 - 1] < col) return b1;
    return lower_bound<int64_t, int32_t>(colB, b0, b1, col);
}

struct RgwSeg {
    int64_t lo, hi;
    int32_t r;
};

__global__ void __launch_bounds__(RGW_THREADS)
k_rgw_symbolic(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nr, const int64_t *__restrict__ dlo,
               const int64_t *__restrict__ dhi, const int64_t *__restrict__ woff,
               unsigned long long *__restrict__ ctr, int32_t m, const int64_t *__restrict__ rpA,
               const int32_t *__restrict__ colA, const int64_t *__restrict__ rpB, const int32_t *__restrict__ colB,
               const int64_t *__restrict__ tile_ptrC, const int32_t *__restrict__ row_nc,
               const int32_t *__restrict__ colidxC, int item_tiles, uint16_t *__restrict__ maskC,
               uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC, int64_t *__restrict__ gcur,
               int64_t gcap) {
    // this workgroup's cursors: positions [0, gcap), then (entries left, column) pairs
    int64_t *const gpos = gcur + (int64_t)blockIdx.x * gcap * 2;
    int2 *const gcn = reinterpret_cast<int2 *>(gpos + gcap);
    __shared__ uint32_t s_m[RGW_COLS * 8];  // two 16-bit row masks per word
    __shared__ RgwSeg s_q[RGW_THREADS];     // long segments of one pass
    __shared__ int64_t s_rp[17];
    __shared__ int64_t s_pb[RGW_GMAX + 1];
    __shared__ RgwItem s_it;
    __shared__ int64_t s_item;
    __shared__ int s_nq;
    bin_rows(dlo, dhi, row_begin, nr);
    const int64_t total = woff[nr];
    if ((int64_t)blockIdx.x >= total) return;  // more workgroups than items: no counter traffic
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            s_item = (int64_t)atomicAdd(&ctr[0], 1ull);
            s_nq = 0;
        }
        __syncthreads();
        const int64_t item = s_item;
        if (item >= total) break;  // uniform: every thread read the same item
        rgw_setup(item, rows, row_begin, nr, woff, m, rpA, tile_ptrC, row_nc, colidxC, item_tiles, s_it, s_rp, s_pb);
        const int64_t e0 = s_rp[0], e1 = s_rp[16];
        bool fresh = true;  // no window of this item processed yet: cursors not yet set
        for (int w = 0; w < s_it.nw; w++) {
            const int32_t c0 = s_it.c0 + w * RGW_COLS;
            const int32_t lo_col = c0 * 16;
            const int64_t hi_col = (int64_t)(c0 + RGW_COLS) * 16;
            const int64_t p0 = s_pb[w], np = s_pb[w + 1] - p0;
            // no C tile in this window (uniform): no B entry of the tile-row lies
            // in it either, so the cursors stay valid
            if (np == 0) continue;
            for (int i = tid; i < RGW_COLS * 8; i += RGW_THREADS) s_m[i] = 0u;
            __syncthreads();
            for (int64_t base = e0; base < e1; base += RGW_THREADS * RGW_ILP) {  // uniform trip count
                // thread t takes entries base + t + 256u (u < RGW_ILP).  Each entry's
                // cursor -- its next B position, the entries left in its B row and
                // the column there -- carries from one window to the next within the
                // item (the windows follow in column order), so only the item's
                // first window looks B up; an entry whose next column lies past the
                // window touches no B memory.
                int64_t pos[RGW_ILP];
                int32_t rem[RGW_ILP], nx[RGW_ILP], k[RGW_ILP];
                bool dirty[RGW_ILP];
#pragma unroll
                for (int u = 0; u < RGW_ILP; u++) {
                    const int64_t e = base + tid + RGW_THREADS * u, ci = e - e0;
                    pos[u] = 0;
                    rem[u] = 0;
                    nx[u] = INT_MAX;
                    dirty[u] = false;
                    k[u] = -1;
                    if (e < e1) {
                        if (!fresh && ci < gcap) {
                            pos[u] = gpos[ci];
                            const int2 c = gcn[ci];
                            rem[u] = c.x;
                            nx[u] = c.y;
                        } else {
                            k[u] = colA[e];
                        }
                    }
                }
                int64_t b0[RGW_ILP], b1[RGW_ILP];
#pragma unroll
                for (int u = 0; u < RGW_ILP; u++) {
                    b0[u] = k[u] >= 0 ? rpB[k[u]] : 0;
                    b1[u] = k[u] >= 0 ? rpB[k[u] + 1] : 0;
                }
#pragma unroll
                for (int u = 0; u < RGW_ILP; u++) {
                    if (k[u] < 0) continue;
                    const int64_t e = base + tid + RGW_THREADS * u;
                    pos[u] = rgw_seek(colB, b0[u], b1[u], lo_col);
                    rem[u] = (int32_t)(b1[u] - pos[u]);
                    nx[u] = rem[u] > 0 ? colB[pos[u]] : INT_MAX;
                    dirty[u] = e - e0 < gcap;
                }
#pragma unroll
                for (int u = 0; u < RGW_ILP; u++) {
                    if (rem[u] <= 0 || nx[u] >= hi_col) continue;  // no entry in the window
                    const int64_t e = base + tid + RGW_THREADS * u;
                    int r = 0;
#pragma unroll
                    for (int st = 8; st > 0; st >>= 1)
                        if (r + st <= 15 && s_rp[r + st] <= e) r += st;
                    const int64_t lo = pos[u], hi = lo + rem[u];
                    int64_t next = hi;      // the first position at or past hi_col
                    int32_t nxt = INT_MAX;  // and its column
                    int qi = RGW_THREADS;   // long: more than RGW_LONG entries inside the window
                    if (hi - lo > RGW_LONG && colB[lo + RGW_LONG] < hi_col) qi = atomicAdd(&s_nq, 1);
                    if (qi < RGW_THREADS) {
                        next = lower_bound<int64_t, int32_t>(colB, lo + RGW_LONG, hi, (int32_t)min<int64_t>(hi_col, INT_MAX));
                        s_q[qi] = RgwSeg{lo, next, r};
                        nxt = next < hi ? colB[next] : INT_MAX;
                    } else {
                        for (int64_t f0 = lo; f0 < hi; f0 += RGW_PF) {  // up to RGW_PF loads issued together
                            int32_t jj[RGW_PF];
#pragma unroll
                            for (int t = 0; t < RGW_PF; t++) jj[t] = INT_MAX;
#pragma unroll
                            for (int t = 0; t < RGW_PF0; t++)
                                if (f0 + t < hi) jj[t] = colB[f0 + t];
                            if (f0 + RGW_PF0 < hi && jj[RGW_PF0 - 1] < hi_col) {
#pragma unroll
                                for (int t = RGW_PF0; t < RGW_PF; t++)
                                    if (f0 + t < hi) jj[t] = colB[f0 + t];
                            }
                            int in = 0;
#pragma unroll
                            for (int t = 0; t < RGW_PF; t++) {
                                const int tc = (jj[t] >> 4) - c0;
                                const bool inw = f0 + t < hi && jj[t] < hi_col;
                                if (inw && (unsigned)tc < (unsigned)RGW_COLS)
                                    atomicOr(&s_m[tc * 8 + (r >> 1)], (0x8000u >> (jj[t] & 15)) << ((r & 1) * 16));
                                in += inw ? 1 : 0;
                            }
                            if (in < RGW_PF && f0 + in < hi) {  // left the window at f0 + in (sorted columns)
                                next = f0 + in;
#pragma unroll
                                for (int t = 0; t < RGW_PF; t++)
                                    if (t == in) nxt = jj[t];
                                break;
                            }
                        }
                    }
                    pos[u] = next;
                    rem[u] = (int32_t)(hi - next);
                    nx[u] = nxt;
                    dirty[u] = e - e0 < gcap;
                }
#pragma unroll
                for (int u = 0; u < RGW_ILP; u++) {
                    const int64_t ci = base + tid + RGW_THREADS * u - e0;
                    if (!dirty[u]) continue;
                    gpos[ci] = pos[u];
                    gcn[ci] = make_int2(rem[u], nx[u]);
                }
                __syncthreads();
                const int nq = min(s_nq, RGW_THREADS);
                for (int qi = 0; qi < nq; qi++) {  // long segments: the whole block walks each
                    const RgwSeg sg = s_q[qi];
                    for (int64_t f = sg.lo + tid; f < sg.hi; f += RGW_THREADS) {
                        const int32_t j = colB[f];
                        if (j >= hi_col) break;
                        const int tc = (j >> 4) - c0;
                        if ((unsigned)tc < (unsigned)RGW_COLS)
                            atomicOr(&s_m[tc * 8 + (sg.r >> 1)], (0x8000u >> (j & 15)) << ((sg.r & 1) * 16));
                    }
                }
                __syncthreads();
                if (tid == 0) s_nq = 0;
                __syncthreads();
            }
            // the window's C tiles: masks, row offsets, value counts (16 threads per tile)
            for (int64_t it = tid; it < np * 16; it += RGW_THREADS) {
                const int64_t t = p0 + (it >> 4);
                const int rr = (int)(it & 15);
                const int tc = colidxC[t] - c0;
                const uint32_t rm = (s_m[tc * 8 + (rr >> 1)] >> ((rr & 1) * 16)) & 0xFFFFu;
                maskC[t * 16 + rr] = (uint16_t)rm;
                const int pc = __popc(rm);
                const int inc = group16_incl_sum(pc);
                ptrRowC[t * 16 + rr] = (uint8_t)(inc - pc);
                if (rr == 15) tile_nnzC[t + 1] = inc;
            }
            __syncthreads();
            fresh = false;
        }
    }
}

// A sub-window's LDS: nt tiles' row masks msk[nt * 16], row offsets
// prw[nt * 16] and value offsets tbs[nt], then the values, in buf; and the map
// tile column - c0 -> sub-window tile (0xFFFF: none).
struct RgwNumLDS {
    alignas(16) unsigned char buf[RGN_BYTES];
    int64_t cur[RGN_CUR];  // per A entry (first RGN_CUR of the tile-row): next B position
    int2 cn[RGN_CUR];      // and (entries left in its B row, the column there)
    uint16_t map[RGN_COLS];
};
__host__ __device__ constexpr int rgn_val_offset(int nt) { return (nt * RGN_TILE_BYTES + 15) & ~15; }

// Values of one numeric work item: the tiles of one tile-row whose value
// offsets start in one block of this many values.  Each item searches B once
// per A entry of its tile-row, so small items cost searches, and large ones
// leave the grid's tail unbalanced: about 128 items a workgroup over C's
// values, between two and eight sub-windows' worth (k_rgw_count_num decides
// from nnz(C) on the device and leaves the figure in ctr[2] for the kernel).
template <typename V>
constexpr int64_t rgw_item_min() { return 2 * RGN_BYTES / (int)sizeof(V); }
template <typename V>
constexpr int64_t rgw_item_max() { return 8 * RGN_BYTES / (int)sizeof(V); }

// Numeric work items of each listed row: woffn[q + 1] = ceil(row values / item values).
__global__ void __launch_bounds__(256)
k_rgw_count_num(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nr, const int64_t *__restrict__ dlo,
                const int64_t *__restrict__ dhi, const int64_t *__restrict__ tile_ptrC,
                const int32_t *__restrict__ row_nc, const int64_t *__restrict__ tile_nnzC,
                const int64_t *__restrict__ nnzC, int64_t hmin, int64_t hmax, int64_t per,
                int64_t *__restrict__ woffn, unsigned long long *__restrict__ ctr) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t cover = nr;
    const int64_t item_values = min(hmax, max(hmin, nnzC ? *nnzC / per : hmin));
    bin_rows(dlo, dhi, row_begin, nr);
    if (q == 0) {
        woffn[0] = 0;
        ctr[2] = (unsigned long long)item_values;
    }
    if (q >= nr) {
        if (q < cover) woffn[q + 1] = 0;  // past a device-sized bin's rows: no items
        return;
    }
    int64_t cs, ce;
    c_row_range(tile_ptrC, row_nc, rows[row_begin + q], cs, ce);
    const int64_t nv = ce > cs ? tile_nnzC[ce] - tile_nnzC[cs] : 0;
    woffn[q + 1] = ce > cs ? max<int64_t>(1, (nv + item_values - 1) / item_values) : 0;
}

struct RgwNumItem {
    int32_t bi;
    int64_t ce, ta, tb;  // row's end tile; the item's tiles [ta, tb)
};

// Head of a batch of 64 A entries (lane l: entry eh + l, if below eb): its B
// segment in the sub-window -- the carried cursor (position, entries left in
// the B row, the column there), or a search in the item's first sub-window --
// and its first RGW_PF0 products' columns and values.  skip: the carried
// column lies past the sub-window, so the entry has no product in it and
// nothing of B is loaded.
template <typename V>
__device__ __forceinline__ void rgn_head(int64_t eh, int64_t eb, int64_t rp0, bool first, const int64_t *cur,
                                         const int2 *ccn, const int64_t *gb, const int2 *gcn, int64_t gcap,
                                         const V *__restrict__ valA, const int32_t *__restrict__ colA,
                                         const int64_t *__restrict__ rpB, const int32_t *__restrict__ colB,
                                         const V *__restrict__ valB, int32_t lo_col, int64_t hi_col, int64_t &lo,
                                         int64_t &hi, V &a, bool &skip, int32_t *jj, V *bv) {
    const int64_t e = eh + lane_id();
    const int64_t ci = e - rp0;
    const bool in_lds = ci < RGN_CUR;
    const bool cached = in_lds || ci - RGN_CUR < gcap;
    lo = 0;
    hi = 0;
    a = V(0);
    skip = false;
    if (e < eb) {
        if (cached && !first) {
            lo = in_lds ? cur[ci] : gb[ci - RGN_CUR];
            const int2 c = in_lds ? ccn[ci] : gcn[ci - RGN_CUR];
            hi = lo + c.x;
            skip = c.y >= hi_col;
        } else {
            const int32_t k = colA[e];
            hi = rpB[k + 1];
            lo = rgw_seek(colB, rpB[k], hi, lo_col);
        }
        if (!skip) a = valA[e];
    }
#pragma unroll
    for (int t = 0; t < RGW_PF0; t++) {
        jj[t] = INT_MAX;
        bv[t] = V(0);
        if (!skip && lo + t < hi) {
            jj[t] = colB[lo + t];
            bv[t] = valB[lo + t];
        }
    }
}

// Output slot of product column j for scalar row r in the staged sub-window, or -1.
__device__ __forceinline__ int rgw_slot(const uint16_t *map, const uint16_t *msk, const uint8_t *prw,
                                        const int32_t *tbs, int32_t j, int32_t c0, int nt, int nv, int r) {
    const int tc = (j >> 4) - c0;
    if ((unsigned)tc >= (unsigned)RGN_COLS) return -1;
    const int lt = map[tc];
    if (lt >= nt) return -1;
    const uint32_t mrow = msk[lt * 16 + r];
    const int c = j & 15;
    if (!(mrow & (0x8000u >> c))) return -1;
    const int off = tbs[lt] + (int)prw[lt * 16 + r] + __popc(mrow & rgw_below(c));
    return off < nv ? off : -1;
}

template <typename V, bool ANY>
__global__ void __launch_bounds__(RGW_THREADS)
k_rgw_numeric(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nr, const int64_t *__restrict__ dlo,
              const int64_t *__restrict__ dhi, const int64_t *__restrict__ woffn,
              unsigned long long *__restrict__ ctr, int32_t m, const int64_t *__restrict__ rpA,
              const int32_t *__restrict__ colA, const V *__restrict__ valA, const int64_t *__restrict__ rpB,
              const int32_t *__restrict__ colB, const V *__restrict__ valB, const int64_t *__restrict__ tile_ptrC,
              const int32_t *__restrict__ row_nc, const int32_t *__restrict__ colidxC,
              const uint16_t *__restrict__ maskC, const uint8_t *__restrict__ ptrRowC,
              const int64_t *__restrict__ tile_nnzC, V *__restrict__ valC, uint8_t *__restrict__ colC,
              int32_t *__restrict__ overflow, int64_t val_cap, int64_t val_done, int64_t *__restrict__ gcur,
              int64_t gcap) {
    const int64_t H = (int64_t)ctr[2];  // values per work item (k_rgw_count_num)
    // this workgroup's global cursor area: positions [0, gcap), then (entries left, column) pairs
    int64_t *const gb = gcur + (int64_t)blockIdx.x * gcap * 2;
    int2 *const gcn = reinterpret_cast<int2 *>(gb + gcap);
    __shared__ RgwNumLDS s;
    __shared__ int64_t s_rp[17];
    __shared__ RgwNumItem s_it;
    __shared__ int64_t s_item, s_sp1;
    __shared__ int s_row;
    bin_rows(dlo, dhi, row_begin, nr);
    const int64_t total = woffn[nr];
    if ((int64_t)blockIdx.x >= total) return;  // more workgroups than items: no counter traffic
    const int tid = threadIdx.x, l = lane_id();
    for (;;) {
        if (tid == 0) s_item = (int64_t)atomicAdd(&ctr[1], 1ull);
        __syncthreads();
        const int64_t item = s_item;
        if (item >= total) break;  // uniform
        if (tid == 0) {
            int64_t lo = 0, hi = nr - 1;  // largest q with woffn[q] <= item
            while (lo < hi) {
                const int64_t mid = (lo + hi + 1) >> 1;
                if (woffn[mid] <= item) lo = mid; else hi = mid - 1;
            }
            const int32_t bi = rows[row_begin + lo];
            int64_t cs, ce;
            c_row_range(tile_ptrC, row_nc, bi, cs, ce);
            const int64_t j = item - woffn[lo], vb = tile_nnzC[cs];
            s_it.bi = bi;
            s_it.ce = ce;
            s_it.ta = lower_bound<int64_t, int64_t>(tile_nnzC, cs, ce, vb + j * H);
            s_it.tb = (item + 1 == woffn[lo + 1]) ? ce
                                                   : lower_bound<int64_t, int64_t>(tile_nnzC, s_it.ta, ce, vb + (j + 1) * H);
        }
        __syncthreads();
        if (tid <= 16) {
            const int64_t ri = (int64_t)s_it.bi * 16 + tid;
            s_rp[tid] = rpA[ri < m ? ri : (int64_t)m];
        }
        __syncthreads();
        if (tile_nnzC[s_it.ce] > val_cap) {  // C's values outgrew their capacity (uniform)
            if (tid == 0) atomicOr(overflow, 1);
            __syncthreads();
            continue;
        }
        if (tile_nnzC[s_it.ce] <= val_done) {  // a re-run: this tile-row's values are already written
            __syncthreads();
            continue;
        }
        const int64_t tb = s_it.tb;
        for (int64_t sp0 = s_it.ta; sp0 < tb;) {  // uniform sub-window loop
            const bool first = sp0 == s_it.ta;
            if (tid == 0) {
                s_sp1 = sp0 + 1;
                s_row = 0;
            }
            for (int q = tid; q < RGN_COLS; q += RGW_THREADS) s.map[q] = 0xFFFF;
            __syncthreads();
            {  // the sub-window ends at the largest end whose tiles [sp0, end) and their values fit
               // RGN_BYTES and span fewer than RGN_COLS tile columns (both tests are monotone in
               // end): every thread tests a few candidate ends at once, one coalesced load each
                const int64_t v0 = tile_nnzC[sp0];
                const int32_t cl = colidxC[sp0] + RGN_COLS;
                const int64_t hmax = min<int64_t>(tb, sp0 + RGN_BYTES / RGN_TILE_BYTES);
                int64_t best = sp0 + 1;
                for (int64_t end = sp0 + 2 + tid; end <= hmax; end += RGW_THREADS) {
                    if (rgn_val_offset((int)(end - sp0)) + (tile_nnzC[end] - v0) * (int64_t)sizeof(V) > RGN_BYTES ||
                        colidxC[end - 1] >= cl)
                        break;
                    best = end;
                }
                if (best > sp0 + 1)
                    atomicMax(reinterpret_cast<unsigned long long *>(&s_sp1), (unsigned long long)best);
            }
            __syncthreads();
            const int64_t sp1 = s_sp1;
            const int nt = (int)(sp1 - sp0);
            uint16_t *const msk = reinterpret_cast<uint16_t *>(s.buf);  // row masks of tile q: [16 q, 16 q + 16)
            uint8_t *const prw = s.buf + nt * 32;                          // row offsets within the tile
            int32_t *const tbs = reinterpret_cast<int32_t *>(s.buf + nt * 48);  // tile value offsets
            V *const vals = reinterpret_cast<V *>(s.buf + rgn_val_offset(nt));
            const int64_t vbase = tile_nnzC[sp0];
            const int nv = (int)min<int64_t>(tile_nnzC[sp1] - vbase, (RGN_BYTES - rgn_val_offset(nt)) / (int)sizeof(V));
            const int32_t c0 = colidxC[sp0];
            {  // the tiles' masks and row offsets copied as they are, 16 bytes a lane, by
               // loads that write LDS directly: lane i of a wave writes the wave's base
               // + 16 i, the base being piece q - l, the same for the whole wave
                const u32x4 *const gm = reinterpret_cast<const u32x4 *>(maskC) + sp0 * 2;
                const u32x4 *const gp = reinterpret_cast<const u32x4 *>(ptrRowC) + sp0;
                for (int q = tid; q < nt * 2; q += RGW_THREADS)
                    __builtin_amdgcn_global_load_lds(gm + q, reinterpret_cast<u32x4 *>(msk) + (q - l), 16, 0, 0);
                for (int q = tid; q < nt; q += RGW_THREADS)
                    __builtin_amdgcn_global_load_lds(gp + q, reinterpret_cast<u32x4 *>(prw) + (q - l), 16, 0, 0);
                int32_t cc[RGN_FILL];
                int64_t tn[RGN_FILL];
#pragma unroll
                for (int f = 0; f < RGN_FILL; f++) {
                    const int q = tid + f * RGW_THREADS;
                    if (q < nt) {
                        cc[f] = colidxC[sp0 + q];
                        tn[f] = tile_nnzC[sp0 + q];
                    }
                }
#pragma unroll
                for (int f = 0; f < RGN_FILL; f++) {
                    const int q = tid + f * RGW_THREADS;
                    if (q < nt) {
                        tbs[q] = (int32_t)(tn[f] - vbase);
                        s.map[cc[f] - c0] = (uint16_t)q;
                    }
                }
            }
            for (int q = tid; q < nv; q += RGW_THREADS) vals[q] = V(0);
            __syncthreads();
            const int32_t lo_col = c0 * 16;
            const int64_t hi_col = (int64_t)colidxC[sp1 - 1] * 16 + 16;
            // one wave per scalar row, rows taken in turn from s_row
            for (;;) {
                int r = 0;
                if (l == 0) r = atomicAdd(&s_row, 1);
                r = __shfl(r, 0, WAVE);
                if (r >= 16) break;  // wave-uniform
                const int32_t i = s_it.bi * 16 + r;
                if (i >= m) continue;
                const int64_t ea = s_rp[r], eb = s_rp[r + 1];
                // the next batch's head is loaded while this batch's products are added
                int64_t plo, phi;
                V pa;
                bool pskip;
                int32_t pjj[RGW_PF0];
                V pbv[RGW_PF0];
                rgn_head<V>(ea, eb, s_rp[0], first, s.cur, s.cn, gb, gcn, gcap, valA, colA, rpB, colB, valB, lo_col,
                            hi_col, plo, phi, pa, pskip, pjj, pbv);
                for (int64_t e0 = ea; e0 < eb; e0 += WAVE) {  // 64 A entries of row r at a time
                    // lane l owns entry e0 + l: its B segment in the sub-window,
                    // the first RGN_PF products' B values and output slots (each
                    // stage's loads issued together, all lanes at once)
                    // Within an item the sub-windows follow in column order and no B
                    // entry falls between two of them (it would have a C tile), so
                    // an entry's segment starts where its previous one ended: the
                    // first RGN_CUR entries of the tile-row keep that position in
                    // LDS and search only in the item's first sub-window.
                    const int64_t e = e0 + l;
                    const int64_t ci = e - s_rp[0];
                    const bool in_lds = ci < RGN_CUR;
                    const bool cached = in_lds || ci - RGN_CUR < gcap;
                    const int64_t lo = plo, hi = phi;
                    const V a = pa;
                    const bool skip = pskip;  // the cursor stays as it is
                    int64_t next = lo;        // where the segment ends (the next sub-window's start)
                    int32_t nxt = INT_MAX;    // and the column there
                    int sl[RGN_PF];
                    V bv[RGN_PF];
                    bool more = false, any = false;
#pragma unroll
                    for (int t = 0; t < RGN_PF; t++) sl[t] = -1;
                    {
                        int32_t jj[RGN_PF], tc[RGN_PF], lt[RGN_PF], tbo[RGN_PF];
                        uint32_t mr[RGN_PF], po[RGN_PF];
                        // two stages: the first RGW_PF0 entries (loaded with the batch
                        // head), the rest only while the segment stays in the sub-window
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) {
                            jj[t] = t < RGW_PF0 ? pjj[t < RGW_PF0 ? t : 0] : INT_MAX;
                            bv[t] = t < RGW_PF0 ? pbv[t < RGW_PF0 ? t : 0] : V(0);
                        }
                        if (lo + RGW_PF0 < hi && jj[RGW_PF0 - 1] < hi_col) {
#pragma unroll
                            for (int t = RGW_PF0; t < RGN_PF; t++)
                                if (lo + t < hi) {
                                    jj[t] = colB[lo + t];
                                    bv[t] = valB[lo + t];
                                }
                        }
                        uint32_t ok = 0;
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) {
                            tc[t] = (jj[t] >> 4) - c0;
                            const bool v = lo + t < hi && jj[t] < hi_col && (unsigned)tc[t] < (unsigned)RGN_COLS;
                            ok |= (uint32_t)v << t;
                            tc[t] = v ? tc[t] : 0;
                        }
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) lt[t] = s.map[tc[t]];
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) {
                            if (lt[t] >= nt) ok &= ~(1u << t);
                            lt[t] = lt[t] < nt ? lt[t] : 0;
                        }
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) {
                            mr[t] = msk[lt[t] * 16 + r];
                            po[t] = prw[lt[t] * 16 + r];
                            tbo[t] = tbs[lt[t]];
                        }
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) {
                            const int c = jj[t] & 15;
                            const uint32_t mrow = mr[t];
                            const int off = tbo[t] + (int)po[t] + __popc(mrow & rgw_below(c));
                            sl[t] = ((ok >> t) & 1u) && (mrow & (0x8000u >> c)) && off < nv ? off : -1;
                        }
                        any = jj[0] < hi_col;
                        more = lo + RGN_PF < hi && jj[RGN_PF - 1] < hi_col;
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++) next += (lo + t < hi && jj[t] < hi_col) ? 1 : 0;
                        // the first loaded column past the sub-window (entries are in column order)
#pragma unroll
                        for (int t = RGN_PF - 1; t >= 0; t--)
                            if (lo + t < hi && jj[t] >= hi_col) nxt = jj[t];
                    }
                    if (e0 + WAVE < eb)  // wave-uniform
                        rgn_head<V>(e0 + WAVE, eb, s_rp[0], first, s.cur, s.cn, gb, gcn, gcap, valA, colA, rpB, colB,
                                    valB, lo_col, hi_col, plo, phi, pa, pskip, pjj, pbv);
                    if (ANY) {  // every lane adds its staged products at once, in any order
#pragma unroll
                        for (int t = 0; t < RGN_PF; t++)
                            if (sl[t] >= 0) atomicAdd(&vals[sl[t]], a * bv[t]);
                    }
                    uint64_t mk = ANY ? __ballot(more) : __ballot(any);
                    const uint64_t mm = __ballot(more);
                    while (mk) {  // entries in ascending k; one entry's outputs are distinct
                        const int sgi = __ffsll((unsigned long long)mk) - 1;
                        mk &= mk - 1;
                        if (!ANY && l == sgi) {  // the owner adds its staged products (all reads, then all writes)
                            V old[RGN_PF];
#pragma unroll
                            for (int t = 0; t < RGN_PF; t++) old[t] = vals[sl[t] >= 0 ? sl[t] : 0];
#pragma unroll
                            for (int t = 0; t < RGN_PF; t++)
                                if (sl[t] >= 0) vals[sl[t]] = rgw_fma<V>(a, bv[t], old[t]);
                        }
                        if ((mm >> sgi) & 1) {  // the rest of a long segment: the whole wave, 64 at a time
                            const V sa = __shfl(a, sgi, WAVE);
                            const int64_t slo = __shfl(lo, sgi, WAVE) + RGN_PF, shi = __shfl(hi, sgi, WAVE);
                            int64_t f0 = slo;
                            int32_t pj = INT_MAX;  // the column where the segment left the sub-window
                            for (; f0 < shi; f0 += WAVE) {  // wave-uniform trip
                                const int64_t f = f0 + l;
                                const int32_t j = f < shi ? colB[f] : INT_MAX;
                                const V b = f < shi ? valB[f] : V(0);
                                const int off = j < hi_col ? rgw_slot(s.map, msk, prw, tbs, j, c0, nt, nv, r) : -1;
                                if (off >= 0) vals[off] = rgw_fma<V>(sa, b, vals[off]);
                                const uint64_t past = __ballot(j >= hi_col);
                                if (past) {  // the segment left the sub-window
                                    const int q = __ffsll((unsigned long long)past) - 1;
                                    pj = __shfl(j, q, WAVE);
                                    f0 += q;
                                    break;
                                }
                            }
                            if (l == sgi) {
                                next = min(f0, shi);
                                nxt = next < shi ? pj : INT_MAX;
                            }
                        }
                    }
                    if (e < eb && cached && !skip) {
                        const int2 c = make_int2((int32_t)(hi - next), nxt);
                        if (in_lds) {
                            s.cur[ci] = next;
                            s.cn[ci] = c;
                        } else {
                            gb[ci - RGN_CUR] = next;
                            gcn[ci - RGN_CUR] = c;
                        }
                    }
                }
            }
            __syncthreads();
            // values of the sub-window (contiguous), then its col bytes (16 threads per tile)
            for (int q = tid; q < nv; q += RGW_THREADS) valC[vbase + q] = vals[q];
            for (int q = tid; q < nt * 16; q += RGW_THREADS) {
                const int lt = q >> 4, rr = q & 15;
                uint32_t mrow = msk[q];
                int o = tbs[lt] + (int)prw[q];
                while (mrow) {
                    const int c = __clz(mrow) - 16;
                    mrow ^= 0x8000u >> c;
                    if (o < nv) colC[vbase + o] = (uint8_t)((rr << 4) | c);
                    o++;
                }
            }
            __syncthreads();
            sp0 = sp1;
        }
    }
}

static unsigned rgw_grid(Context *ctx, int per_cu) {
    return (unsigned)std::max(1, ctx->prop.multiProcessorCount * per_cu);
}

static int64_t rgn_gcap(const tsg_tile_matrix *A) { return std::min<int64_t>(RGN_GCUR, A->nnz); }
static int64_t rgs_gcap(const tsg_tile_matrix *A) { return std::min<int64_t>(RGS_GCUR, A->nnz); }

// The cursor area both row-by-row kernels use (allocated once per product,
// at the size of the larger of their grids' areas).
static int alloc_rg_cursors(Context *ctx, const tsg_tile_matrix *A, int64_t **gcur) {
    if (*gcur) return TSG_OK;
    const int64_t need = std::max((int64_t)rgw_grid(ctx, 3) * rgn_gcap(A), (int64_t)rgw_grid(ctx, 4) * rgs_gcap(A));
    return need > 0 ? palloc(ctx, gcur, (size_t)need * 2) : TSG_OK;
}

int launch_rg_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                       const int32_t *rows, const BinRange &br, int64_t *woff, int64_t **gcur, const SlotLayout &sl) {
    const int64_t begin = br.begin, count = br.count;
    if (count <= 0) return TSG_OK;
    TSG_CHECK(alloc_rg_cursors(ctx, A, gcur));
    unsigned long long *ctr = reinterpret_cast<unsigned long long *>(woff + count + 1);
    // about 128 work items a workgroup over C's tiles, RGW_ITEM_TILES at most
    // (each item searches B once per A entry of its tile-row; smaller C need
    // smaller items for the grid's tail to stay even)
    const unsigned grid = rgw_grid(ctx, 4);
    const int item_tiles = (int)std::max<int64_t>(1024, std::min<int64_t>(RGW_ITEM_TILES, C->numtile / ((int64_t)grid * 128)));
    k_rgw_count<<<(unsigned)((count + 256) / 256), 256, 0, ctx->stream>>>(rows, begin, count, br.lo, br.hi,
                                                                         C->tile_ptr, sl.row_nc, C->tile_columnidx,
                                                                         item_tiles, woff);
    TSG_HIP(hipGetLastError());
    TSG_CHECK(scan_inclusive_i64(ctx, woff, count + 1));
    k_rgw_zero<<<1, 64, 0, ctx->stream>>>(ctr);
    k_rgw_symbolic<<<grid, RGW_THREADS, 0, ctx->stream>>>(
        rows, begin, count, br.lo, br.hi, woff, ctr, A->m, A->row_ptr, A->row_col, B->row_ptr, B->row_col, C->tile_ptr,
        sl.row_nc, C->tile_columnidx, item_tiles, C->mask, C->tile_csr_Ptr, C->tile_nnz, *gcur,
        *gcur ? rgs_gcap(A) : 0);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int prepare_rg_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *C, const int32_t *rows,
                       const BinRange &br, int64_t *woff, int64_t *woffn, int64_t **gcur, const int64_t *nnzC,
                       const SlotLayout &sl) {
    const int64_t begin = br.begin, count = br.count;
    if (count <= 0) return TSG_OK;
    TSG_CHECK(alloc_rg_cursors(ctx, A, gcur));
    const bool f32 = C->dtype == TSG_F32;
    const int64_t hmin = f32 ? rgw_item_min<float>() : rgw_item_min<double>();
    const int64_t hmax = f32 ? rgw_item_max<float>() : rgw_item_max<double>();
    unsigned long long *ctr = reinterpret_cast<unsigned long long *>(woff + count + 1);
    k_rgw_count_num<<<(unsigned)((count + 256) / 256), 256, 0, ctx->stream>>>(
        rows, begin, count, br.lo, br.hi, C->tile_ptr, sl.row_nc, C->tile_nnz, nnzC, hmin, hmax,
        (int64_t)rgw_grid(ctx, 3) * 128, woffn, ctr);
    TSG_HIP(hipGetLastError());
    return scan_inclusive_i64(ctx, woffn, count + 1);
}

int launch_rg_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                      const int32_t *rows, const BinRange &br, int64_t *woff, const int64_t *woffn,
                      int64_t *gcur, const SlotLayout &sl, bool any_order) {
    const int64_t begin = br.begin, count = br.count;
    if (count <= 0) return TSG_OK;
    const int64_t gcap = gcur ? rgn_gcap(A) : 0;
    unsigned long long *ctr = reinterpret_cast<unsigned long long *>(woff + count + 1);
    k_rgw_zero<<<1, 64, 0, ctx->stream>>>(ctr);
#define TSG_RG_NUMERIC(V, ANY)                                                                                       \
    k_rgw_numeric<V, ANY><<<rgw_grid(ctx, 3), RGW_THREADS, 0, ctx->stream>>>(                                                    \
        rows, begin, count, br.lo, br.hi, woffn, ctr, A->m, A->row_ptr, A->row_col, (const V *)A->row_val,           \
        B->row_ptr, B->row_col, (const V *)B->row_val, C->tile_ptr, sl.row_nc, C->tile_columnidx, C->mask,           \
        C->tile_csr_Ptr, C->tile_nnz, (V *)C->tile_csr_Val, C->tile_csr_Col, sl.overflow, sl.val_cap, sl.val_done,  \
        gcur, gcap)
    if (A->dtype == TSG_F32) {
        if (any_order) TSG_RG_NUMERIC(float, true);
        else TSG_RG_NUMERIC(float, false);
    } else {
        if (any_order) TSG_RG_NUMERIC(double, true);
        else TSG_RG_NUMERIC(double, false);
    }
#undef TSG_RG_NUMERIC
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

}  // namespace tsg
