// convert.hip -- tile-format conversion on the GPU.
//
// tsg_csr2tile replaces the host csr2tile (src/tile_format.h:93-222) and
// tsg_build_csc_tiles replaces the host build_csc_tiles
// (src/tile_format.h:424-546).  Output arrays are identical to what those host
// functions produce for CSR input whose rows are sorted by column (the form
// load_mtx, the --aat transpose and every generator give); duplicate columns
// are kept in input order like the reference.
//
// csr2tile, one wave64 per 16-row tile-row, entries read coalesced from the
// contiguous CSR range [rowptr[16bi], rowptr[16bi+16]):
//   K1 count   -- LDS bitmap over the tile-row's tile-column window; popcount
//   K2 fill    -- same bitmap: ascending tile_columnidx; the first entry of each
//                 (tile,row) segment walks its segment (at most 16 distinct
//                 columns) to produce that row's count and bitmask, so no
//                 atomics are needed
//   K3 rowscan -- per tile, 16-lane exclusive scan of row counts -> tile_csr_Ptr
//                 and the tile's nnz; one int64 scan -> tile_nnz
//   K4 scatter -- each entry's slot = tile_nnz + tile_csr_Ptr[row] + position in
//                 its segment; col byte (row<<4)|col
// build_csc_tiles: tile-level transpose.  Tiles are counted per tile column,
// dropped into unordered column slots, each column's tile rows sorted with the
// same LDS-bitmap pass, then every tile's 16-row payload (row pointers, masks,
// values, local col bytes) is gathered into its CSC position.
#include <climits>
#include <cstring>

#include "tsg_internal.h"

namespace tsg {

constexpr int CV_WAVES = 4;
constexpr int CV_WIN_WORDS = 1024;  // 32,768 tile columns per pass

// ---------------------------------------------------------------- csr2tile
// Tile column of column c for tile size TS (16, a shift) or ts (TS = 0, the
// parameterized tile sizes 1-15); negative (invalid) columns map below 0.
template <int TS>
__device__ __forceinline__ int32_t tile_of(int32_t c, int ts) {
    if (TS == 16) return c >> 4;
    return c >= 0 ? c / ts : -1;
}

template <int MODE, int TS>  // MODE 0 = count, 1 = fill; TS = 16, or 0 for a runtime tile size ts
__global__ void __launch_bounds__(CV_WAVES *WAVE)
k_c2t(const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int32_t m, int32_t n, int32_t tilem,
      int64_t *__restrict__ tile_ptr, int32_t *__restrict__ tile_colidx, uint8_t *__restrict__ rowcnt,
      uint16_t *__restrict__ mask, int32_t *__restrict__ flags, int ts_rt) {
    const int ts = TS ? TS : ts_rt;
    __shared__ uint32_t s_bm[CV_WAVES][CV_WIN_WORDS];
    __shared__ int32_t s_wpre[CV_WAVES][CV_WIN_WORDS];
    __shared__ int64_t s_rp[CV_WAVES][17];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int32_t bi = blockIdx.x * CV_WAVES + w;
    if (bi >= tilem) return;
    const int32_t rs = bi * ts, re = min(rs + ts, m);
    if (lane <= ts) s_rp[w][lane] = rp[min(rs + lane, re)];  // rows past m repeat the end
    wave_sync();
    const int64_t js = s_rp[w][0], je = s_rp[w][re - rs];
    int32_t cmin = INT_MAX, cmax = -1;
    if (lane < re - rs) {
        int64_t a = s_rp[w][lane], b = s_rp[w][lane + 1];
        if (b > a) { cmin = tile_of<TS>(col[a], ts); cmax = tile_of<TS>(col[b - 1], ts); }
    }
    cmin = wave_min(cmin);
    cmax = wave_max(cmax);
    uint32_t *bm = s_bm[w];
    int64_t count = 0;
    int64_t tbase = MODE ? tile_ptr[bi] : 0;
    bool bad = false, dup = false;
    for (int64_t w0 = cmin; w0 <= cmax; w0 += (int64_t)CV_WIN_WORDS * 32) {
        const int32_t wbits = (int32_t)min<int64_t>((int64_t)CV_WIN_WORDS * 32, cmax - w0 + 1);
        const int nwords = (wbits + 31) >> 5;
        for (int i = lane; i < nwords; i += WAVE) bm[i] = 0;
        wave_sync();
        for (int64_t j = js + lane; j < je; j += WAVE) {
            int64_t off = (int64_t)tile_of<TS>(col[j], ts) - w0;
            if (off >= 0 && off < wbits) atomicOr(&bm[off >> 5], 1u << (off & 31));
        }
        wave_sync();
        if (MODE == 0) {
            int32_t cnt = 0;
            for (int i = lane; i < nwords; i += WAVE) cnt += __popc(bm[i]);
            count += wave_sum(cnt);
        } else {
            // per-word exclusive tile ranks and ascending tile_columnidx
            const int wpl = (nwords + WAVE - 1) / WAVE;
            const int i0 = min(nwords, lane * wpl), i1 = min(nwords, i0 + wpl);
            int32_t cnt = 0;
            for (int i = i0; i < i1; i++) cnt += __popc(bm[i]);
            int32_t incl = wave_incl_sum(cnt);
            int32_t tot = __shfl(incl, WAVE - 1, WAVE);
            int32_t run = incl - cnt;
            for (int i = i0; i < i1; i++) {
                s_wpre[w][i] = run;
                uint32_t b = bm[i];
                while (b) {
                    int bit = __ffs(b) - 1;
                    b &= b - 1;
                    tile_colidx[tbase + run] = (int32_t)(w0 + i * 32 + bit);
                    run++;
                }
            }
            wave_sync();
            // segment heads: (tile,row) counts and masks
            const uint32_t top = 1u << (ts - 1);  // mask bit of local column 0
            for (int64_t j = js + lane; j < je; j += WAVE) {
                const int32_t cj = col[j];
                const int32_t tc = tile_of<TS>(cj, ts);
                int64_t off = (int64_t)tc - w0;
                if (off < 0 || off >= wbits) continue;
                int r = 0;
                while (r < ts - 1 && s_rp[w][r + 1] <= j) r++;
                const int64_t rowbeg = s_rp[w][r], rowend = s_rp[w][r + 1];
                if (j > rowbeg && tile_of<TS>(col[j - 1], ts) == tc) continue;  // not a segment head
                const int wi = (int)(off >> 5);
                const int32_t t = s_wpre[w][wi] + __popc(bm[wi] & ((1u << (off & 31)) - 1u));
                const int64_t T = tbase + t;
                uint32_t mk = 0;
                int32_t n = 0, prev = -1;
                for (int64_t q = j; q < rowend; q++) {
                    const int32_t cq = col[q];
                    if (tile_of<TS>(cq, ts) != tc) break;
                    if (cq < prev) bad = true;
                    prev = cq;
                    mk |= top >> (cq - tc * ts);
                    n++;
                }
                if (n != __popc(mk)) dup = true;  // a column repeated inside the row
                rowcnt[T * ts + r] = (uint8_t)n;
                mask[T * ts + r] = (uint16_t)mk;
            }
            tbase += tot;
            wave_sync();
        }
    }
    bool range_bad = false;
    if (MODE == 0) {
        // rows must be sorted by column: a descending step is rejected;
        // column indices must lie in [0, n)
        for (int64_t j = js + lane; j < je; j += WAVE) {
            if (col[j] < 0 || col[j] >= n) range_bad = true;
            if (j == js) continue;
            int r = 0;
            while (r < ts - 1 && s_rp[w][r + 1] <= j) r++;
            if (j > s_rp[w][r] && col[j - 1] > col[j]) bad = true;
        }
        if (lane == 0) {
            tile_ptr[bi + 1] = count;
            if (bi == 0) tile_ptr[0] = 0;
            if (count > HEAVY_LIST) atomicOr(flags, 8);  // a long tile-row
        }
    }
    if (__ballot(bad) && lane == 0) atomicOr(flags, 1);
    if (__ballot(dup) && lane == 0) atomicOr(flags, 2);
    if (__ballot(range_bad) && lane == 0) atomicOr(flags, 4);
}

// Per tile: exclusive scan of its ts (<= 16) row counts (in place) and the tile's nnz.
__global__ void __launch_bounds__(256) k_tile_rowscan(int64_t numtile, uint8_t *__restrict__ ptr,
                                                      int64_t *__restrict__ tile_nnz, int ts) {
    const int l = threadIdx.x & 15;
    for (int64_t T = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 4; T < numtile; T += (int64_t)gridDim.x * 16) {
        const int c = l < ts ? ptr[T * ts + l] : 0;
        const int incl = group16_incl_sum(c);
        if (l < ts) ptr[T * ts + l] = (uint8_t)(incl - c);
        if (l == 15) tile_nnz[T + 1] = incl;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_nnz[0] = 0;
}

template <typename V, int TS>
__global__ void __launch_bounds__(CV_WAVES *WAVE)
k_c2t_scatter(const int64_t *__restrict__ rp, const int32_t *__restrict__ col, const V *__restrict__ val,
              int32_t m, int32_t tilem, const int64_t *__restrict__ tile_ptr,
              const int32_t *__restrict__ tile_colidx, const int64_t *__restrict__ tile_nnz,
              const uint8_t *__restrict__ ptr, V *__restrict__ tval, uint8_t *__restrict__ tcol,
              int64_t *__restrict__ perm, int ts_rt) {
    const int ts = TS ? TS : ts_rt;
    __shared__ int64_t s_rp[CV_WAVES][17];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int32_t bi = blockIdx.x * CV_WAVES + w;
    if (bi >= tilem) return;
    const int32_t rs = bi * ts, re = min(rs + ts, m);
    if (lane <= ts) s_rp[w][lane] = rp[min(rs + lane, re)];  // rows past m repeat the end
    wave_sync();
    const int64_t js = s_rp[w][0], je = s_rp[w][re - rs];
    const int64_t tb = tile_ptr[bi], te = tile_ptr[bi + 1];
    for (int64_t j = js + lane; j < je; j += WAVE) {
        const int32_t cj = col[j];
        const int32_t tc = tile_of<TS>(cj, ts);
        int r = 0;
        while (r < ts - 1 && s_rp[w][r + 1] <= j) r++;
        const int64_t T = lower_bound<int64_t, int32_t>(tile_colidx, tb, te, tc);
        const int64_t seg = lower_bound<int64_t, int32_t>(col, s_rp[w][r], j, tc * ts);
        const int64_t pos = tile_nnz[T] + ptr[T * ts + r] + (j - seg);
        tval[pos] = val[j];
        tcol[pos] = (uint8_t)((r << 4) | (cj - tc * ts));  // (row<<4)|col for every tile size
        perm[j] = pos;
    }
}

// ---------------------------------------------------------------- helpers
__global__ void __launch_bounds__(256) k_fill_rowidx(const int64_t *__restrict__ tile_ptr, int32_t tilem,
                                                     int32_t *__restrict__ rowidx) {
    for (int64_t bi = ((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE; bi < tilem; bi += (int64_t)gridDim.x * 4)
        for (int64_t t = tile_ptr[bi] + lane_id(); t < tile_ptr[bi + 1]; t += WAVE) rowidx[t] = (int32_t)bi;
}

// ---------------------------------------------------------------- build_csc_tiles
__global__ void __launch_bounds__(256) k_csc_count(int64_t numtile, const int32_t *__restrict__ colidx,
                                                   int64_t *__restrict__ cptr) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < numtile) atomicAdd((unsigned long long *)&cptr[colidx[t] + 1], 1ull);
}

__global__ void __launch_bounds__(256) k_csc_slot(int64_t numtile, const int32_t *__restrict__ colidx,
                                                  const int32_t *__restrict__ rowidx,
                                                  int64_t *__restrict__ cursor, int32_t *__restrict__ tmp) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= numtile) return;
    const int64_t s = (int64_t)atomicAdd((unsigned long long *)&cursor[colidx[t]], 1ull);
    tmp[s] = rowidx[t];
}

// One wave per tile column: sort its tile rows with an LDS bitmap.
__global__ void __launch_bounds__(CV_WAVES *WAVE)
k_csc_sort(int32_t tilen, const int64_t *__restrict__ cptr, const int32_t *__restrict__ tmp,
           int32_t *__restrict__ crow) {
    __shared__ uint32_t s_bm[CV_WAVES][CV_WIN_WORDS];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int32_t bt = blockIdx.x * CV_WAVES + w;
    if (bt >= tilen) return;
    const int64_t s = cptr[bt], e = cptr[bt + 1];
    int32_t vmin = INT_MAX, vmax = -1;
    for (int64_t i = s + lane; i < e; i += WAVE) { vmin = min(vmin, tmp[i]); vmax = max(vmax, tmp[i]); }
    vmin = wave_min(vmin);
    vmax = wave_max(vmax);
    uint32_t *bm = s_bm[w];
    int64_t out = s;
    for (int64_t w0 = vmin; w0 <= vmax; w0 += (int64_t)CV_WIN_WORDS * 32) {
        const int32_t wbits = (int32_t)min<int64_t>((int64_t)CV_WIN_WORDS * 32, vmax - w0 + 1);
        const int nwords = (wbits + 31) >> 5;
        for (int i = lane; i < nwords; i += WAVE) bm[i] = 0;
        wave_sync();
        for (int64_t i = s + lane; i < e; i += WAVE) {
            int64_t off = (int64_t)tmp[i] - w0;
            if (off >= 0 && off < wbits) atomicOr(&bm[off >> 5], 1u << (off & 31));
        }
        wave_sync();
        const int wpl = (nwords + WAVE - 1) / WAVE;
        const int i0 = min(nwords, lane * wpl), i1 = min(nwords, i0 + wpl);
        int32_t cnt = 0;
        for (int i = i0; i < i1; i++) cnt += __popc(bm[i]);
        int32_t incl = wave_incl_sum(cnt);
        int32_t tot = __shfl(incl, WAVE - 1, WAVE);
        int64_t pos = out + incl - cnt;
        for (int i = i0; i < i1; i++) {
            uint32_t b = bm[i];
            while (b) {
                int bit = __ffs(b) - 1;
                b &= b - 1;
                crow[pos++] = (int32_t)(w0 + i * 32 + bit);
            }
        }
        out += tot;
        wave_sync();
    }
}

// pos_of[t] = CSC position p of CSR tile t, inv[p] = t (when inv is given),
// and the tile's entry count at nnz_new[p + 1].
__global__ void __launch_bounds__(256)
k_csc_perm(int64_t numtile, const int32_t *__restrict__ colidx, const int32_t *__restrict__ rowidx,
           const int64_t *__restrict__ cptr, const int32_t *__restrict__ crow,
           const int64_t *__restrict__ nnz_old, int64_t *__restrict__ inv, int64_t *__restrict__ nnz_new,
           int64_t *__restrict__ pos_of) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= numtile) return;
    const int32_t tc = colidx[t];
    const int64_t p = lower_bound<int64_t, int32_t>(crow, cptr[tc], cptr[tc + 1], rowidx[t]);
    if (inv) inv[p] = t;
    pos_of[t] = p;
    nnz_new[p + 1] = nnz_old[t + 1] - nnz_old[t];
    if (t == 0) nnz_new[0] = 0;
}

// 16 lanes per CSC tile: row pointers, masks (ts <= 16 rows), then values
// and local cols (the low 4 bits of (row<<4)|col for every tile size).
template <typename V>
__global__ void __launch_bounds__(256)
k_csc_gather(int64_t numtile, const int64_t *__restrict__ inv, const int64_t *__restrict__ nnz_old,
             const int64_t *__restrict__ nnz_new, const uint8_t *__restrict__ ptr_old,
             const uint16_t *__restrict__ mask_old, const uint8_t *__restrict__ col_old,
             const V *__restrict__ val_old, uint8_t *__restrict__ ptr_new, uint16_t *__restrict__ mask_new,
             uint8_t *__restrict__ col_new, V *__restrict__ val_new, int ts) {
    const int l = threadIdx.x & 15;
    for (int64_t p = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 4; p < numtile; p += (int64_t)gridDim.x * 16) {
        const int64_t t = inv[p];
        if (l < ts) {
            ptr_new[p * ts + l] = ptr_old[t * ts + l];
            mask_new[p * ts + l] = mask_old[t * ts + l];
        }
        const int64_t so = nnz_old[t], sn = nnz_new[p], n = nnz_old[t + 1] - so;
        for (int64_t e = l; e < n; e += 16) {
            val_new[sn + e] = val_old[so + e];
            col_new[sn + e] = col_old[so + e] & 15;
        }
    }
}

// row_nnz_ptr[k] = first payload entry of tile-row k (tile_nnz at its first tile).
__global__ void __launch_bounds__(256) k_row_nnz_ptr(int32_t tilem, const int64_t *__restrict__ tile_ptr,
                                                     const int64_t *__restrict__ tile_nnz, int64_t *__restrict__ rnp) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k <= tilem) rnp[k] = tile_nnz[tile_ptr[k]];
}

// The re-layout of 16 x 16 tiles without repeated columns in one pass over
// the CSR-order tiles, 16 lanes per tile t (lane l = row l and column l):
//  - CSC order: tile t's row offsets, row masks, values and local columns
//    (the low 4 bits of the col byte) move to CSC position p = pos_of[t];
//  - column-major copy at t (read by the staged numeric kernels): column
//    masks (bit 15-r = row r), exclusive column offsets, and each column's
//    values in ascending row order.
// Entry (r, c) of the tile goes to column c's first value plus the rows
// above r that hold column c.  Each entry is read once; the group reads the
// tile's 16 row masks and offsets as three 16-byte words.
template <typename V>
__global__ void __launch_bounds__(256)
k_csc_relayout(int64_t numtile, const int64_t *__restrict__ pos_of, const int64_t *__restrict__ nnz_old,
               const int64_t *__restrict__ nnz_new, const uint8_t *__restrict__ ptr_old,
               const uint16_t *__restrict__ mask_old, const uint8_t *__restrict__ col_old,
               const V *__restrict__ val_old, uint8_t *__restrict__ ptr_new, uint16_t *__restrict__ mask_new,
               uint8_t *__restrict__ col_new, V *__restrict__ val_new, uint16_t *__restrict__ cmask,
               uint8_t *__restrict__ cptr, V *__restrict__ cval) {
    const int l = threadIdx.x & 15;
    const uint32_t cbit = 0x8000u >> l;
    // grid-stride over tiles; whole 16-lane groups move together
    for (int64_t t = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 4; t < numtile; t += (int64_t)gridDim.x * 16) {
        const int64_t p = pos_of[t];
        const int64_t so = nnz_old[t], n = nnz_old[t + 1] - so;
        const uint4 mlo = reinterpret_cast<const uint4 *>(mask_old)[t * 2];
        const uint4 mhi = reinterpret_cast<const uint4 *>(mask_old)[t * 2 + 1];
        const uint4 po = reinterpret_cast<const uint4 *>(ptr_old)[t];
        const int64_t sn = nnz_new[p];
        const uint32_t mw[8] = {mlo.x, mlo.y, mlo.z, mlo.w, mhi.x, mhi.y, mhi.z, mhi.w};
        const uint32_t pw[4] = {po.x, po.y, po.z, po.w};
        uint32_t cm = 0, rm = 0, ro = 0;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t m = (mw[r >> 1] >> ((r & 1) * 16)) & 0xFFFFu;
            if (m & cbit) cm |= 0x8000u >> r;
            if (r == l) {
                rm = m;
                ro = (pw[r >> 2] >> ((r & 3) * 8)) & 0xFFu;
            }
        }
        const int pc = __popc(cm);
        const int co = group16_incl_sum(pc) - pc;
        ptr_new[p * 16 + l] = (uint8_t)ro;
        mask_new[p * 16 + l] = (uint16_t)rm;
        cmask[t * 16 + l] = (uint16_t)cm;
        cptr[t * 16 + l] = (uint8_t)co;
        // entries 16 at a time; n is uniform in the group, so every lane
        // takes part in the exchanges of column masks and offsets
        for (int64_t e0 = 0; e0 < n; e0 += 16) {
            const int64_t e = e0 + l;
            const bool in = e < n;
            V v = V(0);
            int cb = 0;
            if (in) {
                v = val_old[so + e];
                cb = col_old[so + e];
            }
            const int r = cb >> 4, c = cb & 15;
            const uint32_t cmc = __shfl(cm, c, 16);
            const int coc = __shfl(co, c, 16);
            if (in) {
                val_new[sn + e] = v;
                col_new[sn + e] = (uint8_t)c;
                cval[so + coc + __popc(cmc >> (16 - r))] = v;  // rows 0..r-1 of column c
            }
        }
    }
}

// Value positions after the re-layout: perm[e] (a position in the CSR-order
// tiles) becomes the position in the CSC-order tiles; perm_col[e] the position
// in the column-major copy.  The entry's row and column inside its tile come
// from its CSR-order col byte (r<<4)|c.
__global__ void __launch_bounds__(256)
k_val_perm_b(int64_t nnz, int64_t numtile, const int64_t *__restrict__ nnz_old, const int64_t *__restrict__ pos_of,
             const int64_t *__restrict__ nnz_new, const uint8_t *__restrict__ col_old,
             const uint16_t *__restrict__ cmask, const uint8_t *__restrict__ cptr, int64_t *__restrict__ perm,
             int64_t *__restrict__ perm_col) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * 256) {
    const int64_t old = perm[e];
    int64_t lo = 0, hi = numtile - 1;  // last tile whose first entry is at or before `old`
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (nnz_old[mid] <= old) lo = mid; else hi = mid - 1;
    }
    const int64_t t = lo, off = old - nnz_old[t];
    perm[e] = nnz_new[pos_of[t]] + off;
    if (perm_col) {
        const int rc = col_old[old], r = rc >> 4, c = rc & 15;
        perm_col[e] = nnz_old[t] + cptr[t * 16 + c] + __popc(cmask[t * 16 + c] & ((0xFFFF0000u >> r) & 0xFFFFu));
    }
    }
}

// New values in CSR order into every copy.
template <typename V>
__global__ void __launch_bounds__(256)
k_set_values(int64_t nnz, const V *__restrict__ vals, const int64_t *__restrict__ perm,
             const int64_t *__restrict__ perm_col, V *__restrict__ tval, V *__restrict__ cval, V *__restrict__ rval) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * 256) {
        const V v = vals[e];
        tval[perm[e]] = v;
        if (perm_col) cval[perm_col[e]] = v;
        if (rval) rval[e] = v;
    }
}

static unsigned blocks_for(int64_t threads, int per) { return (unsigned)((threads + per - 1) / per); }

}  // namespace tsg

using namespace tsg;

// csr2tile (ts = 16) and csr2tile_ts (src/tile_format.h:225-324, ts 1-15):
// the same kernels, the tile size a compile-time 16 or a runtime value.
static int csr2tile_impl(Context *ctx, const tsg_csr *A, tsg_tile_matrix *T, int ts) {
    if (A->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "CSR must be on the device");
    if (A->m < 0 || A->n < 0 || A->nnz < 0) return fail(TSG_ERR_ARG, "negative sizes");
    (void)hipSetDevice(ctx->device);
    memset(T, 0, sizeof *T);
    T->m = A->m; T->n = A->n; T->nnz = A->nnz; T->dtype = A->dtype; T->location = TSG_DEVICE;
    T->tile_size = ts == 16 ? 0 : ts;
    T->tilem = (int32_t)(((int64_t)A->m + ts - 1) / ts); T->tilen = (int32_t)(((int64_t)A->n + ts - 1) / ts);
    if ((int64_t)A->m > (int64_t)ts * MAX_TILE_ROWS || (int64_t)A->n > (int64_t)ts * MAX_TILE_ROWS) {
        const int32_t m = A->m, n = A->n;
        memset(T, 0, sizeof *T);
        return fail(TSG_ERR_OVERFLOW, "%d x %d: more than %lld tile-rows or tile-columns", m, n,
                    (long long)MAX_TILE_ROWS);
    }
    hipStream_t s = ctx->stream;
    int32_t *flags = nullptr;
    int rc = palloc(ctx, &flags, 1);
    if (!rc) rc = palloc(ctx, &T->tile_ptr, (size_t)T->tilem + 1);
    if (rc) { pool_free(ctx, flags); return rc; }
    auto cleanup = [&](int code) {
        pool_free(ctx, flags);
        void *ps[] = {T->tile_ptr, T->tile_columnidx, T->tile_nnz, T->tile_csr_Val, T->tile_csr_Col,
                      T->tile_csr_Ptr, T->mask, T->row_ptr, T->row_col, T->row_val, T->val_perm};
        for (void *p : ps) pool_free(ctx, p);
        memset(T, 0, sizeof *T);
        return code;
    };
    if (hipMemsetAsync(flags, 0, sizeof(int32_t), s) != hipSuccess) return cleanup(fail(TSG_ERR_HIP, "memset"));
    if (T->tilem == 0) {
        (void)hipMemsetAsync(T->tile_ptr, 0, sizeof(int64_t), s);
    } else {
        if (ts == 16)
            k_c2t<0, 16><<<blocks_for(T->tilem, CV_WAVES), CV_WAVES * WAVE, 0, s>>>(
                A->rowptr, A->col, A->m, A->n, T->tilem, T->tile_ptr, nullptr, nullptr, nullptr, flags, ts);
        else
            k_c2t<0, 0><<<blocks_for(T->tilem, CV_WAVES), CV_WAVES * WAVE, 0, s>>>(
                A->rowptr, A->col, A->m, A->n, T->tilem, T->tile_ptr, nullptr, nullptr, nullptr, flags, ts);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(TSG_ERR_HIP, "k_c2t count launch"));
    }
    if ((rc = scan_inclusive_i64(ctx, T->tile_ptr, (int64_t)T->tilem + 1))) return cleanup(rc);
    if ((rc = read_i64(ctx, T->tile_ptr + T->tilem, &T->numtile))) return cleanup(rc);
    int32_t hf = 0;
    if ((rc = read_i32(ctx, flags, &hf))) return cleanup(rc);
    if (hf & 4) return cleanup(fail(TSG_ERR_ARG, "CSR column index outside [0, %d)", A->n));
    if (hf & 1) return cleanup(fail(TSG_ERR_UNSUPPORTED, "CSR rows must be sorted by column"));
    if (T->numtile > INT32_MAX) return cleanup(fail(TSG_ERR_OVERFLOW, "numtile exceeds int32 tile ids"));
    const int64_t nt = T->numtile;
    if ((rc = palloc(ctx, &T->tile_columnidx, (size_t)nt))) return cleanup(rc);
    if ((rc = palloc(ctx, &T->tile_nnz, (size_t)nt + 1))) return cleanup(rc);
    if ((rc = palloc(ctx, &T->tile_csr_Ptr, (size_t)nt * ts))) return cleanup(rc);
    if ((rc = palloc(ctx, &T->mask, (size_t)nt * ts))) return cleanup(rc);
    if ((rc = pool_alloc(ctx, &T->tile_csr_Val, (size_t)T->nnz * vsize(T->dtype)))) return cleanup(rc);
    if ((rc = palloc(ctx, &T->tile_csr_Col, (size_t)T->nnz))) return cleanup(rc);
    if ((rc = palloc(ctx, &T->val_perm, (size_t)T->nnz))) return cleanup(rc);
    if (hipMemsetAsync(T->tile_csr_Ptr, 0, (size_t)nt * ts, s) != hipSuccess ||
        hipMemsetAsync(T->mask, 0, (size_t)nt * ts * 2, s) != hipSuccess)
        return cleanup(fail(TSG_ERR_HIP, "memset"));
    if (T->tilem > 0) {
        if (ts == 16)
            k_c2t<1, 16><<<blocks_for(T->tilem, CV_WAVES), CV_WAVES * WAVE, 0, s>>>(
                A->rowptr, A->col, A->m, A->n, T->tilem, T->tile_ptr, T->tile_columnidx, T->tile_csr_Ptr, T->mask,
                flags, ts);
        else
            k_c2t<1, 0><<<blocks_for(T->tilem, CV_WAVES), CV_WAVES * WAVE, 0, s>>>(
                A->rowptr, A->col, A->m, A->n, T->tilem, T->tile_ptr, T->tile_columnidx, T->tile_csr_Ptr, T->mask,
                flags, ts);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(TSG_ERR_HIP, "k_c2t fill launch"));
    }
    if (nt > 0) {
        k_tile_rowscan<<<grid_capped(nt, 16), 256, 0, s>>>(nt, T->tile_csr_Ptr, T->tile_nnz, ts);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(TSG_ERR_HIP, "rowscan launch"));
    } else {
        (void)hipMemsetAsync(T->tile_nnz, 0, sizeof(int64_t), s);
    }
    if ((rc = scan_inclusive_i64(ctx, T->tile_nnz, nt + 1))) return cleanup(rc);
    if (T->tilem > 0 && T->nnz > 0) {
        const unsigned g = blocks_for(T->tilem, CV_WAVES);
        if (T->dtype == TSG_F32) {
            auto k = ts == 16 ? k_c2t_scatter<float, 16> : k_c2t_scatter<float, 0>;
            k<<<g, CV_WAVES * WAVE, 0, s>>>(A->rowptr, A->col, (const float *)A->val, A->m, T->tilem, T->tile_ptr,
                                            T->tile_columnidx, T->tile_nnz, T->tile_csr_Ptr,
                                            (float *)T->tile_csr_Val, T->tile_csr_Col, T->val_perm, ts);
        } else {
            auto k = ts == 16 ? k_c2t_scatter<double, 16> : k_c2t_scatter<double, 0>;
            k<<<g, CV_WAVES * WAVE, 0, s>>>(A->rowptr, A->col, (const double *)A->val, A->m, T->tilem, T->tile_ptr,
                                            T->tile_columnidx, T->tile_nnz, T->tile_csr_Ptr,
                                            (double *)T->tile_csr_Val, T->tile_csr_Col, T->val_perm, ts);
        }
        if (hipGetLastError() != hipSuccess) return cleanup(fail(TSG_ERR_HIP, "scatter launch"));
    }
    if (ts != 16) {  // no scalar CSR copy: no product runs at this tile size
        if ((rc = read_i32(ctx, flags, &hf))) return cleanup(rc);
        T->flags = ((hf & 2) ? 1 : 0) | ((hf & 8) ? 2 : 0) | 4;  // bit 2: tile-row lengths checked
        pool_free(ctx, flags);
        return TSG_OK;
    }
    // keep the scalar CSR beside the tiles (row-by-row product of sparse heavy rows)
    const size_t vb = (size_t)T->nnz * vsize(T->dtype);
    if ((rc = palloc(ctx, &T->row_ptr, (size_t)T->m + 1))) return cleanup(rc);
    if ((rc = palloc(ctx, &T->row_col, (size_t)T->nnz))) return cleanup(rc);
    if ((rc = pool_alloc(ctx, &T->row_val, vb))) return cleanup(rc);
    if (hipMemcpyAsync(T->row_ptr, A->rowptr, sizeof(int64_t) * ((size_t)T->m + 1), hipMemcpyDeviceToDevice, s) !=
            hipSuccess ||
        (T->nnz > 0 &&
         (hipMemcpyAsync(T->row_col, A->col, sizeof(int32_t) * (size_t)T->nnz, hipMemcpyDeviceToDevice, s) !=
              hipSuccess ||
          hipMemcpyAsync(T->row_val, A->val, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)))
        return cleanup(fail(TSG_ERR_HIP, "copy of the scalar CSR failed"));
    if ((rc = read_i32(ctx, flags, &hf))) return cleanup(rc);
    T->flags = ((hf & 2) ? 1 : 0) | ((hf & 8) ? 2 : 0) | 4;  // bit 2: tile-row lengths checked
    pool_free(ctx, flags);
    return TSG_OK;
}

extern "C" int tsg_csr2tile(tsg_context *c, const tsg_csr *A, tsg_tile_matrix *T) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !A || !T) return fail(TSG_ERR_ARG, "null argument");
    return csr2tile_impl(ctx, A, T, 16);
}

extern "C" int tsg_csr2tile_ts(tsg_context *c, const tsg_csr *A, int32_t ts, tsg_tile_matrix *T) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !A || !T) return fail(TSG_ERR_ARG, "null argument");
    if (ts < 1 || ts > 16)
        return fail(TSG_ERR_ARG, "tile size %d outside 1-16 (uint16 row masks and the (row<<4)|col byte)", ts);
    return csr2tile_impl(ctx, A, T, ts);
}

extern "C" int tsg_build_csc_tiles(tsg_context *c, tsg_tile_matrix *B) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !B) return fail(TSG_ERR_ARG, "null argument");
    if (B->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles must be on the device");
    if (B->is_csc) return fail(TSG_ERR_ARG, "already in CSC-tile order");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t nt = B->numtile;
    const int ts = B->tile_size ? B->tile_size : 16;  // build_csc_tiles_ts (src/tile_format.h:327-417) when not 16
    if (ts == 16 && (!aligned16(B->mask) || !aligned16(B->tile_csr_Ptr)))
        return fail(TSG_ERR_ARG, "row masks and row offsets must start on 16-byte boundaries (tiles are read as 16-byte words)");
    int32_t *rowidx = B->tile_rowidx, *tmp = nullptr, *crow = nullptr;
    int64_t *cptr = nullptr, *cursor = nullptr, *inv = nullptr, *nnz_new = nullptr, *pos_of = nullptr;
    uint8_t *ptr_new = nullptr, *col_new = nullptr;
    uint16_t *mask_new = nullptr;
    void *val_new = nullptr;
    bool own_rowidx = false;
    int rc = TSG_OK;
    if (!rowidx) {
        rc = palloc(ctx, &rowidx, (size_t)nt);
        own_rowidx = true;
        if (!rc && B->tilem > 0) {
            k_fill_rowidx<<<grid_capped(B->tilem, 4), 256, 0, s>>>(B->tile_ptr, B->tilem, rowidx);
            if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "rowidx launch");
        }
    }
    // the column-major copy of the CSR-order tiles for the staged numeric
    // kernels (16 x 16 tiles without repeated columns: the staged kernels are
    // not used otherwise), and the tile-row starts of the payload
    const bool cm_copy = !(B->flags & 1) && ts == 16;
    uint16_t *ccm = nullptr;
    uint8_t *ccp = nullptr;
    void *ccv = nullptr;
    int64_t *rnp = nullptr, *vpc = nullptr;
    if (!rc) rc = palloc(ctx, &cptr, (size_t)B->tilen + 1);
    if (!rc) rc = palloc(ctx, &cursor, (size_t)B->tilen + 1);
    if (!rc) rc = palloc(ctx, &tmp, (size_t)nt);
    if (!rc) rc = palloc(ctx, &crow, (size_t)nt);
    if (!rc && !cm_copy) rc = palloc(ctx, &inv, (size_t)nt);
    if (!rc) rc = palloc(ctx, &pos_of, (size_t)nt);
    if (!rc) rc = palloc(ctx, &nnz_new, (size_t)nt + 1);
    if (!rc) rc = palloc(ctx, &ptr_new, (size_t)nt * ts);
    if (!rc) rc = palloc(ctx, &mask_new, (size_t)nt * ts);
    if (!rc) rc = palloc(ctx, &col_new, (size_t)B->nnz);
    if (!rc) rc = pool_alloc(ctx, &val_new, (size_t)B->nnz * vsize(B->dtype));
    if (!rc && cm_copy) rc = palloc(ctx, &ccm, (size_t)nt * 16);
    if (!rc && cm_copy) rc = palloc(ctx, &ccp, (size_t)nt * 16);
    if (!rc && cm_copy) rc = pool_alloc(ctx, &ccv, (size_t)B->nnz * vsize(B->dtype));
    if (!rc && ts == 16) rc = palloc(ctx, &rnp, (size_t)B->tilem + 1);
    if (!rc && cm_copy && B->val_perm && B->nnz > 0) rc = palloc(ctx, &vpc, (size_t)B->nnz);
    if (!rc && hipMemsetAsync(cptr, 0, sizeof(int64_t) * ((size_t)B->tilen + 1), s) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "memset");
    if (!rc && nt > 0) {
        k_csc_count<<<blocks_for(nt, 256), 256, 0, s>>>(nt, B->tile_columnidx, cptr);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "csc count launch");
    }
    if (!rc) rc = scan_inclusive_i64(ctx, cptr, (int64_t)B->tilen + 1);
    if (!rc && hipMemcpyAsync(cursor, cptr, sizeof(int64_t) * ((size_t)B->tilen + 1), hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = fail(TSG_ERR_HIP, "copy");
    if (!rc && nt > 0) {
        k_csc_slot<<<blocks_for(nt, 256), 256, 0, s>>>(nt, B->tile_columnidx, rowidx, cursor, tmp);
        k_csc_sort<<<blocks_for(B->tilen, CV_WAVES), CV_WAVES * WAVE, 0, s>>>(B->tilen, cptr, tmp, crow);
        k_csc_perm<<<blocks_for(nt, 256), 256, 0, s>>>(nt, B->tile_columnidx, rowidx, cptr, crow, B->tile_nnz, inv, nnz_new,
                                                        pos_of);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "csc perm launches");
    } else if (!rc) {
        (void)hipMemsetAsync(nnz_new, 0, sizeof(int64_t), s);
    }
    if (!rc) rc = scan_inclusive_i64(ctx, nnz_new, nt + 1);
    if (!rc && nt > 0) {
        const bool f32 = B->dtype == TSG_F32;
        if (cm_copy && f32)
            k_csc_relayout<float><<<grid_capped(nt, 16), 256, 0, s>>>(
                nt, pos_of, B->tile_nnz, nnz_new, B->tile_csr_Ptr, B->mask, B->tile_csr_Col,
                (const float *)B->tile_csr_Val, ptr_new, mask_new, col_new, (float *)val_new, ccm, ccp, (float *)ccv);
        else if (cm_copy)
            k_csc_relayout<double><<<grid_capped(nt, 16), 256, 0, s>>>(
                nt, pos_of, B->tile_nnz, nnz_new, B->tile_csr_Ptr, B->mask, B->tile_csr_Col,
                (const double *)B->tile_csr_Val, ptr_new, mask_new, col_new, (double *)val_new, ccm, ccp,
                (double *)ccv);
        else if (f32)
            k_csc_gather<float><<<grid_capped(nt, 16), 256, 0, s>>>(
                nt, inv, B->tile_nnz, nnz_new, B->tile_csr_Ptr, B->mask, B->tile_csr_Col,
                (const float *)B->tile_csr_Val, ptr_new, mask_new, col_new, (float *)val_new, ts);
        else
            k_csc_gather<double><<<grid_capped(nt, 16), 256, 0, s>>>(
                nt, inv, B->tile_nnz, nnz_new, B->tile_csr_Ptr, B->mask, B->tile_csr_Col,
                (const double *)B->tile_csr_Val, ptr_new, mask_new, col_new, (double *)val_new, ts);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "csc gather launch");
    }
    if (!rc && rnp) {
        k_row_nnz_ptr<<<blocks_for((int64_t)B->tilem + 1, 256), 256, 0, s>>>(B->tilem, B->tile_ptr, B->tile_nnz, rnp);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "row_nnz_ptr launch");
    }
    // value positions of the CSR entries in the CSC-order tiles and in the column-major copy
    if (!rc && B->val_perm && B->nnz > 0) {
        k_val_perm_b<<<grid_capped(B->nnz, 256), 256, 0, s>>>(B->nnz, nt, B->tile_nnz, pos_of, nnz_new,
                                                              B->tile_csr_Col, ccm, ccp, B->val_perm, vpc);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "value positions launch");
    }
    // the work arrays go back to the pool in stream order: later work on this
    // stream that reuses them runs after the kernels above
    pool_free(ctx, cursor);
    pool_free(ctx, tmp);
    pool_free(ctx, inv);
    if (rc) {
        void *ps[] = {cptr, crow, nnz_new, ptr_new, mask_new, col_new, val_new, rnp, ccm, ccp, ccv, vpc, pos_of};
        for (void *q : ps) pool_free(ctx, q);
        if (own_rowidx) pool_free(ctx, rowidx);
        return rc;
    }
    pool_free(ctx, pos_of);
    // a within-tile CSC built in row-tile order no longer matches the data arrays
    pool_free(ctx, B->tile_csc_Ptr); pool_free(ctx, B->tile_csc_Row); pool_free(ctx, B->tile_csc_Val);
    B->tile_csc_Ptr = nullptr; B->tile_csc_Row = nullptr; B->tile_csc_Val = nullptr;
    B->val_perm_col = vpc;
    pool_free(ctx, B->tile_csr_Col);
    pool_free(ctx, B->tile_csr_Ptr);
    pool_free(ctx, B->tile_csr_Val);
    if (ts == 16) {
        B->csr_tile_nnz = B->tile_nnz;
        B->csr_mask = B->mask;
    } else {  // the row-staged kernels (the only readers of the CSR-order copy) need 16-row tiles
        pool_free(ctx, B->tile_nnz);
        pool_free(ctx, B->mask);
    }
    B->csr_col_mask = ccm;
    B->csr_col_Ptr = ccp;
    B->csr_col_Val = ccv;
    B->row_nnz_ptr = rnp;
    B->tile_nnz = nnz_new;
    B->tile_csr_Ptr = ptr_new;
    B->mask = mask_new;
    B->tile_csr_Col = col_new;
    B->tile_csr_Val = val_new;
    B->csc_tile_ptr = cptr;
    B->csc_tile_rowidx = crow;
    B->tile_rowidx = rowidx;
    B->is_csc = 1;
    return TSG_OK;
}

// ---------------------------------------------------------------- within-tile CSC
namespace tsg {

// A 16-lane group per tile, lane = local column c.  Without repeated columns,
// column c holds row r exactly when bit 15-c of mask[r] is set, and that entry
// sits at tile_csr_Ptr[r] + (set bits of mask[r] left of c); rows ascending
// then give the column's entries in the order the reference's counting
// scatter writes them (src/tile_format.h:567-582, :609-625).  With repeated
// columns (flags bit 0) the lane walks the tile's col bytes in storage order
// and keeps those of column c, which is that same order.
template <typename V, bool DUPS>
__global__ void __launch_bounds__(256)
k_tile_csc(int64_t nt, const int64_t *__restrict__ tile_nnz, const uint8_t *__restrict__ ptr,
           const uint16_t *__restrict__ mask, const uint8_t *__restrict__ colb, const V *__restrict__ val,
           uint8_t *__restrict__ cptr, uint8_t *__restrict__ crow, V *__restrict__ cval) {
    const int c = lane_id() & 15;
    const uint32_t bit = 0x8000u >> c, left = (0xFFFF0000u >> c) & 0xFFFFu;
    for (int64_t t = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4); t < nt; t += (int64_t)gridDim.x * 16) {
        const int64_t base = tile_nnz[t];
        const int tn = (int)(tile_nnz[t + 1] - base);
        uint32_t m[16];
#pragma unroll
        for (int r = 0; r < 16; r++) m[r] = mask[t * 16 + r];
        int cnt = 0;
        if (DUPS) {
            for (int e = 0; e < tn; e++) cnt += (colb[base + e] & 15) == c;
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) cnt += (m[r] & bit) ? 1 : 0;
        }
        int o = group16_incl_sum(cnt) - cnt;
        cptr[t * 16 + c] = (uint8_t)o;
        if (DUPS) {
            int r = 0;
            for (int e = 0; e < tn; e++) {
                while (r < 15 && e >= (int)ptr[t * 16 + r + 1]) r++;  // the row holding entry e
                if ((colb[base + e] & 15) == c) {
                    crow[base + o] = (uint8_t)r;
                    cval[base + o] = val[base + e];
                    o++;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if (m[r] & bit) {
                    const int e = (int)ptr[t * 16 + r] + __popc(m[r] & left);
                    crow[base + o] = (uint8_t)r;
                    cval[base + o] = val[base + e];
                    o++;
                }
            }
        }
    }
}

template <typename V>
static int launch_tile_csc_t(Context *ctx, const tsg_tile_matrix *T, uint8_t *cptr, uint8_t *crow, void *cval) {
    const unsigned grid = grid_capped(T->numtile, 16);
    if (T->flags & 1)
        k_tile_csc<V, true><<<grid, 256, 0, ctx->stream>>>(T->numtile, T->tile_nnz, T->tile_csr_Ptr, T->mask,
                                                           T->tile_csr_Col, (const V *)T->tile_csr_Val, cptr, crow,
                                                           (V *)cval);
    else
        k_tile_csc<V, false><<<grid, 256, 0, ctx->stream>>>(T->numtile, T->tile_nnz, T->tile_csr_Ptr, T->mask,
                                                            T->tile_csr_Col, (const V *)T->tile_csr_Val, cptr, crow,
                                                            (V *)cval);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

// Fills the within-tile CSC arrays of T (allocated by the caller).
static int launch_tile_csc(Context *ctx, const tsg_tile_matrix *T, uint8_t *cptr, uint8_t *crow, void *cval) {
    if (T->numtile == 0) return TSG_OK;
    if (T->dtype == TSG_F32) return launch_tile_csc_t<float>(ctx, T, cptr, crow, cval);
    return launch_tile_csc_t<double>(ctx, T, cptr, crow, cval);
}

}  // namespace tsg

extern "C" int tsg_build_tile_csc(tsg_context *c, tsg_tile_matrix *T) {
    using namespace tsg;
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !T) return fail(TSG_ERR_ARG, "null argument");
    if (T->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles must be on the device");
    if (!T->mask || !T->tile_csr_Ptr || !T->tile_nnz) return fail(TSG_ERR_ARG, "tiles lack masks or row offsets");
    if (T->tile_size && T->tile_size != 16) return fail(TSG_ERR_UNSUPPORTED, "within-tile CSC needs 16 x 16 tiles");
    (void)hipSetDevice(ctx->device);
    uint8_t *cptr = nullptr, *crow = nullptr;
    void *cval = nullptr;
    int rc = palloc(ctx, &cptr, (size_t)T->numtile * 16);
    if (!rc) rc = palloc(ctx, &crow, (size_t)T->nnz);
    if (!rc) rc = pool_alloc(ctx, &cval, (size_t)T->nnz * vsize(T->dtype));
    if (!rc) rc = launch_tile_csc(ctx, T, cptr, crow, cval);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "within-tile CSC failed");
    if (rc) {
        pool_free(ctx, cptr); pool_free(ctx, crow); pool_free(ctx, cval);
        return rc;
    }
    pool_free(ctx, T->tile_csc_Ptr); pool_free(ctx, T->tile_csc_Row); pool_free(ctx, T->tile_csc_Val);
    T->tile_csc_Ptr = cptr;
    T->tile_csc_Row = crow;
    T->tile_csc_Val = cval;
    return TSG_OK;
}

// ---------------------------------------------------------------- tile2csr
namespace tsg {

// One wave per tile-row.  For each of the 16 rows the wave scans the row
// popcounts of the tile-row's tiles (ascending tile column) to place each
// tile's segment, so CSR rows come out sorted by column.
template <bool FILL, typename V>
__global__ void __launch_bounds__(CV_WAVES *WAVE)
k_t2c(int32_t m, int32_t tilem, const int64_t *__restrict__ tile_ptr, const int32_t *__restrict__ colidx,
      const int64_t *__restrict__ tile_nnz, const uint8_t *__restrict__ ptr, const uint16_t *__restrict__ mask,
      const uint8_t *__restrict__ tcol, const V *__restrict__ tval, int64_t *__restrict__ rowptr,
      int32_t *__restrict__ col, V *__restrict__ val) {
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int32_t bi = blockIdx.x * CV_WAVES + w;
    if (bi >= tilem) return;
    const int64_t tb = tile_ptr[bi], te = tile_ptr[bi + 1];
    for (int r = 0; r < 16 && bi * 16 + r < m; r++) {
        const int32_t row = bi * 16 + r;
        int64_t base = FILL ? rowptr[row] : 0;
        for (int64_t t0 = tb; t0 < te; t0 += WAVE) {
            const int64_t T = t0 + lane;
            int pc = 0;
            if (T < te) pc = __popc((uint32_t)mask[T * 16 + r]);
            const int incl = wave_incl_sum(pc);
            if (FILL && pc) {
                const int64_t src = tile_nnz[T] + ptr[T * 16 + r];
                const int64_t dst = base + incl - pc;
                const int32_t cb = colidx[T] * 16;
                for (int q = 0; q < pc; q++) {
                    col[dst + q] = cb + (tcol[src + q] & 15);
                    val[dst + q] = tval[src + q];
                }
            }
            base += __shfl(incl, WAVE - 1, WAVE);
        }
        if (!FILL && lane == 0) {
            rowptr[row + 1] = base;
            if (row == 0) rowptr[0] = 0;
        }
    }
}

template <typename V>
static int tile2csr_t(Context *ctx, const tsg_tile_matrix *T, tsg_csr *o) {
    hipStream_t s = ctx->stream;
    const unsigned grid = blocks_for(T->tilem, CV_WAVES);
    if (T->tilem > 0) {
        k_t2c<false, V><<<grid, CV_WAVES * WAVE, 0, s>>>(T->m, T->tilem, T->tile_ptr, T->tile_columnidx,
                                                         T->tile_nnz, T->tile_csr_Ptr, T->mask, T->tile_csr_Col,
                                                         (const V *)T->tile_csr_Val, o->rowptr, nullptr, nullptr);
        TSG_HIP(hipGetLastError());
    } else {
        TSG_HIP(hipMemsetAsync(o->rowptr, 0, sizeof(int64_t), s));
    }
    TSG_CHECK(scan_inclusive_i64(ctx, o->rowptr, (int64_t)T->m + 1));
    if (T->tilem > 0 && T->nnz > 0) {
        k_t2c<true, V><<<grid, CV_WAVES * WAVE, 0, s>>>(T->m, T->tilem, T->tile_ptr, T->tile_columnidx,
                                                        T->tile_nnz, T->tile_csr_Ptr, T->mask, T->tile_csr_Col,
                                                        (const V *)T->tile_csr_Val, o->rowptr, o->col, (V *)o->val);
        TSG_HIP(hipGetLastError());
    }
    TSG_HIP(hipStreamSynchronize(s));
    return TSG_OK;
}

}  // namespace tsg

extern "C" int tsg_tile2csr(tsg_context *c, const tsg_tile_matrix *T, tsg_csr *o) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !T || !o) return fail(TSG_ERR_ARG, "null argument");
    if (T->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles must be on the device");
    if (T->is_csc) return fail(TSG_ERR_ARG, "tile2csr needs row-tile order (is_csc = 0)");
    if (T->tile_size && T->tile_size != 16) return fail(TSG_ERR_UNSUPPORTED, "tile2csr needs 16 x 16 tiles");
    (void)hipSetDevice(ctx->device);
    memset(o, 0, sizeof *o);
    o->m = T->m; o->n = T->n; o->nnz = T->nnz; o->dtype = T->dtype; o->location = TSG_DEVICE;
    int rc = palloc(ctx, &o->rowptr, (size_t)T->m + 1);
    if (!rc) rc = palloc(ctx, &o->col, (size_t)T->nnz);
    if (!rc) rc = pool_alloc(ctx, &o->val, (size_t)T->nnz * vsize(T->dtype));
    if (!rc) rc = T->dtype == TSG_F32 ? tile2csr_t<float>(ctx, T, o) : tile2csr_t<double>(ctx, T, o);
    if (rc) {
        pool_free(ctx, o->rowptr); pool_free(ctx, o->col); pool_free(ctx, o->val);
        memset(o, 0, sizeof *o);
    }
    return rc;
}

extern "C" int tsg_tile_set_values(tsg_context *c, tsg_tile_matrix *T, const void *vals) {
    using namespace tsg;
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !T || (!vals && T->nnz > 0)) return fail(TSG_ERR_ARG, "null argument");
    if (T->location != TSG_DEVICE) return fail(TSG_ERR_ARG, "tiles are not on the device");
    if (!T->val_perm) return fail(TSG_ERR_ARG, "no value positions (a matrix made by tsg_csr2tile keeps them)");
    if (T->csr_col_Val && !T->val_perm_col) return fail(TSG_ERR_ARG, "column-major copy without value positions");
    (void)hipSetDevice(ctx->device);
    if (T->nnz == 0) return TSG_OK;
    const unsigned grid = grid_capped(T->nnz, 256);
    if (T->dtype == TSG_F32)
        k_set_values<float><<<grid, 256, 0, ctx->stream>>>(T->nnz, (const float *)vals, T->val_perm, T->val_perm_col,
                                                            (float *)T->tile_csr_Val, (float *)T->csr_col_Val,
                                                            (float *)T->row_val);
    else
        k_set_values<double><<<grid, 256, 0, ctx->stream>>>(T->nnz, (const double *)vals, T->val_perm,
                                                             T->val_perm_col, (double *)T->tile_csr_Val,
                                                             (double *)T->csr_col_Val, (double *)T->row_val);
    TSG_HIP(hipGetLastError());
    if (T->tile_csc_Val) TSG_CHECK(launch_tile_csc(ctx, T, T->tile_csc_Ptr, T->tile_csc_Row, T->tile_csc_Val));
    return TSG_OK;
}
