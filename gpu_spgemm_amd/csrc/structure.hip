// structure.hip -- Step 1: tile structure of C.
//
// Replaces step1_count_kernel / step1_fill_kernel
// (src/spgemm_raw_kernels.cuh:23-102).  C tile-row bi is the union, over the
// tile columns k of A's tile-row bi, of B's CSR-tile columns of tile-row k
// (the reference passes B's CSR tile_ptr/tile_columnidx here,
// src/spgemm_raw_v2.cu:127-128), in ascending order.
//
// MI355X design: one wave64 per tile-row.  The candidate columns of all A
// tiles of the row are flattened across the 64 lanes (an inclusive prefix of
// B row lengths in LDS, one lane per candidate), and inserted into an LDS
// bitmap that covers only the row's own column window [cmin, cmax] (8 KB per
// wave = 65,536 tile columns per pass; wider rows take several passes).  The
// reference's bitmap spans all of tilenB, which needs 128 KB of shared memory
// per block on the 2048x2048 Laplacian; this window needs 17 words there.
// Count = popcount; fill = wave prefix over per-lane word chunks, so each lane
// writes its columns in ascending order at its own offset.
#include <climits>

#include "tsg_internal.h"

namespace tsg {

constexpr int ST_WAVES = 4;
constexpr int ST_WIN_WORDS = 2048;  // 65,536 bits

template <bool FILL>
__global__ void __launch_bounds__(ST_WAVES *WAVE)
k_structure(const int64_t *__restrict__ tile_ptrA, const int32_t *__restrict__ colidxA, int32_t tilemA,
            const int64_t *__restrict__ tile_ptrB, const int32_t *__restrict__ colidxB,
            int64_t *__restrict__ tile_ptrC, int32_t *__restrict__ rowidxC,
            int32_t *__restrict__ colidxC, const int32_t *__restrict__ rows, int64_t row_begin,
            int64_t nrows, int32_t *__restrict__ row_nc, int32_t *__restrict__ overflow) {
    __shared__ uint32_t s_bm[ST_WAVES][ST_WIN_WORDS];
    __shared__ int64_t s_bs[ST_WAVES][WAVE];
    __shared__ int32_t s_pre[ST_WAVES][WAVE];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t idx = (int64_t)blockIdx.x * ST_WAVES + w;
    if (idx >= (rows ? nrows : (int64_t)tilemA)) return;  // whole wave leaves; no block barrier below
    const int32_t bi = rows ? rows[row_begin + idx] : (int32_t)idx;
    uint32_t *bm = s_bm[w];
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];

    int32_t cmin = INT_MAX, cmax = -1;
    for (int64_t i = as + lane; i < ae; i += WAVE) {
        int32_t k = colidxA[i];
        int64_t bs = tile_ptrB[k], be = tile_ptrB[k + 1];
        if (be > bs) {
            cmin = min(cmin, colidxB[bs]);
            cmax = max(cmax, colidxB[be - 1]);
        }
    }
    cmin = wave_min(cmin);
    cmax = wave_max(cmax);

    int64_t count = 0;
    int64_t out = FILL ? tile_ptrC[bi] : 0;
    // slot layout: the row's slots end at tile_ptrC[bi + 1]; an overfull row
    // writes nothing past them and raises the overflow flag
    const int64_t slot_end = (FILL && row_nc) ? tile_ptrC[bi + 1] : INT64_MAX;
    for (int64_t w0 = cmin; w0 <= cmax; w0 += (int64_t)ST_WIN_WORDS * 32) {
        const int32_t wbits = (int32_t)min<int64_t>((int64_t)ST_WIN_WORDS * 32, cmax - w0 + 1);
        const int nwords = (wbits + 31) >> 5;
        for (int i = lane; i < nwords; i += WAVE) bm[i] = 0;
        wave_sync();
        for (int64_t c0 = as; c0 < ae; c0 += WAVE) {
            int64_t i = c0 + lane;
            int32_t len = 0;
            int64_t bs = 0;
            if (i < ae) {
                int32_t k = colidxA[i];
                bs = tile_ptrB[k];
                len = (int32_t)(tile_ptrB[k + 1] - bs);
            }
            int32_t incl = wave_incl_sum(len);
            int32_t total = __shfl(incl, WAVE - 1, WAVE);
            s_pre[w][lane] = incl;
            s_bs[w][lane] = bs;
            wave_sync();
            for (int32_t t = lane; t < total; t += WAVE) {
                int lo = 0, hi = WAVE - 1;  // first owner with inclusive prefix > t
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if (s_pre[w][mid] > t) hi = mid; else lo = mid + 1;
                }
                int32_t excl = lo ? s_pre[w][lo - 1] : 0;
                int64_t off = (int64_t)colidxB[s_bs[w][lo] + (t - excl)] - w0;
                if (off >= 0 && off < wbits) atomicOr(&bm[off >> 5], 1u << (off & 31));
            }
            wave_sync();
        }
        if (!FILL) {
            int32_t cnt = 0;
            for (int i = lane; i < nwords; i += WAVE) cnt += __popc(bm[i]);
            count += wave_sum(cnt);
        } else {
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
                    if (pos < slot_end) {
                        rowidxC[pos] = bi;
                        colidxC[pos] = (int32_t)(w0 + i * 32 + bit);
                    }
                    pos++;
                }
            }
            out += tot;
        }
        wave_sync();
    }
    if (!FILL && lane == 0) {
        tile_ptrC[bi + 1] = count;
        if (bi == 0) tile_ptrC[0] = 0;
    }
    if (FILL && row_nc && lane == 0) {
        // an overfull row keeps only its slots' worth (its results are discarded)
        const int64_t n = min<int64_t>(out, slot_end) - tile_ptrC[bi];
        row_nc[bi] = (int32_t)min<int64_t>(n, INT32_MAX);
        if (out > slot_end) atomicOr(overflow, 1);
    }
}

int launch_structure_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                           int64_t *tile_ptrC) {
    if (A->tilem == 0) {
        TSG_HIP(hipMemsetAsync(tile_ptrC, 0, sizeof(int64_t), ctx->stream));
        return TSG_OK;
    }
    unsigned grid = (unsigned)((A->tilem + ST_WAVES - 1) / ST_WAVES);
    k_structure<false><<<grid, ST_WAVES * WAVE, 0, ctx->stream>>>(
        A->tile_ptr, A->tile_columnidx, A->tilem, B->tile_ptr, B->tile_columnidx, tile_ptrC, nullptr,
        nullptr, nullptr, 0, 0, nullptr, nullptr);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int launch_structure_fill(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                          const int64_t *tile_ptrC, int32_t *rowidxC, int32_t *colidxC) {
    if (A->tilem == 0) return TSG_OK;
    unsigned grid = (unsigned)((A->tilem + ST_WAVES - 1) / ST_WAVES);
    k_structure<true><<<grid, ST_WAVES * WAVE, 0, ctx->stream>>>(
        A->tile_ptr, A->tile_columnidx, A->tilem, B->tile_ptr, B->tile_columnidx,
        const_cast<int64_t *>(tile_ptrC), rowidxC, colidxC, nullptr, 0, 0, nullptr, nullptr);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int launch_structure_fill_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                               const tsg_tile_matrix *C, const int32_t *rows, int64_t begin, int64_t count,
                               const SlotLayout &sl) {
    if (count <= 0) return TSG_OK;
    unsigned grid = (unsigned)((count + ST_WAVES - 1) / ST_WAVES);
    k_structure<true><<<grid, ST_WAVES * WAVE, 0, ctx->stream>>>(
        A->tile_ptr, A->tile_columnidx, A->tilem, B->tile_ptr, B->tile_columnidx, C->tile_ptr, C->tile_rowidx,
        C->tile_columnidx, rows, begin, count, sl.row_nc, sl.overflow);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

}  // namespace tsg
