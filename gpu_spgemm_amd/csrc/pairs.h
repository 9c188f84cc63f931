// pairs.h -- matched tile pairs of one C tile, for a 16-lane group.
//
// For C tile (bi,bj) the pairs are the k present both in A's tile-row bi
// (tile_columnidx) and in B's tile-column bj (csc_tile_rowidx).  As in the
// reference (src/spgemm_raw_kernels.cuh:146-158) the shorter list drives and
// each of its entries is binary-searched in the longer one; unlike the
// reference, matches are compacted with a wave ballot in lane order, so pairs
// are visited in ascending k with no LDS atomics and no 32-pair cap (the
// reference's merge fallback, :188-195, is never needed) -- which makes the
// floating-point summation order deterministic.
#pragma once
#include "tsg_internal.h"

namespace tsg {

// All 16 lanes of an aligned group call this together with group-uniform
// bi/bj; f(a_tile, b_tile) runs on all 16 lanes for each pair, ascending k.
template <typename F>
__device__ __forceinline__ void for_each_pair(const int64_t *__restrict__ tile_ptrA,
                                              const int32_t *__restrict__ colidxA,
                                              const int64_t *__restrict__ csc_ptrB,
                                              const int32_t *__restrict__ csc_rowB, int32_t bi,
                                              int32_t bj, F &&f) {
    const int l = lane_id() & 15, gbase = lane_id() & 48;
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];
    const int64_t bs = csc_ptrB[bj], be = csc_ptrB[bj + 1];
    const bool a_short = (ae - as) <= (be - bs);
    const int32_t *sl = a_short ? colidxA + as : csc_rowB + bs;
    const int32_t *ll = a_short ? csc_rowB + bs : colidxA + as;
    const int64_t sn = a_short ? ae - as : be - bs, ln = a_short ? be - bs : ae - as;
    const int64_t sbase = a_short ? as : bs, lbase = a_short ? bs : as;
    if (ln == 0) return;
    for (int64_t c0 = 0; c0 < sn; c0 += 16) {
        const int64_t i = c0 + l;
        int64_t p = -1;
        if (i < sn) {
            const int32_t key = sl[i];
            const int64_t q = lower_bound<int64_t, int32_t>(ll, 0, ln, key);
            if (q < ln && ll[q] == key) p = q;
        }
        uint32_t gm = (uint32_t)(__ballot(p >= 0) >> gbase) & 0xFFFFu;
        while (gm) {
            const int src = __ffs(gm) - 1;
            gm &= gm - 1;
            const int64_t pl = __shfl(p, gbase + src, WAVE);
            const int64_t ps = c0 + src;
            const int64_t a_idx = a_short ? sbase + ps : lbase + pl;
            const int64_t b_idx = a_short ? lbase + pl : sbase + ps;
            f(a_idx, b_idx);
        }
    }
}

}  // namespace tsg
