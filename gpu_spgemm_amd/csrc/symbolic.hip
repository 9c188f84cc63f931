// symbolic.hip -- Step 2: per-row bitmasks of every C tile.
//
// Replaces step2_symbolic_kernel and step2_build_ptrRowC_kernel
// (src/spgemm_raw_kernels.cuh:105-215):
//   maskC[r]  = OR over matched pairs (a,b), OR over the columns kk present in
//               A-tile row r, of B-tile row mask kk
//   tile_nnzC = sum_r popcount(maskC[r])   (count stored at [ti+1] for the scan)
//   ptrRowC[r]= exclusive prefix of the row popcounts (uint8)
//
// MI355X design: a 16-lane group per C tile (four C tiles per wave64), lane =
// tile row.  Each pair stages B's 16 row masks (32 B, one coalesced load) in
// LDS; each lane walks the set bits of its A row mask (bit 15-kk) with clz and
// ORs the staged rows -- no per-entry column bytes are read.  The per-tile
// exclusive scan of popcounts is a 16-lane shuffle scan in the same kernel,
// so ptrRowC costs no extra launch.  No pair cache is written: the numeric
// step recomputes the (cheap, L2-resident) intersection, so there is no
// capacity to overflow and no host read of a pair count.
#include "pairs.h"

namespace tsg {

constexpr int SY_THREADS = 256;
constexpr int SY_GROUPS = SY_THREADS / 16;

// Symbolic step of one C tile by a 16-lane group (lane = tile row); s_mb is
// the group's 16-entry LDS slot for B's row masks.
__device__ __forceinline__ void sym_tile(int64_t ti, const int64_t *__restrict__ tile_ptrA,
                                         const int32_t *__restrict__ colidxA, const uint16_t *__restrict__ maskA,
                                         const int64_t *__restrict__ csc_ptrB, const int32_t *__restrict__ csc_rowB,
                                         const uint16_t *__restrict__ maskB, const int32_t *__restrict__ rowidxC,
                                         const int32_t *__restrict__ colidxC, uint16_t *__restrict__ maskC,
                                         uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC,
                                         uint16_t *s_mb) {
    const int l = lane_id() & 15;
    const int32_t bi = rowidxC[ti], bj = colidxC[ti];
    uint32_t rm = 0;
    for_each_pair(tile_ptrA, colidxA, csc_ptrB, csc_rowB, bi, bj, [&](int64_t a, int64_t b) {
        s_mb[l] = maskB[b * 16 + l];
        wave_sync();
        uint32_t ma = maskA[a * 16 + l];
        while (ma) {
            const int kk = __clz(ma) - 16;
            rm |= s_mb[kk];
            ma ^= 0x8000u >> kk;
        }
        wave_sync();
    });
    maskC[ti * 16 + l] = (uint16_t)rm;
    const int pc = __popc(rm);
    const int incl = group16_incl_sum(pc);
    ptrRowC[ti * 16 + l] = (uint8_t)(incl - pc);
    if (l == 15) tile_nnzC[ti + 1] = incl;
}

__global__ void __launch_bounds__(SY_THREADS)
k_symbolic(const int64_t *__restrict__ tile_ptrA, const int32_t *__restrict__ colidxA,
           const uint16_t *__restrict__ maskA, const int64_t *__restrict__ csc_ptrB,
           const int32_t *__restrict__ csc_rowB, const uint16_t *__restrict__ maskB,
           int64_t numtileC, const int32_t *__restrict__ rowidxC, const int32_t *__restrict__ colidxC,
           uint16_t *__restrict__ maskC, uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC) {
    __shared__ uint16_t s_mb[SY_GROUPS][16];
    const int g = threadIdx.x >> 4;
    // grid-stride over tiles (groups leave independently; only wave-level syncs below)
    for (int64_t ti = (int64_t)blockIdx.x * SY_GROUPS + g; ti < numtileC; ti += (int64_t)gridDim.x * SY_GROUPS)
        sym_tile(ti, tile_ptrA, colidxA, maskA, csc_ptrB, csc_rowB, maskB, rowidxC, colidxC, maskC, ptrRowC,
                 tile_nnzC, s_mb[g]);
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_nnzC[0] = 0;
}

// Heavy tile-rows: one workgroup per listed row, its 16 groups stride over the
// row's C tiles.
__global__ void __launch_bounds__(SY_THREADS)
k_symbolic_list(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nrows, const int64_t *__restrict__ dlo,
                const int64_t *__restrict__ dhi, int64_t c0, const int64_t *__restrict__ tile_ptrC,
                const int64_t *__restrict__ tile_ptrA, const int32_t *__restrict__ colidxA,
                const uint16_t *__restrict__ maskA, const int64_t *__restrict__ csc_ptrB,
                const int32_t *__restrict__ csc_rowB, const uint16_t *__restrict__ maskB,
                const int32_t *__restrict__ rowidxC, const int32_t *__restrict__ colidxC,
                uint16_t *__restrict__ maskC, uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC,
                const int32_t *__restrict__ row_nc) {
    __shared__ uint16_t s_mb[SY_GROUPS][16];
    const int g = threadIdx.x >> 4;
    const int32_t bi = bin_row(rows, row_begin, nrows, dlo, dhi, c0 + blockIdx.x);
    if (bi < 0) return;
    int64_t cs, ce;
    c_row_range(tile_ptrC, row_nc, bi, cs, ce);
    for (int64_t ti = cs + g; ti < ce; ti += SY_GROUPS)
        sym_tile(ti, tile_ptrA, colidxA, maskA, csc_ptrB, csc_rowB, maskB, rowidxC, colidxC, maskC, ptrRowC,
                 tile_nnzC, s_mb[g]);
}

int launch_symbolic_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const int32_t *rows,
                         const BinRange &br, const tsg_tile_matrix *C, const SlotLayout &sl) {
    const int64_t count = br.count;
    for (int64_t c0 = 0; c0 < count; c0 += ROW_CHUNK) {  // work-items per launch stay below 2^32
        const unsigned n = (unsigned)(count - c0 < ROW_CHUNK ? count - c0 : ROW_CHUNK);
        k_symbolic_list<<<n, SY_THREADS, 0, ctx->stream>>>(
            rows, br.begin, count, br.lo, br.hi, c0, C->tile_ptr, A->tile_ptr, A->tile_columnidx, A->mask, B->csc_tile_ptr,
            B->csc_tile_rowidx, B->mask, C->tile_rowidx, C->tile_columnidx, C->mask, C->tile_csr_Ptr, C->tile_nnz,
            sl.row_nc);
        TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

int launch_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, int64_t numtileC,
                    const int32_t *rowidxC, const int32_t *colidxC, uint16_t *maskC, uint8_t *ptrRowC,
                    int64_t *tile_nnzC) {
    if (numtileC == 0) {
        TSG_HIP(hipMemsetAsync(tile_nnzC, 0, sizeof(int64_t), ctx->stream));
        return TSG_OK;
    }
    k_symbolic<<<grid_capped(numtileC, SY_GROUPS), SY_THREADS, 0, ctx->stream>>>(
        A->tile_ptr, A->tile_columnidx, A->mask, B->csc_tile_ptr, B->csc_tile_rowidx, B->mask, numtileC,
        rowidxC, colidxC, maskC, ptrRowC, tile_nnzC);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

}  // namespace tsg
