// numeric.hip -- Step 3: values of C.
//
// Replaces step3_numeric_kernel_v2 (src/spgemm_raw_kernels_v2.cuh:10-139):
// for every matched pair (ascending k), every A entry (r,kk,va) and every B
// entry (kk,c,vb): acc[r][c] += va*vb, then the accumulated values of the set
// bits of maskC[r] are written, ascending c, at tile_nnzC + ptrRowC[r] with
// col byte (r<<4)|c (:120-137).  Empty tiles write nothing (:33).
//
// MI355X design: a 16-lane group per C tile, lane = tile row, 16 register
// accumulators per lane.  The reference selects an accumulator with a
// 16-way switch per product (:54-67); here each A entry reads B row kk's mask
// and offset and runs a fully unrolled loop over the 16 columns whose index is
// a compile-time constant, so the accumulator is a fixed register and only the
// loads and FMAs of present columns execute (predicated).  One fused
// multiply-add per product, pairs in ascending k and A entries in storage
// order: the summation order is fixed, so results are reproducible run to run.
#include "pairs.h"

namespace tsg {

constexpr int NU_THREADS = 256;
constexpr int NU_GROUPS = NU_THREADS / 16;

template <typename V>
__device__ __forceinline__ V fmav(V a, V b, V c);
template <>
__device__ __forceinline__ double fmav<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fmav<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename V>
__global__ void __launch_bounds__(NU_THREADS)
k_numeric_rows(const int64_t *__restrict__ tile_ptrA, const int32_t *__restrict__ colidxA,
               const int64_t *__restrict__ tile_nnzA, const uint8_t *__restrict__ ptrRowA,
               const uint8_t *__restrict__ colA, const V *__restrict__ valA,
               const int64_t *__restrict__ csc_ptrB, const int32_t *__restrict__ csc_rowB,
               const int64_t *__restrict__ tile_nnzB, const uint8_t *__restrict__ ptrRowB,
               const uint16_t *__restrict__ maskB, const V *__restrict__ valB, int64_t numtileC,
               const int32_t *__restrict__ rowidxC, const int32_t *__restrict__ colidxC,
               const int64_t *__restrict__ tile_nnzC, const uint16_t *__restrict__ maskC,
               const uint8_t *__restrict__ ptrRowC, V *__restrict__ valC, uint8_t *__restrict__ colC) {
    const int g = threadIdx.x >> 4, l = threadIdx.x & 15;
    const int64_t ti = (int64_t)blockIdx.x * NU_GROUPS + g;
    if (ti >= numtileC) return;
    const int64_t c0 = tile_nnzC[ti];
    if (tile_nnzC[ti + 1] == c0) return;  // structurally empty tile
    const int32_t bi = rowidxC[ti], bj = colidxC[ti];
    V acc[16];
#pragma unroll
    for (int c = 0; c < 16; c++) acc[c] = V(0);
    for_each_pair(tile_ptrA, colidxA, csc_ptrB, csc_rowB, bi, bj, [&](int64_t a, int64_t b) {
        const int64_t an0 = tile_nnzA[a];
        const int e0 = ptrRowA[a * 16 + l];
        const int e1 = l == 15 ? (int)(tile_nnzA[a + 1] - an0) : (int)ptrRowA[a * 16 + l + 1];
        const int64_t bn0 = tile_nnzB[b];
        for (int e = e0; e < e1; e++) {
            const int kk = colA[an0 + e] & 15;
            const V va = valA[an0 + e];
            const uint32_t mb = maskB[b * 16 + kk];
            int64_t pb = bn0 + ptrRowB[b * 16 + kk];
#pragma unroll
            for (int c = 0; c < 16; c++) {
                if (mb & (0x8000u >> c)) {
                    acc[c] = fmav<V>(va, valB[pb], acc[c]);
                    pb++;
                }
            }
        }
    });
    const uint32_t mc = maskC[ti * 16 + l];
    int64_t o = c0 + ptrRowC[ti * 16 + l];
#pragma unroll
    for (int c = 0; c < 16; c++) {
        if (mc & (0x8000u >> c)) {
            valC[o] = acc[c];
            colC[o] = (uint8_t)((l << 4) | c);
            o++;
        }
    }
}

template <typename V>
static int launch_rows(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C) {
    int64_t grid = (C->numtile + NU_GROUPS - 1) / NU_GROUPS;
    if (grid > INT32_MAX) return fail(TSG_ERR_OVERFLOW, "numeric grid too large");
    k_numeric_rows<V><<<(unsigned)grid, NU_THREADS, 0, ctx->stream>>>(
        A->tile_ptr, A->tile_columnidx, A->tile_nnz, A->tile_csr_Ptr, A->tile_csr_Col,
        (const V *)A->tile_csr_Val, B->csc_tile_ptr, B->csc_tile_rowidx, B->tile_nnz, B->tile_csr_Ptr,
        B->mask, (const V *)B->tile_csr_Val, C->numtile, C->tile_rowidx, C->tile_columnidx, C->tile_nnz,
        C->mask, C->tile_csr_Ptr, (V *)C->tile_csr_Val, C->tile_csr_Col);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int launch_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts) {
    (void)opts;
    if (C->numtile == 0 || C->nnz == 0) return TSG_OK;
    if (A->dtype == TSG_F32) return launch_rows<float>(ctx, A, B, C);
    return launch_rows<double>(ctx, A, B, C);
}

}  // namespace tsg
