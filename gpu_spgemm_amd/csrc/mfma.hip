// mfma.hip -- numeric step with matrix cores for dense tile pairs.
//
// A wave (64 lanes) per C tile, pairs in ascending k as the C-tile kernels
// (src/spgemm_raw_kernels_v2.cuh:10-139 computes the same sums per tile).  For
// a pair whose A and B tiles are dense enough -- nnz(A tile) * nnz(B tile)
// >= 16 * min_products, the expected product count of the pair -- the two
// tiles are expanded to dense 16x16 operands and multiplied by four
// v_mfma_f64_16x16x4 (FP64) or v_mfma_f32_16x16x4 (FP32) steps of four k each;
// sparser pairs add their products with one FMA each on the lanes that own the
// outputs.  Lane l holds A operand element (l%16, kb*4 + l/16), B operand
// element (kb*4 + l/16, l%16) and accumulator j = 0..3 of column l%16 at row
// 4*j + l/16 (FP64) or 4*(l/16) + j (FP32): the matrix cores' layouts.
//
// Summation: per output, pairs ascending and k ascending within a pair, as
// everywhere else; a dense pair also adds the products of its absent entries
// (zeros), which leaves finite sums unchanged except that a -0 sum becomes +0.
// The matrix cores accumulate each k step as a fused multiply-add: on the
// parity cases (tests/test_gpu_mfma.py) the values are bit-identical to the
// one-FMA-per-product kernels; they are checked against the FP64 recompute
// within the stated tolerance.
#include "pairs.h"

namespace tsg {

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

template <typename V>
struct MfAcc;
template <>
struct MfAcc<double> {
    typedef mf_d4 T;
    static __device__ __forceinline__ T mma(double a, double b, T c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
    static __device__ __forceinline__ int row(int quad, int j) { return j * 4 + quad; }
};
template <>
struct MfAcc<float> {
    typedef mf_f4 T;
    static __device__ __forceinline__ T mma(float a, float b, T c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    static __device__ __forceinline__ int row(int quad, int j) { return quad * 4 + j; }
};

__device__ __forceinline__ uint32_t mf_below(int c) { return (0xFFFF0000u >> c) & 0xFFFFu; }

constexpr int MF_WAVES = 4;

template <typename V>
__global__ void __launch_bounds__(MF_WAVES *WAVE)
k_numeric_mfma(int64_t numtileC, int64_t min_prod16, const int64_t *__restrict__ tile_ptrA,
               const int32_t *__restrict__ colidxA, const int64_t *__restrict__ tile_nnzA,
               const uint8_t *__restrict__ ptrA, const uint16_t *__restrict__ maskA, const V *__restrict__ valA,
               const int64_t *__restrict__ csc_ptrB, const int32_t *__restrict__ csc_rowB,
               const int64_t *__restrict__ tile_nnzB, const uint8_t *__restrict__ ptrB,
               const uint16_t *__restrict__ maskB, const V *__restrict__ valB, const int32_t *__restrict__ rowidxC,
               const int32_t *__restrict__ colidxC, const int64_t *__restrict__ tile_nnzC,
               const uint16_t *__restrict__ maskC, const uint8_t *__restrict__ ptrRowC, V *__restrict__ valC,
               uint8_t *__restrict__ colC) {
    typedef typename MfAcc<V>::T Acc;
    // grid-stride over tiles, one wave per tile (the matrix cores need the whole wave)
    for (int64_t ti = (int64_t)blockIdx.x * MF_WAVES + threadIdx.x / WAVE; ti < numtileC;
         ti += (int64_t)gridDim.x * MF_WAVES) {
    const int64_t c0 = tile_nnzC[ti];
    if (tile_nnzC[ti + 1] == c0) continue;  // structurally empty tile
    const int l = lane_id();
    const int col = l & 15, quad = l >> 4;  // output column; operand k within a 4-step, output row block
    const int32_t bi = rowidxC[ti], bj = colidxC[ti];
    Acc acc = {V(0), V(0), V(0), V(0)};
    for_each_pair(tile_ptrA, colidxA, csc_ptrB, csc_rowB, bi, bj, [&](int64_t a, int64_t b) {
        const int64_t an0 = tile_nnzA[a], bn0 = tile_nnzB[b];
        const int64_t na = tile_nnzA[a + 1] - an0, nb = tile_nnzB[b + 1] - bn0;
        if (na * nb >= min_prod16) {  // dense pair: matrix cores
            const uint32_t ma = maskA[a * 16 + col];  // A row `col`
            const int pa = ptrA[a * 16 + col];
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                const int k = kb * 4 + quad;
                const uint32_t kbit = 0x8000u >> k;
                const V av = (ma & kbit) ? valA[an0 + pa + __popc(ma & mf_below(k))] : V(0);
                const uint32_t mb = maskB[b * 16 + k];  // B row k, column `col`
                const uint32_t cbit = 0x8000u >> col;
                const V bv = (mb & cbit) ? valB[bn0 + ptrB[b * 16 + k] + __popc(mb & mf_below(col))] : V(0);
                acc = MfAcc<V>::mma(av, bv, acc);
            }
        } else {  // sparse pair: one FMA per product on the owning lanes
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = MfAcc<V>::row(quad, j);
                uint32_t ma = maskA[a * 16 + r];
                const int64_t ea = an0 + ptrA[a * 16 + r];
                const uint32_t mrow = ma;
                while (ma) {
                    const int k = __clz(ma) - 16;
                    ma ^= 0x8000u >> k;
                    const uint32_t mb = maskB[b * 16 + k];
                    if (mb & (0x8000u >> col)) {
                        const V av = valA[ea + __popc(mrow & mf_below(k))];
                        const V bv = valB[bn0 + ptrB[b * 16 + k] + __popc(mb & mf_below(col))];
                        acc[j] = MfAcc<V>::fma(av, bv, acc[j]);
                    }
                }
            }
        }
    });
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int r = MfAcc<V>::row(quad, j);
        const uint32_t mc = maskC[ti * 16 + r];
        if (mc & (0x8000u >> col)) {
            const int64_t o = c0 + ptrRowC[ti * 16 + r] + __popc(mc & mf_below(col));
            valC[o] = acc[j];
            colC[o] = (uint8_t)((r << 4) | col);
        }
    }
    }
}

int launch_numeric_mfma(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                        const tsg_options *opts) {
    if (C->numtile == 0 || C->nnz == 0) return TSG_OK;
    if ((A->flags | B->flags) & 1) return fail(TSG_ERR_UNSUPPORTED, "the matrix-core kernel needs rows without duplicate columns");
    const int64_t minp = (opts && opts->mfma_min_products > 0) ? opts->mfma_min_products : 256;
    const unsigned grid = grid_capped(C->numtile, MF_WAVES);
    if (A->dtype == TSG_F32)
        k_numeric_mfma<float><<<grid, MF_WAVES * WAVE, 0, ctx->stream>>>(
            C->numtile, minp * 16, A->tile_ptr, A->tile_columnidx, A->tile_nnz, A->tile_csr_Ptr, A->mask,
            (const float *)A->tile_csr_Val, B->csc_tile_ptr, B->csc_tile_rowidx, B->tile_nnz, B->tile_csr_Ptr,
            B->mask, (const float *)B->tile_csr_Val, C->tile_rowidx, C->tile_columnidx, C->tile_nnz, C->mask,
            C->tile_csr_Ptr, (float *)C->tile_csr_Val, C->tile_csr_Col);
    else
        k_numeric_mfma<double><<<grid, MF_WAVES * WAVE, 0, ctx->stream>>>(
            C->numtile, minp * 16, A->tile_ptr, A->tile_columnidx, A->tile_nnz, A->tile_csr_Ptr, A->mask,
            (const double *)A->tile_csr_Val, B->csc_tile_ptr, B->csc_tile_rowidx, B->tile_nnz, B->tile_csr_Ptr,
            B->mask, (const double *)B->tile_csr_Val, C->tile_rowidx, C->tile_columnidx, C->tile_nnz, C->mask,
            C->tile_csr_Ptr, (double *)C->tile_csr_Val, C->tile_csr_Col);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

}  // namespace tsg
