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
//   k_rg_symbolic: C's tile columns of the row are already written (Step 1
//     fill); each product sets bit c = j % 16 of row mask (i % 16) in C tile
//     j / 16 (atomic OR on the row's own tiles, zeroed first), then the row's
//     tiles get their row offsets and value counts.  The masks equal the
//     reference's OR over tile pairs: a column is present exactly when some
//     product reaches it.
//   k_rg_numeric: the row's values are zeroed and their col bytes written;
//     then thread (r, s) takes scalar row r and slice s of the row's C tiles
//     (16 slices of consecutive tiles, hence of consecutive columns): it walks
//     the row's A entries in ascending k, finds the slice's column range in
//     B row k by binary search, and accumulates each product there with one
//     FMA into the output's slot.  Every output belongs to one thread and sums
//     its products in ascending k -- the order of the tiled method (tile pairs
//     ascending, entries ascending within a pair) -- so values are
//     bit-identical to the other kernels.
//
// A C tile column is mapped to its position in the row through an LDS table
// over the row's tile-column window (RG_MAP entries), or by binary search in
// the row's sorted tile columns when the window is wider.
#include <algorithm>

#include "tsg_internal.h"

namespace tsg {

constexpr int RG_THREADS = 256;   // 16 threads per scalar row
constexpr int RG_MAP = 8192;      // tile columns mapped through LDS (16 KB)

__device__ __forceinline__ uint32_t rg_below(int c) { return (0xFFFF0000u >> c) & 0xFFFFu; }

template <typename V>
__device__ __forceinline__ V rg_fma(V a, V b, V c);
template <>
__device__ __forceinline__ double rg_fma<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float rg_fma<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Position of C tile column bj in [cs, ce): LDS map when the window fits,
// else binary search in the sorted global list.  In the slot layout a row
// whose tiles overflowed their slots holds only some of its columns (its
// result is discarded); `verify` then returns -1 for a column not found.
__device__ __forceinline__ int64_t rg_tile(int32_t bj, bool use_map, const uint16_t *map, int32_t cmin,
                                           int32_t cmax, const int32_t *__restrict__ colidxC, int64_t cs,
                                           int64_t ce, bool verify) {
    int64_t t;
    if (use_map) {
        if (verify && (bj < cmin || bj > cmax)) return -1;
        t = cs + map[bj - cmin];
    } else {
        int64_t lo = cs, hi = ce - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (colidxC[mid] < bj) lo = mid + 1; else hi = mid;
        }
        t = lo;
    }
    if (verify && (t < cs || t >= ce || colidxC[t] != bj)) return -1;
    return t;
}

// Builds the map (every thread of the block calls it; map entries of absent
// columns stay unset and are only read under `verify`).
__device__ __forceinline__ bool rg_build_map(uint16_t *map, const int32_t *__restrict__ colidxC, int64_t cs,
                                             int64_t ce, int32_t &cmin, int32_t &cmax) {
    const int64_t nc = ce - cs;
    cmin = colidxC[cs];
    cmax = colidxC[ce - 1];
    const bool use = (int64_t)cmax - cmin < RG_MAP && nc <= 65535;
    if (use)
        for (int64_t q = threadIdx.x; q < nc; q += blockDim.x) map[colidxC[cs + q] - cmin] = (uint16_t)q;
    return use;
}

__global__ void __launch_bounds__(RG_THREADS)
k_rg_symbolic(const int32_t *__restrict__ rows, int64_t row_begin, int32_t m,
              const int64_t *__restrict__ rpA, const int32_t *__restrict__ colA, const int64_t *__restrict__ rpB,
              const int32_t *__restrict__ colB, const int64_t *__restrict__ tile_ptrC,
              const int32_t *__restrict__ row_nc, const int32_t *__restrict__ colidxC, uint16_t *__restrict__ maskC,
              uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC) {
    __shared__ uint16_t map[RG_MAP];
    const int32_t bi = rows[row_begin + blockIdx.x];
    int64_t cs, ce;
    c_row_range(tile_ptrC, row_nc, bi, cs, ce);
    if (ce <= cs) return;  // whole block leaves: no C tiles, no products
    uint4 *mw = reinterpret_cast<uint4 *>(maskC);
    for (int64_t q = threadIdx.x; q < (ce - cs) * 2; q += RG_THREADS) mw[cs * 2 + q] = make_uint4(0, 0, 0, 0);
    int32_t cmin = 0, cmax = 0;
    const bool use_map = rg_build_map(map, colidxC, cs, ce, cmin, cmax);
    const bool verify = row_nc != nullptr;
    __threadfence();
    __syncthreads();
    // products: 16 threads per scalar row, each a strided share of the row's A entries
    const int r = threadIdx.x >> 4, q = threadIdx.x & 15;
    const int32_t i = bi * 16 + r;
    uint32_t *mw32 = reinterpret_cast<uint32_t *>(maskC);
    if (i < m) {
        for (int64_t e = rpA[i] + q; e < rpA[i + 1]; e += 16) {
            const int32_t k = colA[e];
            for (int64_t f = rpB[k]; f < rpB[k + 1]; f++) {
                const int32_t j = colB[f];
                const int64_t t = rg_tile(j >> 4, use_map, map, cmin, cmax, colidxC, cs, ce, verify);
                if (t >= 0) atomicOr(&mw32[t * 8 + (r >> 1)], (0x8000u >> (j & 15)) << ((r & 1) * 16));
            }
        }
    }
    __threadfence();
    __syncthreads();
    // row offsets and value counts of every tile of the row (16 threads per tile)
    for (int64_t it = threadIdx.x; it < (ce - cs) * 16; it += RG_THREADS) {
        const int64_t t = cs + (it >> 4);
        const int rr = (int)(it & 15);
        const uint32_t w = __hip_atomic_load(&mw32[t * 8 + (rr >> 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t rm = (w >> ((rr & 1) * 16)) & 0xFFFFu;
        const int pc = __popc(rm);
        const int inc = group16_incl_sum(pc);
        ptrRowC[t * 16 + rr] = (uint8_t)(inc - pc);
        if (rr == 15) tile_nnzC[t + 1] = inc;
    }
}

template <typename V>
__global__ void __launch_bounds__(RG_THREADS)
k_rg_numeric(const int32_t *__restrict__ rows, int64_t row_begin, int32_t m, const int64_t *__restrict__ rpA,
             const int32_t *__restrict__ colA, const V *__restrict__ valA, const int64_t *__restrict__ rpB,
             const int32_t *__restrict__ colB, const V *__restrict__ valB, const int64_t *__restrict__ tile_ptrC,
             const int32_t *__restrict__ row_nc, const int32_t *__restrict__ colidxC,
             const uint16_t *__restrict__ maskC, const uint8_t *__restrict__ ptrRowC,
             const int64_t *__restrict__ tile_nnzC, V *__restrict__ valC, uint8_t *__restrict__ colC,
             int32_t *__restrict__ overflow, int64_t val_cap) {
    __shared__ uint16_t map[RG_MAP];
    const int32_t bi = rows[row_begin + blockIdx.x];
    int64_t cs, ce;
    c_row_range(tile_ptrC, row_nc, bi, cs, ce);
    if (ce <= cs) return;
    if (tile_nnzC[ce] > val_cap) {  // slot layout: C's values outgrew their capacity
        if (threadIdx.x == 0) atomicOr(overflow, 1);
        return;
    }
    // zero the row's values and write their col bytes (16 threads per tile)
    for (int64_t it = threadIdx.x; it < (ce - cs) * 16; it += RG_THREADS) {
        const int64_t t = cs + (it >> 4);
        const int rr = (int)(it & 15);
        uint32_t rm = maskC[t * 16 + rr];
        int64_t o = tile_nnzC[t] + ptrRowC[t * 16 + rr];
        while (rm) {
            const int c = __clz(rm) - 16;
            rm ^= 0x8000u >> c;
            valC[o] = V(0);
            colC[o] = (uint8_t)((rr << 4) | c);
            o++;
        }
    }
    int32_t cmin = 0, cmax = 0;
    const bool use_map = rg_build_map(map, colidxC, cs, ce, cmin, cmax);
    const bool verify = row_nc != nullptr;
    __threadfence();
    __syncthreads();
    // thread (r, s): scalar row r, C tiles [t_lo, t_hi) of the row; products in ascending k
    const int r = threadIdx.x >> 4, sl = threadIdx.x & 15;
    const int32_t i = bi * 16 + r;
    const int64_t nc = ce - cs;
    const int64_t t_lo = cs + nc * sl / 16, t_hi = cs + nc * (sl + 1) / 16;
    if (i < m && t_lo < t_hi) {
        const int32_t col_lo = colidxC[t_lo] * 16, col_hi = colidxC[t_hi - 1] * 16 + 15;
        {
            for (int64_t e = rpA[i]; e < rpA[i + 1]; e++) {
                const int32_t k = colA[e];
                const V a = valA[e];
                int64_t lo = rpB[k], hi = rpB[k + 1];
                const int64_t fe = hi;
                while (lo < hi) {  // first entry of B row k at or after col_lo
                    const int64_t mid = (lo + hi) >> 1;
                    if (colB[mid] < col_lo) lo = mid + 1; else hi = mid;
                }
                for (int64_t f = lo; f < fe; f++) {
                    const int32_t j = colB[f];
                    if (j > col_hi) break;
                    const int c = j & 15;
                    const int64_t t = rg_tile(j >> 4, use_map, map, cmin, cmax, colidxC, cs, ce, verify);
                    if (t < 0) continue;
                    const int64_t o = tile_nnzC[t] + ptrRowC[t * 16 + r] + __popc(maskC[t * 16 + r] & rg_below(c));
                    valC[o] = rg_fma<V>(a, valB[f], valC[o]);
                }
            }
        }
    }
}

int launch_rg_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                       const int32_t *rows, int64_t begin, int64_t count, const SlotLayout &sl) {
    for (int64_t c0 = 0; c0 < count; c0 += ROW_CHUNK) {  // work-items per launch stay below 2^32
        const unsigned n = (unsigned)std::min<int64_t>(ROW_CHUNK, count - c0);
        k_rg_symbolic<<<n, RG_THREADS, 0, ctx->stream>>>(
            rows, begin + c0, A->m, A->row_ptr, A->row_col, B->row_ptr, B->row_col, C->tile_ptr, sl.row_nc,
            C->tile_columnidx, C->mask, C->tile_csr_Ptr, C->tile_nnz);
        TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

int launch_rg_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                      const int32_t *rows, int64_t begin, int64_t count, const SlotLayout &sl) {
    for (int64_t c0 = 0; c0 < count; c0 += ROW_CHUNK) {  // work-items per launch stay below 2^32
    const unsigned n = (unsigned)std::min<int64_t>(ROW_CHUNK, count - c0);
    if (A->dtype == TSG_F32)
        k_rg_numeric<float><<<n, RG_THREADS, 0, ctx->stream>>>(
            rows, begin + c0, A->m, A->row_ptr, A->row_col, (const float *)A->row_val, B->row_ptr, B->row_col,
            (const float *)B->row_val, C->tile_ptr, sl.row_nc, C->tile_columnidx, C->mask, C->tile_csr_Ptr,
            C->tile_nnz, (float *)C->tile_csr_Val, C->tile_csr_Col, sl.overflow, sl.val_cap);
    else
        k_rg_numeric<double><<<n, RG_THREADS, 0, ctx->stream>>>(
            rows, begin + c0, A->m, A->row_ptr, A->row_col, (const double *)A->row_val, B->row_ptr, B->row_col,
            (const double *)B->row_val, C->tile_ptr, sl.row_nc, C->tile_columnidx, C->mask, C->tile_csr_Ptr,
            C->tile_nnz, (double *)C->tile_csr_Val, C->tile_csr_Col, sl.overflow, sl.val_cap);
    TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

}  // namespace tsg
