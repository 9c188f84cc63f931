// numeric.hip -- Step 3: values of C, one C tile per 16-lane group.
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
//
// k_numeric_outer is the outer-product form of the same step over A's
// within-tile CSC (numeric_variant 3), with identical results.
//
// This is the C-tile-centric method of the reference.  It runs every tile in
// the reference-method variant and the heavy tile-rows (one workgroup per row)
// in the binned pipeline; light rows go to the row-staged kernels.
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
struct NumArgs {
    const int64_t *tile_ptrA; const int32_t *colidxA; const int64_t *tile_nnzA; const uint8_t *ptrRowA;
    const uint8_t *colA; const V *valA;
    const int64_t *csc_ptrB; const int32_t *csc_rowB; const int64_t *tile_nnzB; const uint8_t *ptrRowB;
    const uint16_t *maskB; const uint8_t *colB; const V *valB;
    const int64_t *tile_ptrC; const int32_t *rowidxC; const int32_t *colidxC; const int64_t *tile_nnzC;
    const uint16_t *maskC; const uint8_t *ptrRowC; V *valC; uint8_t *colC;
    const int32_t *row_nc; int32_t *overflow; int64_t val_cap;  // slot layout / value capacity (learned scheduler)
    int64_t val_done;  // a numeric re-run: values below this are already written
};

template <typename V>
static NumArgs<V> num_args(const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C) {
    return NumArgs<V>{A->tile_ptr, A->tile_columnidx, A->tile_nnz, A->tile_csr_Ptr, A->tile_csr_Col,
                      (const V *)A->tile_csr_Val, B->csc_tile_ptr, B->csc_tile_rowidx, B->tile_nnz,
                      B->tile_csr_Ptr, B->mask, B->tile_csr_Col, (const V *)B->tile_csr_Val, C->tile_ptr, C->tile_rowidx,
                      C->tile_columnidx, C->tile_nnz, C->mask, C->tile_csr_Ptr, (V *)C->tile_csr_Val,
                      C->tile_csr_Col, nullptr, nullptr, INT64_MAX, -1};
}

// Numeric step of one C tile by a 16-lane group (lane = tile row).  DUPS:
// some row repeats a column, so B row entries are walked by their col bytes
// in storage order (as the reference does) instead of by mask bits.
template <typename V, bool DUPS>
__device__ __forceinline__ void num_tile(int64_t ti, const NumArgs<V> &p) {
    const int l = lane_id() & 15;
    const int64_t c0 = p.tile_nnzC[ti];
    if (p.tile_nnzC[ti + 1] == c0) return;  // structurally empty tile
    if (p.tile_nnzC[ti + 1] > p.val_cap) {   // C's values outgrew their capacity
        if ((lane_id() & 15) == 0) atomicOr(p.overflow, 1);
        return;
    }
    if (p.tile_nnzC[ti + 1] <= p.val_done) return;  // a re-run: already written
    const int32_t bi = p.rowidxC[ti], bj = p.colidxC[ti];
    V acc[16];
#pragma unroll
    for (int c = 0; c < 16; c++) acc[c] = V(0);
    for_each_pair(p.tile_ptrA, p.colidxA, p.csc_ptrB, p.csc_rowB, bi, bj, [&](int64_t a, int64_t b) {
        const int64_t an0 = p.tile_nnzA[a];
        const int e0 = p.ptrRowA[a * 16 + l];
        const int e1 = l == 15 ? (int)(p.tile_nnzA[a + 1] - an0) : (int)p.ptrRowA[a * 16 + l + 1];
        const int64_t bn0 = p.tile_nnzB[b];
        for (int e = e0; e < e1; e++) {
            const int kk = p.colA[an0 + e] & 15;
            const V va = p.valA[an0 + e];
            int64_t pb = bn0 + p.ptrRowB[b * 16 + kk];
            if (DUPS) {
                const int64_t pe = bn0 + (kk == 15 ? (int)(p.tile_nnzB[b + 1] - bn0) : (int)p.ptrRowB[b * 16 + kk + 1]);
                for (; pb < pe; pb++) {
                    const int cb = p.colB[pb] & 15;
                    const V vb = p.valB[pb];
#pragma unroll
                    for (int c = 0; c < 16; c++)
                        if (c == cb) acc[c] = fmav<V>(va, vb, acc[c]);
                }
                continue;
            }
            const uint32_t mb = p.maskB[b * 16 + kk];
#pragma unroll
            for (int c = 0; c < 16; c++) {
                if (mb & (0x8000u >> c)) {
                    acc[c] = fmav<V>(va, p.valB[pb], acc[c]);
                    pb++;
                }
            }
        }
    });
    const uint32_t mc = p.maskC[ti * 16 + l];
    int64_t o = c0 + p.ptrRowC[ti * 16 + l];
#pragma unroll
    for (int c = 0; c < 16; c++) {
        if (mc & (0x8000u >> c)) {
            p.valC[o] = acc[c];
            p.colC[o] = (uint8_t)((l << 4) | c);
            o++;
        }
    }
}

// Outer-product form of the same step (numeric_variant 3), over A's
// within-tile CSC (tile_csc_Ptr/Row/Val, the reference's tile_a_csc_* of
// src/tile_format.h:39-41 built by build_a_tile_csc :589-627): a 16-lane
// group per C tile, lane = tile column c, 16 register accumulators indexed by
// row.  For every pair (ascending k tile) and every local k where A's column
// k and B's row k both hold entries (from the tiles' masks), A's column k
// (its rows, read from the CSC arrays) times B's row k adds one product to
// each (r, c) it meets.  Output (r, c) therefore sees its products in the
// order the row form gives them (pairs ascending, k ascending, repeated
// entries in storage order), so the two forms produce identical bits.
// A tile's occupied columns (OR of its 16 row masks, bit 15-c) and its
// non-empty rows (bit 15-r), from two 16-byte loads of the masks.
__device__ __forceinline__ void tile_cols_rows(const uint16_t *__restrict__ m, int64_t t, uint32_t &cols,
                                               uint32_t &rows) {
    const uint4 lo = reinterpret_cast<const uint4 *>(m)[t * 2], hi = reinterpret_cast<const uint4 *>(m)[t * 2 + 1];
    const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};  // w[i] = rows 2i (low), 2i+1
    cols = 0;
    rows = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        cols |= (w[i] | (w[i] >> 16)) & 0xFFFFu;
        rows |= ((w[i] & 0xFFFFu) ? 0x8000u : 0u) >> (2 * i);
        rows |= ((w[i] >> 16) ? 0x8000u : 0u) >> (2 * i + 1);
    }
}

template <typename V, bool DUPS>
__device__ __forceinline__ void num_tile_outer(int64_t ti, const NumArgs<V> &p, const uint16_t *__restrict__ amask,
                                               const uint8_t *__restrict__ acp, const uint8_t *__restrict__ acr,
                                               const V *__restrict__ acv) {
    const int c = lane_id() & 15;
    const uint32_t bit = 0x8000u >> c, left = (0xFFFF0000u >> c) & 0xFFFFu;
    const int64_t c0 = p.tile_nnzC[ti];
    if (p.tile_nnzC[ti + 1] == c0) return;  // structurally empty tile
    const int32_t bi = p.rowidxC[ti], bj = p.colidxC[ti];
    V acc[16];
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = V(0);
    for_each_pair(p.tile_ptrA, p.colidxA, p.csc_ptrB, p.csc_rowB, bi, bj, [&](int64_t a, int64_t b) {
        const int64_t an0 = p.tile_nnzA[a];
        const int an = (int)(p.tile_nnzA[a + 1] - an0);
        const int64_t bn0 = p.tile_nnzB[b];
        const int bn = (int)(p.tile_nnzB[b + 1] - bn0);
        uint32_t colsA, rowsA, colsB, rowsB;
        tile_cols_rows(amask, a, colsA, rowsA);
        tile_cols_rows(p.maskB, b, colsB, rowsB);
        uint32_t ks = colsA & rowsB;  // k with entries in A's column k and B's row k
        while (ks) {
            const int k = __clz(ks) - 16;  // ascending k
            ks ^= 0x8000u >> k;
            const int e0 = acp[a * 16 + k];
            const int e1 = k == 15 ? an : (int)acp[a * 16 + k + 1];
            if (DUPS) {
                const int q0 = p.ptrRowB[b * 16 + k], q1 = k == 15 ? bn : (int)p.ptrRowB[b * 16 + k + 1];
                for (int e = e0; e < e1; e++) {
                    const int r = acr[an0 + e];
                    const V va = acv[an0 + e];
                    for (int q = q0; q < q1; q++) {
                        if ((p.colB[bn0 + q] & 15) != c) continue;
                        const V vb = p.valB[bn0 + q];
#pragma unroll
                        for (int rr = 0; rr < 16; rr++)
                            if (rr == r) acc[rr] = fmav<V>(va, vb, acc[rr]);
                    }
                }
                continue;
            }
            const uint32_t mb = p.maskB[b * 16 + k];
            if (!(mb & bit)) continue;  // B[k, c] absent
            const V vb = p.valB[bn0 + p.ptrRowB[b * 16 + k] + __popc(mb & left)];
            uint32_t ca = 0;  // rows of A's column k
            for (int e = e0; e < e1; e++) ca |= 0x8000u >> acr[an0 + e];
            int e = e0;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if (ca & (0x8000u >> r)) {
                    acc[r] = fmav<V>(acv[an0 + e], vb, acc[r]);
                    e++;
                }
            }
        }
    });
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t mc = p.maskC[ti * 16 + r];
        if (mc & bit) {
            const int64_t o = c0 + p.ptrRowC[ti * 16 + r] + __popc(mc & left);
            p.valC[o] = acc[r];
            p.colC[o] = (uint8_t)((r << 4) | c);
        }
    }
}

template <typename V, bool DUPS>
__global__ void __launch_bounds__(NU_THREADS)
k_numeric_outer(int64_t numtileC, NumArgs<V> p, const uint16_t *__restrict__ amask, const uint8_t *__restrict__ acp,
                const uint8_t *__restrict__ acr, const V *__restrict__ acv) {
    for (int64_t ti = (int64_t)blockIdx.x * NU_GROUPS + (threadIdx.x >> 4); ti < numtileC;
         ti += (int64_t)gridDim.x * NU_GROUPS)  // grid-stride over tiles
        num_tile_outer<V, DUPS>(ti, p, amask, acp, acr, acv);
}

template <typename V, bool DUPS>
__global__ void __launch_bounds__(NU_THREADS) k_numeric_tiles(int64_t numtileC, NumArgs<V> p) {
    for (int64_t ti = (int64_t)blockIdx.x * NU_GROUPS + (threadIdx.x >> 4); ti < numtileC;
         ti += (int64_t)gridDim.x * NU_GROUPS)  // grid-stride over tiles
        num_tile<V, DUPS>(ti, p);
}

// Heavy tile-rows: one workgroup per listed row, its 16 groups stride over the
// row's C tiles.
template <typename V, bool DUPS>
__global__ void __launch_bounds__(NU_THREADS) k_numeric_list(const int32_t *__restrict__ rows, int64_t row_begin,
                                                             int64_t nrows, const int64_t *__restrict__ dlo,
                                                             const int64_t *__restrict__ dhi, int64_t c0,
                                                             NumArgs<V> p) {
    const int32_t bi = bin_row(rows, row_begin, nrows, dlo, dhi, c0 + blockIdx.x);
    if (bi < 0) return;
    int64_t cs, ce;
    c_row_range(p.tile_ptrC, p.row_nc, bi, cs, ce);
    for (int64_t ti = cs + (threadIdx.x >> 4); ti < ce; ti += NU_GROUPS) num_tile<V, DUPS>(ti, p);
}

template <typename V>
static int launch_list_t(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                         const int32_t *rows, const BinRange &br, const SlotLayout &sl) {
    const int64_t count = br.count;
    NumArgs<V> a = num_args<V>(A, B, C);
    a.row_nc = sl.row_nc;
    a.overflow = sl.overflow;
    a.val_cap = sl.val_cap;
    a.val_done = sl.val_done;
    for (int64_t c0 = 0; c0 < count; c0 += ROW_CHUNK) {  // work-items per launch stay below 2^32
        const unsigned n = (unsigned)(count - c0 < ROW_CHUNK ? count - c0 : ROW_CHUNK);
        if ((A->flags | B->flags) & 1)
            k_numeric_list<V, true><<<n, NU_THREADS, 0, ctx->stream>>>(rows, br.begin, count, br.lo, br.hi, c0, a);
        else
            k_numeric_list<V, false><<<n, NU_THREADS, 0, ctx->stream>>>(rows, br.begin, count, br.lo, br.hi, c0, a);
        TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

template <typename V>
static int launch_tiles_t(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                          unsigned grid, bool outer) {
    if (outer) {
        const V *acv = (const V *)A->tile_csc_Val;
        if ((A->flags | B->flags) & 1)
            k_numeric_outer<V, true><<<grid, NU_THREADS, 0, ctx->stream>>>(
                C->numtile, num_args<V>(A, B, C), A->mask, A->tile_csc_Ptr, A->tile_csc_Row, acv);
        else
            k_numeric_outer<V, false><<<grid, NU_THREADS, 0, ctx->stream>>>(
                C->numtile, num_args<V>(A, B, C), A->mask, A->tile_csc_Ptr, A->tile_csc_Row, acv);
        TSG_HIP(hipGetLastError());
        return TSG_OK;
    }
    if ((A->flags | B->flags) & 1)
        k_numeric_tiles<V, true><<<grid, NU_THREADS, 0, ctx->stream>>>(C->numtile, num_args<V>(A, B, C));
    else
        k_numeric_tiles<V, false><<<grid, NU_THREADS, 0, ctx->stream>>>(C->numtile, num_args<V>(A, B, C));
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int launch_numeric_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                        const int32_t *rows, const BinRange &br, const SlotLayout &sl) {
    if (br.count <= 0) return TSG_OK;
    if (A->dtype == TSG_F32) return launch_list_t<float>(ctx, A, B, C, rows, br, sl);
    return launch_list_t<double>(ctx, A, B, C, rows, br, sl);
}

int launch_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts) {
    const bool outer = opts && opts->numeric_variant == 3;
    if (outer && !(A->tile_csc_Ptr && A->tile_csc_Row && A->tile_csc_Val))
        return fail(TSG_ERR_ARG, "numeric_variant 3 needs A's within-tile CSC (tsg_build_tile_csc)");
    if (C->numtile == 0 || C->nnz == 0) return TSG_OK;
    const unsigned grid = grid_capped(C->numtile, NU_GROUPS);
    if (A->dtype == TSG_F32) return launch_tiles_t<float>(ctx, A, B, C, grid, outer);
    return launch_tiles_t<double>(ctx, A, B, C, grid, outer);
}

}  // namespace tsg
