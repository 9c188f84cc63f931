// rownum.hip -- numeric step of light tile-rows, one lane per scalar row.
//
// The reference computes each C tile with a half-warp over its matched tile
// pairs (src/spgemm_raw_kernels_v2.cuh:10-139); k_num_staged gives each lane
// one output and walks that output's pairs.  Here a wave takes four C
// tile-rows and each lane one scalar row of them: the lane walks its row's A
// entries in ascending column k (A's scalar CSR) and, for each, the B row k
// (B's scalar CSR), adding every product into the tile-row's C values staged
// in LDS at the output's position -- the C tile of the column (C tile
// columns ascend like B's), the tile's row offset and the rank of the column
// in the row mask.  An output's products therefore arrive in ascending k,
// one FMA each: the summation order of every other path, so the bits are the
// same.  A lane owns its row's outputs, so the adds need no atomics.  A
// tile-row whose values exceed the LDS window runs in windows of whole C
// tiles, the products walked once per window.
#include <climits>

#include "rowstaged.h"

namespace tsg {

constexpr int RN_GROUPS = 4;   // tile-rows per wave; lane = (tile-row, scalar row)
constexpr int RN_NC = 16;      // C tiles a tile-row may hold (the small staged variants' cap)
constexpr int RN_VCAP = 256;   // C values of one tile-row held in LDS at a time
constexpr int RN_EB = 8;       // A entries whose loads issue together
constexpr int RN_BB = 8;       // B row entries whose loads issue together

template <typename V>
struct RowNumLDS {
    V val[RN_GROUPS][RN_VCAP + 16];  // + a slot per lane that takes the products outside the window
    alignas(16) uint16_t cmask[RN_GROUPS][RN_NC * 16];  // row masks of the tile-row's C tiles
    alignas(16) uint8_t cptr[RN_GROUPS][RN_NC * 16];    // row offsets inside each C tile
    int32_t coff[RN_GROUPS][RN_NC + 1];                 // each C tile's first value, from the tile-row's first
    int32_t cj[RN_GROUPS][RN_NC];                       // C tile columns, ascending
};

// Bits of a 16-bit row mask that stand for columns below c (bit 15-c is column c).
__device__ __forceinline__ uint32_t mask_below(int c) { return (0xFFFF0000u >> c) & 0xFFFFu; }

template <typename V>
__device__ __forceinline__ V fma_rn(V a, V b, V c);
template <>
__device__ __forceinline__ double fma_rn<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_rn<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename V>
__global__ void __launch_bounds__(WAVE)
k_num_rows(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nrows, const int64_t *__restrict__ dlo,
           const int64_t *__restrict__ dhi, int32_t m,
           const int64_t *__restrict__ rpA, const int32_t *__restrict__ colA, const V *__restrict__ valA,
           const int64_t *__restrict__ rpB, const int32_t *__restrict__ colB, const V *__restrict__ valB,
           const int64_t *__restrict__ tile_ptrC, const int32_t *__restrict__ colidxC,
           const int64_t *__restrict__ tile_nnzC, const uint16_t *__restrict__ maskC,
           const uint8_t *__restrict__ ptrC, V *__restrict__ valC, uint8_t *__restrict__ colC,
           int32_t *__restrict__ overflow, int64_t val_cap, int64_t val_done) {
    __shared__ RowNumLDS<V> s;
    const int lane = lane_id();
    const int g = lane >> 4, r = lane & 15;
    const int64_t idx = (int64_t)blockIdx.x * RN_GROUPS + g;
    // this group's tile-row and C tile range (uniform in the group)
    const int32_t bi = bin_row(rows, row_begin, nrows, dlo, dhi, idx);
    int64_t cs = 0, ce = 0;
    if (bi >= 0) {
        cs = tile_ptrC[bi];
        ce = tile_ptrC[bi + 1];
    }
    const int nc = (int)(ce - cs);
    bool ok = bi >= 0 && nc <= RN_NC;
    if (bi >= 0 && nc > RN_NC && r == 0) atomicOr(overflow, 1);
    // C headers, lane r = C tile r of the group
    int64_t vs = 0, vend = 0, co = 0;
    int32_t cjv = 0;
    uint4 m0 = make_uint4(0, 0, 0, 0), m1 = m0, pp = m0;
    if (ok) {
        vs = tile_nnzC[cs];
        vend = tile_nnzC[ce];
        if (r < nc) {
            co = tile_nnzC[cs + r];
            cjv = colidxC[cs + r];
            m0 = reinterpret_cast<const uint4 *>(maskC)[(cs + r) * 2];
            m1 = reinterpret_cast<const uint4 *>(maskC)[(cs + r) * 2 + 1];
            pp = reinterpret_cast<const uint4 *>(ptrC)[cs + r];
        }
        if (vend > val_cap) {  // C's values outgrew their capacity (learned scheduler)
            if (r == 0) atomicOr(overflow, 1);
            ok = false;
        } else if (vend <= val_done) {  // a re-run: this tile-row's values are already written
            ok = false;
        }
    }
    if (ok && r < nc) {
        s.coff[g][r] = (int32_t)(co - vs);
        s.cj[g][r] = cjv;
        reinterpret_cast<uint4 *>(&s.cmask[g][r * 16])[0] = m0;
        reinterpret_cast<uint4 *>(&s.cmask[g][r * 16])[1] = m1;
        reinterpret_cast<uint4 *>(&s.cptr[g][r * 16])[0] = pp;
    }
    if (ok && r == 0) s.coff[g][nc] = (int32_t)(vend - vs);
    // this lane's scalar row of A
    const int64_t gr = (int64_t)(bi < 0 ? 0 : bi) * 16 + r;
    int64_t e0 = 0, e1 = 0;
    if (ok && gr < m) {
        e0 = rpA[gr];
        e1 = rpA[gr + 1];
    }
    wave_sync();
    // windows of whole C tiles holding at most RN_VCAP values (a tile holds <= 256)
    int w0 = 0;
    while (true) {
        const bool act = ok && w0 < nc;
        if (!__any(act)) break;
        int w1 = w0, vb = 0, vn = 0;
        if (act) {
            vb = s.coff[g][w0];
            w1 = w0 + 1;
            while (w1 < nc && s.coff[g][w1 + 1] - vb <= RN_VCAP) w1++;
            vn = s.coff[g][w1] - vb;
        }
        for (int i = r; i < vn; i += 16) s.val[g][i] = V(0);
        wave_sync();
        if (act) {
            const int32_t jlo = s.cj[g][w0], jhi = s.cj[g][w1 - 1];
            // A entries RN_EB at a time, their B rows RN_BB entries at a time:
            // each batch's loads are issued together (one latency a batch)
            for (int64_t eb = e0; eb < e1; eb += RN_EB) {
                const int ne = (int)min<int64_t>(RN_EB, e1 - eb);
                int32_t kk[RN_EB];
                V aa[RN_EB];
                int64_t j0[RN_EB], j1[RN_EB];
                // loads past the row's end repeat its last entry (in bounds,
                // no branch) and are then given an empty B row
#pragma unroll
                for (int u = 0; u < RN_EB; u++) {
                    const int64_t e = eb + min(u, ne - 1);
                    kk[u] = colA[e];
                    aa[u] = valA[e];
                }
#pragma unroll
                for (int u = 0; u < RN_EB; u++) {
                    j0[u] = rpB[kk[u]];
                    j1[u] = rpB[kk[u] + 1];
                }
#pragma unroll
                for (int u = 0; u < RN_EB; u++)
                    if (u >= ne) j1[u] = j0[u];
#pragma unroll
                for (int u = 0; u < RN_EB; u++) {
                    // the current C tile's column, first value of this row in it and row
                    // mask, kept in registers while the B row stays in that tile
                    int ct = w0;
                    int32_t jc = -1, base = 0;
                    uint32_t mk = 0;
                    for (int64_t jb = j0[u]; jb < j1[u]; jb += RN_BB) {
                        const int nb = (int)min<int64_t>(RN_BB, j1[u] - jb);
                        int32_t cc[RN_BB];
                        V bv[RN_BB];
#pragma unroll
                        for (int q = 0; q < RN_BB; q++) {  // past the row's end: its last entry again, column INT_MAX
                            const int64_t j = jb + min(q, nb - 1);
                            cc[q] = colB[j];
                            bv[q] = valB[j];
                            if (q >= nb) cc[q] = INT_MAX;
                        }
#pragma unroll
                        for (int q = 0; q < RN_BB; q++) {
                            const int32_t J = cc[q] >> 4;
                            const bool in = J >= jlo && J <= jhi;  // inside the window
                            if (in && J != jc) {
                                while (s.cj[g][ct] < J) ct++;  // J is one of the tile-row's C tile columns
                                const int t = ct * 16 + r;
                                jc = J;
                                base = s.coff[g][ct] - vb + s.cptr[g][t];
                                mk = s.cmask[g][t];
                            }
                            // outside the window: the lane's own spare slot, never read
                            const int pos = in ? base + __popc(mk & mask_below(cc[q] & 15)) : RN_VCAP + r;
                            s.val[g][pos] = fma_rn<V>(aa[u], bv[q], s.val[g][pos]);
                        }
                    }
                }
            }
        }
        wave_sync();
        // the window's values and col bytes: lane r writes its row's run in each tile
        if (act) {
            for (int ct = w0; ct < w1; ct++) {
                const int q = ct * 16 + r;
                uint32_t mr = s.cmask[g][q];
                int o = s.coff[g][ct] + s.cptr[g][q];
                while (mr) {
                    const int c = __clz(mr) - 16;
                    mr ^= 0x8000u >> c;
                    valC[vs + o] = s.val[g][o - vb];
                    colC[vs + o] = (uint8_t)((r << 4) | c);
                    o++;
                }
            }
        }
        wave_sync();
        w0 = w1;
    }
}

bool num_rows_ok(const tsg_tile_matrix *A, const tsg_tile_matrix *B) {
    return A->row_ptr && A->row_col && A->row_val && B->row_ptr && B->row_col && B->row_val &&
           !((A->flags | B->flags) & 1);
}

int launch_num_rows(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                    const int32_t *rows, const BinRange &br, const SlotLayout &sl) {
    const int64_t count = br.count;
    if (count <= 0) return TSG_OK;
    if (sl.row_nc) return fail(TSG_ERR_ARG, "the row-lane numeric kernel runs on the exact C layout only");
    const unsigned grid = (unsigned)((count + RN_GROUPS - 1) / RN_GROUPS);
    if (A->dtype == TSG_F32)
        k_num_rows<float><<<grid, WAVE, 0, ctx->stream>>>(
            rows, br.begin, count, br.lo, br.hi, A->m, A->row_ptr, A->row_col, (const float *)A->row_val, B->row_ptr, B->row_col,
            (const float *)B->row_val, C->tile_ptr, C->tile_columnidx, C->tile_nnz, C->mask, C->tile_csr_Ptr,
            (float *)C->tile_csr_Val, C->tile_csr_Col, sl.overflow, sl.val_cap, sl.val_done);
    else
        k_num_rows<double><<<grid, WAVE, 0, ctx->stream>>>(
            rows, br.begin, count, br.lo, br.hi, A->m, A->row_ptr, A->row_col, (const double *)A->row_val, B->row_ptr, B->row_col,
            (const double *)B->row_val, C->tile_ptr, C->tile_columnidx, C->tile_nnz, C->mask, C->tile_csr_Ptr,
            (double *)C->tile_csr_Val, C->tile_csr_Col, sl.overflow, sl.val_cap, sl.val_done);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

}  // namespace tsg
