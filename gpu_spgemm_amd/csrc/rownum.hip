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
#include <type_traits>

#include "rowstaged.h"

namespace tsg {

constexpr int RN_GROUPS = 4;   // tile-rows per wave; lane = (tile-row, scalar row)
constexpr int RN_NC = 16;      // C tiles a tile-row may hold (the small staged variants' cap)
constexpr int RN_VCAP = 208;   // C values of one tile-row held in LDS at a time (a 2-D 13-point row block: 208)
constexpr int RN_EB = 5;       // A entries whose loads issue together (a 2-D 5-point row: 5)
constexpr int RN_BB = 5;       // B row entries whose loads issue together (a 2-D 5-point row: 5)

template <typename V>
struct RowNumLDS {
    V val[RN_GROUPS][RN_VCAP + 1];  // + one slot for the products outside the window (never read)
    alignas(16) uint16_t cmask[RN_GROUPS][RN_NC * 16];  // row masks of the tile-row's C tiles
    alignas(16) uint8_t cptr[RN_GROUPS][RN_NC * 16];    // row offsets inside each C tile
    int32_t coff[RN_GROUPS][RN_NC + 1];                 // each C tile's first value, from the tile-row's first
    int32_t cj[RN_GROUPS][RN_NC];                       // C tile columns, ascending
    uint8_t colb[RN_GROUPS][RN_VCAP + 1];               // col byte of each value
};

// Bits of a 16-bit row mask that stand for columns below c (bit 15-c is column c).
__device__ __forceinline__ uint32_t mask_below(int c) { return (0xFFFF0000u >> c) & 0xFFFFu; }

template <typename V>
__device__ __forceinline__ V fma_rn(V a, V b, V c);
template <>
__device__ __forceinline__ double fma_rn<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_rn<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// A load at a 32-bit element index from a kernel-argument base: the address
// is the base plus a 32-bit byte offset (no 64-bit address arithmetic).
// Valid while the array holds less than 4 GB.
template <typename T>
__device__ __forceinline__ T ld_narrow(const T *__restrict__ base, uint32_t i) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T ld_idx(const T *__restrict__ base, uint32_t i) { return ld_narrow(base, i); }
template <typename T>
__device__ __forceinline__ T ld_idx(const T *__restrict__ base, int64_t i) { return base[i]; }

// NQ consecutive entries from element j of a 16-byte aligned array, read as
// whole aligned 16-byte words: only words holding one of the first ne (>= 1)
// entries are read (the others repeat the first word's address), so every
// read lies in a word the array occupies.  Entries past ne are the words'
// other contents and must not be used.
template <typename T, int NQ, typename IX>
__device__ __forceinline__ void ld_run(const T *__restrict__ base, IX j, int ne, T (&out)[NQ]) {
    constexpr int PER = 16 / (int)sizeof(T);
    constexpr int NCH = (PER - 1 + NQ + PER - 1) / PER;
    const int sh = (int)(j & (IX)(PER - 1));
    const IX a0 = j - (IX)sh;
    T w[NCH * PER];
#pragma unroll
    for (int k = 0; k < NCH; k++) {
        const IX ak = PER * k < sh + ne ? a0 + (IX)(PER * k) : a0;
        const uint4 u = ld_idx(reinterpret_cast<const uint4 *>(base), (IX)(ak / (IX)PER));
        __builtin_memcpy(&w[k * PER], &u, 16);
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        T v = w[q];
#pragma unroll
        for (int t = 1; t < PER; t++)
            if (sh == t) v = w[q + t];
        out[q] = v;
    }
}

// NARROW: every index into A's and B's scalar CSR arrays fits a 32-bit byte
// offset (nnz and rows below 2^29), so positions are 32-bit.
template <typename V, bool NARROW>
__global__ void __launch_bounds__(WAVE)
k_num_rows(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nrows, const int64_t *__restrict__ dlo,
           const int64_t *__restrict__ dhi, int32_t m,
           const int64_t *__restrict__ rpA, const int32_t *__restrict__ colA, const V *__restrict__ valA,
           const int64_t *__restrict__ rpB, const int32_t *__restrict__ colB, const V *__restrict__ valB,
           const int64_t *__restrict__ tile_ptrC, const int32_t *__restrict__ colidxC,
           const int64_t *__restrict__ tile_nnzC, const uint16_t *__restrict__ maskC,
           const uint8_t *__restrict__ ptrC, V *__restrict__ valC, uint8_t *__restrict__ colC,
           int32_t *__restrict__ overflow, int64_t val_cap, int64_t val_done) {
    using IX = typename std::conditional<NARROW, uint32_t, int64_t>::type;
    __shared__ RowNumLDS<V> s;
    const int lane = lane_id();
    const int g = lane >> 4, r = lane & 15;
    const int64_t idx = (int64_t)blockIdx.x * RN_GROUPS + g;
    // this group's tile-row and C tile range (uniform in the group)
    const int32_t bi = bin_row(rows, row_begin, nrows, dlo, dhi, idx);
    // this lane's scalar row of A, its first RN_EB entries and their B row
    // ranges: issued before the C headers, so the two load chains overlap
    const int64_t gr = (int64_t)(bi < 0 ? 0 : bi) * 16 + r;
    IX e0 = 0, e1 = 0;
    if (bi >= 0 && gr < m) {
        e0 = (IX)ld_idx(rpA, (IX)gr);
        e1 = (IX)ld_idx(rpA, (IX)(gr + 1));
    }
    int64_t cs = 0, ce = 0;
    if (bi >= 0) {
        cs = tile_ptrC[bi];
        ce = tile_ptrC[bi + 1];
    }
    int32_t kk[RN_EB];
    V aa[RN_EB];
    IX j0[RN_EB];
    int32_t jn[RN_EB];  // B row lengths, 0 past the A row's end
    // A entries [eb, eb + RN_EB) (read as aligned 16-byte words) and their B
    // rows' ranges; entries past the row's end are given an empty B row
    // (eb < e1)
    auto load_a = [&](IX eb) {
        const int ne = (int)min<int64_t>(RN_EB, (int64_t)(e1 - eb));
        ld_run<int32_t, RN_EB>(colA, eb, ne, kk);
        ld_run<V, RN_EB>(valA, eb, ne, aa);
#pragma unroll
        for (int u = 0; u < RN_EB; u++)
            if (u >= ne) kk[u] = 0;  // past the row's end: B row 0, taken as empty
#pragma unroll
        for (int u = 0; u < RN_EB; u++) {
            const IX b0 = (IX)ld_idx(rpB, (IX)kk[u]), b1 = (IX)ld_idx(rpB, (IX)(kk[u] + 1));
            j0[u] = b0;
            jn[u] = u < ne ? (int32_t)(b1 - b0) : 0;
        }
    };
#pragma unroll
    for (int u = 0; u < RN_EB; u++) {
        kk[u] = 0;
        aa[u] = V(0);
        j0[u] = 0;
        jn[u] = 0;
    }
    if (e1 > e0) load_a(e0);
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
    wave_sync();
    bool pre = e1 > e0;  // the batch at e0 is loaded
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
            for (IX eb = e0; eb < e1; eb += RN_EB) {
                if (!pre) load_a(eb);
                pre = false;
                // RN_BB entries of B row u from its entry o (o < jn[u]), read as
                // aligned 16-byte words; past the row's end column INT_MAX
                auto load_b = [&](int u, int o, int32_t(&c)[RN_BB], V(&b)[RN_BB]) {
                    const int n = jn[u] - o;
                    const IX j = j0[u] + (IX)o;
                    ld_run<int32_t, RN_BB>(colB, j, min(n, RN_BB), c);
                    ld_run<V, RN_BB>(valB, j, min(n, RN_BB), b);
#pragma unroll
                    for (int q = 0; q < RN_BB; q++)
                        if (q >= n) c[q] = INT_MAX;
                };
                // a batch's products into the window, in order; the current C
                // tile's column, this row's first value in it and its row mask
                // stay in registers while the B row stays in that tile.  The
                // output's col byte goes beside its value.
                auto apply = [&](V a, const int32_t(&c)[RN_BB], const V(&b)[RN_BB], int &ct, int32_t &jc,
                                 int32_t &base, uint32_t &mk) {
                    int pos[RN_BB];
#pragma unroll
                    for (int q = 0; q < RN_BB; q++) {
                        const int32_t J = c[q] >> 4;
                        const bool in = J >= jlo && J <= jhi;  // inside the window
                        if (in && J != jc) {
                            while (s.cj[g][ct] < J) ct++;  // J is one of the tile-row's C tile columns
                            const int t = ct * 16 + r;
                            jc = J;
                            base = s.coff[g][ct] - vb + s.cptr[g][t];
                            mk = s.cmask[g][t];
                        }
                        // outside the window: the group's spare slot, whose
                        // contents (lanes race on it) are never read
                        pos[q] = in ? base + __popc(mk & mask_below(c[q] & 15)) : RN_VCAP;
                    }
                    // a batch's columns are distinct (no row holds a column
                    // twice on this path), so are its positions inside the
                    // window: its reads issue together, then its writes
                    V v[RN_BB];
#pragma unroll
                    for (int q = 0; q < RN_BB; q++) v[q] = s.val[g][pos[q]];
#pragma unroll
                    for (int q = 0; q < RN_BB; q++) {
                        s.val[g][pos[q]] = fma_rn<V>(a, b[q], v[q]);
                        s.colb[g][pos[q]] = (uint8_t)((r << 4) | (c[q] & 15));
                    }
                };
                // the first batch of B row u + 1 is loaded before row u's
                // products are applied (its loads overlap that work)
                int32_t cc[RN_BB], cn[RN_BB];
                V bv[RN_BB], bn[RN_BB];
                if (jn[0] > 0) load_b(0, 0, cc, bv);
#pragma unroll
                for (int u = 0; u < RN_EB; u++) {
                    if (u + 1 < RN_EB && jn[u + 1] > 0) load_b(u + 1, 0, cn, bn);
                    if (jn[u] > 0) {
                        int ct = w0;
                        int32_t jc = -1, base = 0;
                        uint32_t mk = 0;
                        apply(aa[u], cc, bv, ct, jc, base, mk);
                        for (int o = RN_BB; o < jn[u]; o += RN_BB) {  // a B row longer than one batch
                            load_b(u, o, cc, bv);
                            apply(aa[u], cc, bv, ct, jc, base, mk);
                        }
                    }
                    if (u + 1 < RN_EB) {
#pragma unroll
                        for (int q = 0; q < RN_BB; q++) {
                            cc[q] = cn[q];
                            bv[q] = bn[q];
                        }
                    }
                }
            }
        }
        wave_sync();
        // the window's values and col bytes (every position took at least
        // one product): the group copies them out in 16-byte words of values
        // and 4-byte words of col bytes, the unaligned ends one at a time
        if (act) {
            V *vo = valC + (vs + vb);
            uint8_t *co8 = colC + (vs + vb);
            constexpr int K = 16 / (int)sizeof(V);
            const int h = min(vn, (int)(((16 - ((uintptr_t)vo & 15)) & 15) / sizeof(V)));
            const int nv = (vn - h) / K, tv = h + nv * K;
            if (r < h) vo[r] = s.val[g][r];
            if (tv + r < vn) vo[tv + r] = s.val[g][tv + r];
            for (int w = r; w < nv; w += 16) {
                const int i = h + w * K;
                V t[K];
#pragma unroll
                for (int k = 0; k < K; k++) t[k] = s.val[g][i + k];
                uint4 u;
                __builtin_memcpy(&u, t, 16);
                *reinterpret_cast<uint4 *>(vo + i) = u;
            }
            const int hb = min(vn, (int)((4 - ((uintptr_t)co8 & 3)) & 3));
            const int nw = (vn - hb) / 4, tb = hb + nw * 4;
            if (r < hb) co8[r] = s.colb[g][r];
            if (tb + r < vn) co8[tb + r] = s.colb[g][tb + r];
            for (int w = r; w < nw; w += 16) {
                const int i = hb + w * 4;
                const uint32_t word = (uint32_t)s.colb[g][i] | ((uint32_t)s.colb[g][i + 1] << 8) |
                                      ((uint32_t)s.colb[g][i + 2] << 16) | ((uint32_t)s.colb[g][i + 3] << 24);
                *reinterpret_cast<uint32_t *>(co8 + i) = word;
            }
        }
        wave_sync();
        w0 = w1;
    }
}

// ---------------------------------------------------------------- symbolic step of bin R
// The staged symbolic kernel runs a wave per tile-row; a tile-row that fits
// R's caps (8 A tiles, 32 B tiles, 16 C tiles) leaves most of its lanes idle
// there.  k_sym_rows gives such a tile-row 16 lanes (4 tile-rows a wave):
// two B tiles a lane for the C tile columns (a 1024-column LDS bitmap, ranks
// by popcount prefix), then lane r builds row r of every C tile -- the OR of
// B rows k selected by A's row mask r, as step2_symbolic_kernel
// (src/spgemm_raw_kernels.cuh:105-163) -- with no atomics; then lane t
// writes C tile t whole: its row masks, its exclusive row offsets
// (step2_build_ptrRowC_kernel, :166-215) and its count.
constexpr int SR_NA = 8, SR_NB = 32, SR_NC = 16;  // R's caps
constexpr int SR_BPL = SR_NB / 16;                 // B tiles a lane
constexpr int SR_WORDS = 32;  // bitmap: 1024 tile columns (a 2-D 5-point row spans 513)

// A group's arrays; the bitmap and the fallback's arrays share their space
// (a group takes one of the two ways).
struct SymRowsGroup {
    alignas(16) uint16_t bmask[SR_NB * 16];
    alignas(16) uint16_t cmask[SR_NC * 16];
    union {
        uint32_t bm[SR_WORDS];  // bitmap way
        int32_t bj[SR_NB];      // fallback: the B tiles' columns
    };
    int32_t abs_[SR_NA];
    int32_t aseg[SR_NA + 1];
    union {
        uint8_t wpre[SR_WORDS];  // bitmap way: distinct columns before each word (at most 32)
        uint8_t first[SR_NB];    // fallback: first occurrence of a column
    };
    uint8_t bct[SR_NB];
};
struct SymRowsLDS {
    SymRowsGroup q[RN_GROUPS];
};

// Group sizes are read and combined with every lane active; `ok` only gates
// memory traffic, so the DPP row operations always see their whole row.
__global__ void __launch_bounds__(WAVE)
k_sym_rows(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nrows, const int64_t *__restrict__ dlo,
           const int64_t *__restrict__ dhi, const int64_t *__restrict__ tile_ptrA,
           const uint16_t *__restrict__ maskA, const int4 *__restrict__ link, const int32_t *__restrict__ colidxB,
           const uint16_t *__restrict__ csr_maskB, const int64_t *__restrict__ tile_ptrC,
           int32_t *__restrict__ rowidxC, int32_t *__restrict__ colidxC, uint16_t *__restrict__ maskC,
           uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC, int32_t *__restrict__ overflow) {
    __shared__ SymRowsLDS s;
    const int lane = lane_id();
    const int g = lane >> 4, l = lane & 15;
    SymRowsGroup &G = s.q[g];
    const int64_t idx = (int64_t)blockIdx.x * RN_GROUPS + g;
    const int32_t bi = bin_row(rows, row_begin, nrows, dlo, dhi, idx);
    int64_t as = 0, ae = 0, cs = 0, ce = 0;
    if (bi >= 0) {
        as = tile_ptrA[bi];
        ae = tile_ptrA[bi + 1];
        cs = tile_ptrC[bi];
        ce = tile_ptrC[bi + 1];
    }
    const int na = (int)(ae - as), cap = (int)(ce - cs);
    bool ok = bi >= 0 && na <= SR_NA && cap <= SR_NC;
    if (bi >= 0 && !ok && l == 0) atomicOr(overflow, 1);
    // A tiles: lane l loads A tile l's link, and row l of every A tile's
    // mask (the only row it needs later)
    int4 lk = make_int4(0, 0, 0, 0);
    uint32_t am[SR_NA];
#pragma unroll
    for (int a = 0; a < SR_NA; a++) am[a] = (ok && a < na) ? maskA[(as + a) * 16 + l] : 0u;
    if (ok && l < na) {
        lk = link[as + l];
        G.abs_[l] = lk.x;
    }
    const int incl = group16_incl_sum(lk.y);
    const int nB = __shfl(incl, 15, 16);
    if (ok && nB > SR_NB) {
        if (l == 0) atomicOr(overflow, 1);
        ok = false;
    }
    if (ok && l < na) G.aseg[l + 1] = incl;
    if (ok && l == 0) G.aseg[0] = 0;
    wave_sync();
    // B tiles t = l + 16 h: tile column and row masks
    int32_t jv[SR_BPL] = {};
    int jlo = -INT_MAX, jhi = -1;  // jlo holds the negated minimum
#pragma unroll
    for (int h = 0; h < SR_BPL; h++) {
        const int t = l + 16 * h;
        if (ok && t < nB) {
            const int a = owner_of(G.aseg, na, t);
            const int64_t q = (int64_t)G.abs_[a] + (t - G.aseg[a]);
            jv[h] = colidxB[q];
            uint4 m0, m1;
            load_mask16(csr_maskB, q, m0, m1);
            store_mask16(&G.bmask[t * 16], m0, m1);
            jlo = max(jlo, -jv[h]);
            jhi = max(jhi, jv[h]);
        }
    }
    jlo = -group16_max(jlo);
    jhi = group16_max(jhi);
    // distinct tile columns and their ranks
    const bool bitmap = ok && jhi - jlo < SR_WORDS * 32;  // uniform in the group
    int rk[SR_BPL] = {};
    int nc = 0;
    if (bitmap)
        for (int i = l; i < SR_WORDS; i += 16) G.bm[i] = 0u;
    wave_sync();
    if (bitmap) {
#pragma unroll
        for (int h = 0; h < SR_BPL; h++) {
            const int t = l + 16 * h;
            if (t < nB) atomicOr(&G.bm[(jv[h] - jlo) >> 5], 1u << ((jv[h] - jlo) & 31));
        }
    }
    wave_sync();
    uint32_t wd[SR_WORDS / 16];
    int pc = 0;
#pragma unroll
    for (int i = 0; i < SR_WORDS / 16; i++) {
        wd[i] = bitmap ? G.bm[l * (SR_WORDS / 16) + i] : 0u;
        pc += __popc(wd[i]);
    }
    const int winc = group16_incl_sum(pc);
    if (bitmap) {
        nc = __shfl(winc, 15, 16);
        int pre = winc - pc;
#pragma unroll
        for (int i = 0; i < SR_WORDS / 16; i++) {
            G.wpre[l * (SR_WORDS / 16) + i] = (uint8_t)pre;
            pre += __popc(wd[i]);
        }
    } else if (ok) {  // columns span more than the bitmap: the columns in LDS
#pragma unroll
        for (int h = 0; h < SR_BPL; h++)
            if (l + 16 * h < nB) G.bj[l + 16 * h] = jv[h];
    }
    wave_sync();
    if (ok && !bitmap) {  // first occurrences, then ranks by counting
#pragma unroll
        for (int h = 0; h < SR_BPL; h++) {
            const int t = l + 16 * h;
            if (t < nB) {
                bool f = true;
                for (int u = 0; u < t; u++) f = f && G.bj[u] != jv[h];
                G.first[t] = f ? 1 : 0;
            }
        }
    }
    wave_sync();
    if (ok) {
#pragma unroll
        for (int h = 0; h < SR_BPL; h++) {
            const int t = l + 16 * h;
            if (t >= nB) continue;
            if (bitmap) {
                const int off = jv[h] - jlo;
                rk[h] = G.wpre[off >> 5] + __popc(G.bm[off >> 5] & ((1u << (off & 31)) - 1u));
            } else {
                int r = 0;
                for (int u = 0; u < nB; u++) r += (G.first[u] && G.bj[u] < jv[h]) ? 1 : 0;
                rk[h] = r;
            }
        }
        if (!bitmap) {
            int n = 0;
            for (int u = 0; u < nB; u++) n += G.first[u];
            nc = n;
        }
    }
    if (ok && nc > cap) {  // cannot happen on the exact layout: the count said cap
        if (l == 0) atomicOr(overflow, 1);
        ok = false;
    }
    if (ok) {
#pragma unroll
        for (int h = 0; h < SR_BPL; h++) {
            const int t = l + 16 * h;
            if (t < nB) {  // duplicates write the same column at the same rank
                colidxC[cs + rk[h]] = jv[h];
                rowidxC[cs + rk[h]] = bi;
                G.bct[t] = (uint8_t)rk[h];
            }
        }
        for (int ct = 0; ct < nc; ct++) G.cmask[ct * 16 + l] = 0;
    }
    wave_sync();
    // row l of every C tile: OR of B rows k over A's row mask l (no other
    // lane writes it).  A tile by A tile: its row mask l is held in a
    // register and its B tiles are the contiguous range aseg[a] .. aseg[a+1].
    if (ok) {
#pragma unroll
        for (int a = 0; a < SR_NA; a++) {
            if (a >= na) break;
            const uint32_t ma0 = am[a];
            if (!ma0) continue;  // row l of this A tile is empty
            const int b1 = G.aseg[a + 1];
            for (int b = G.aseg[a]; b < b1; b++) {
                const int ct = G.bct[b];
                uint32_t ma = ma0, rm = 0;
                while (ma) {
                    const int kk = __clz(ma) - 16;
                    ma ^= 0x8000u >> kk;
                    rm |= G.bmask[b * 16 + kk];
                }
                if (rm) G.cmask[ct * 16 + l] |= (uint16_t)rm;
            }
        }
    }
    wave_sync();
    // lane l writes C tile l whole: its 16 row masks, their exclusive row
    // offsets (step2_build_ptrRowC_kernel's scan) and its count
    if (ok && l < nc) {
        const uint4 m0 = reinterpret_cast<const uint4 *>(&G.cmask[l * 16])[0];
        const uint4 m1 = reinterpret_cast<const uint4 *>(&G.cmask[l * 16])[1];
        const uint32_t w[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};  // rows 2i (low), 2i+1 (high)
        uint32_t pw[4] = {0u, 0u, 0u, 0u};
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            pw[i >> 2] |= acc << (8 * (i & 3));
            acc += __popc((w[i >> 1] >> (16 * (i & 1))) & 0xFFFFu);
        }
        uint4 *mo = reinterpret_cast<uint4 *>(maskC + (cs + l) * 16);
        mo[0] = m0;
        mo[1] = m1;
        reinterpret_cast<uint4 *>(ptrRowC + (cs + l) * 16)[0] = make_uint4(pw[0], pw[1], pw[2], pw[3]);
        tile_nnzC[cs + l + 1] = acc;
    }
}

int launch_sym_rows(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                    const int32_t *rows, const BinRange &br, const int4 *link, const SlotLayout &sl) {
    const int64_t count = br.count;
    if (count <= 0) return TSG_OK;
    if (sl.row_nc) return fail(TSG_ERR_ARG, "the row-lane symbolic kernel runs on the exact C layout only");
    const unsigned grid = (unsigned)((count + RN_GROUPS - 1) / RN_GROUPS);
    k_sym_rows<<<grid, WAVE, 0, ctx->stream>>>(row_list(rows, br, C->tilem), br.begin, count, br.lo, br.hi, A->tile_ptr, A->mask, link,
                                               B->tile_columnidx, B->csr_mask, C->tile_ptr, C->tile_rowidx,
                                               C->tile_columnidx, C->mask, C->tile_csr_Ptr, C->tile_nnz, sl.overflow);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
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
    // the kernel reads A's and B's scalar CSR entries as aligned 16-byte words
    // (the library allocates them on 256-byte boundaries)
    auto al16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
    if (!al16(A->row_col) || !al16(A->row_val) || !al16(B->row_col) || !al16(B->row_val))
        return fail(TSG_ERR_ARG, "row-lane numeric: scalar CSR arrays must start on 16-byte boundaries");
    // 32-bit byte offsets reach every scalar CSR entry of A and B
    const int64_t lim = (int64_t)1 << 29;
    const bool narrow = A->nnz < lim && B->nnz < lim && A->m < lim && B->m < lim;
#define TSG_NUM_ROWS(V, N)                                                                                        \
    k_num_rows<V, N><<<grid, WAVE, 0, ctx->stream>>>(                                                             \
        row_list(rows, br, C->tilem), br.begin, count, br.lo, br.hi, A->m, A->row_ptr, A->row_col,               \
        (const V *)A->row_val, B->row_ptr, B->row_col, (const V *)B->row_val, C->tile_ptr, C->tile_columnidx,    \
        C->tile_nnz, C->mask, C->tile_csr_Ptr, (V *)C->tile_csr_Val, C->tile_csr_Col, sl.overflow, sl.val_cap, \
        sl.val_done)
    if (A->dtype == TSG_F32) {
        if (narrow)
            TSG_NUM_ROWS(float, true);
        else
            TSG_NUM_ROWS(float, false);
    } else {
        if (narrow)
            TSG_NUM_ROWS(double, true);
        else
            TSG_NUM_ROWS(double, false);
    }
#undef TSG_NUM_ROWS
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

}  // namespace tsg
