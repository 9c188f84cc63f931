// rowstaged.hip -- the binned pipeline: per-tile-row statistics, row binning,
// and the wave-per-tile-row staged symbolic and numeric kernels.
//
// The reference works C tile by C tile (src/spgemm_raw_kernels.cuh:105-202,
// src/spgemm_raw_kernels_v2.cuh:10-139): every tile walks a chain of dependent
// global loads (its row pointers, the two index lists, each pair's headers,
// each pair's values), so on a 5-point Laplacian most of the time is load
// latency.  Here a whole C tile-row is one wave64's job.  A per-call "link"
// array gives, for every A tile (bi,k), the CSR-tile range and payload range
// of B's tile-row k (B keeps its CSR-tile-order data next to its CSC arrays,
// so a B tile-row is one contiguous range).  A wave then stages everything its
// row needs into LDS in three bulk, coalesced load levels:
//   1. the A and C tile-row pointers
//   2. A's tile headers and links, C's headers
//   3. A's values, B's tile headers and values
// and computes from LDS only.  The numeric step gives every lane one output
// nonzero at a time (outputs of a tile-row are contiguous in C, so the stores
// coalesce); an output (r,c) of C tile ct runs ct's matched tile pairs in
// ascending k and, inside a pair, the columns kk of A's row r in ascending
// order, one FMA per product -- the summation order of the C-tile kernels and
// of the oracle, so every path yields identical bits.
//
// A row is staged only when its sizes fit a variant's LDS capacities; rows are
// binned by exact statistics emitted with the Step-1 count, and rows that fit
// no variant run the C-tile kernels, one workgroup per row.
#include <climits>
#include <cstring>

#include "rowstaged.h"

namespace tsg {

// ---------------------------------------------------------------- variants
template <int NA_, int NC_, int NB_, int PA_, int PB_, int WAVES_>
struct Variant {
    static constexpr int NA = NA_, NC = NC_, NB = NB_, PA = PA_, PB = PB_, WAVES = WAVES_;
};
// S: 2-D 5-point-like rows; W: wide A tile-rows (27-point stencil times an
// aggregation); M: medium rows of irregular matrices.
using VarS = Variant<8, 16, 32, 128, 1 << 20, 4>;
using VarW = Variant<32, 16, 32, 512, 1 << 20, 2>;
using VarM = Variant<32, 48, 64, 512, 1 << 20, 2>;
// R: rows of VarS whose scalar rows make about 32 products or fewer; both
// steps run 16 lanes per tile-row (rownum.hip): the symbolic step k_sym_rows,
// the numeric step a lane per scalar row, which wins on short rows (2-D
// 5-point stencils: 25) and loses on long ones (3-D 7-point: 49).  (Fewer
// than ROW_LANES_MIN such rows take VarS's kernels.)
constexpr StageCaps CAPS_R{8, 16, 32, 128, 1 << 20, 32};
constexpr int NUM_STAGED = 4;  // bins 0..3: R, S, W, M
constexpr int64_t ROW_LANES_MIN = 32768;  // tile-rows below which R's rows take VarS's numeric kernel
constexpr int DESC_CAP = 256;  // output descriptors of one chunk of (C tile, row) items
constexpr int NS_K = 4;        // products of one output whose value loads issue together

template <typename Var>
static StageCaps caps_of() {
    return StageCaps{Var::NA, Var::NC, Var::NB, Var::PA, Var::PB};
}

// Largest t in [0,n) with seg[t] <= i (seg ascending, seg[0] = 0).

// Bits of a 16-bit row mask that stand for columns below c (bit 15-c is column c).
__device__ __forceinline__ uint32_t below(int c) { return (0xFFFF0000u >> c) & 0xFFFFu; }

// ---------------------------------------------------------------- A->B links
// link[t] = {first CSR tile of B's tile-row k, its tile count, its payload
// count} and lpay[t] = its first payload entry, for A tile t in column k.
__global__ void __launch_bounds__(256)
k_ab_link(int64_t numtileA, const int32_t *__restrict__ colidxA, const int64_t *__restrict__ tile_ptrB,
          const int64_t *__restrict__ rnpB, int4 *__restrict__ link, int64_t *__restrict__ lpay,
          int32_t *__restrict__ zero, int nzero) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < nzero) zero[t] = 0;  // the plan's flags and counters (no launch of their own)
    if (t >= numtileA) return;
    const int32_t k = colidxA[t];
    const int64_t bs = tile_ptrB[k], be = tile_ptrB[k + 1];
    int64_t ps = 0, pe = 0;
    if (rnpB) { ps = rnpB[k]; pe = rnpB[k + 1]; }
    link[t] = make_int4((int32_t)bs, (int32_t)(be - bs), (int32_t)min<int64_t>(pe - ps, INT_MAX), 0);
    lpay[t] = ps;
}

// ---------------------------------------------------------------- Step 1 count + statistics
constexpr int RS_WAVES = 4;
constexpr uint8_t BIN_REST = 0xFF;  // k_row_light's mark: counted by the wave-per-row kernel (bins are < MAX_BINS)
constexpr int RS_WIN_WORDS = 512;  // 16,384 tile columns per pass

// Ascending bitonic sort of one 64-bit key per lane across the wave.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t key) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= WAVE; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t other = __shfl_xor(key, j, WAVE);
            const bool up = (lane & k) == 0, lower = (lane & j) == 0;
            const uint64_t mn = key < other ? key : other, mx = key < other ? other : key;
            key = (lower == up) ? mn : mx;
        }
    }
    return key;
}

// One tile-row's Step-1 results: its C tile count into tile_ptrC (scanned
// later), its bin (the first staged variant whose capacities hold it, else
// the row-by-row or the C-tile kernels) and, for the learned scheduler, its
// statistics (heavy rows' count comes from k_heavy_structure).
// pa: the tile-row's A entries.
__device__ __forceinline__ void row_result(int32_t bi, int64_t na, int64_t pa, int64_t nb, int64_t pb, int64_t count,
                                           int32_t cmin, int32_t cmax, int64_t f_maxlen, int64_t f_sumspan,
                                           int64_t f_nonempty, const BinTable &bt, int64_t *__restrict__ tile_ptrC,
                                           uint8_t *__restrict__ bin, int32_t *__restrict__ stats) {
    tile_ptrC[bi + 1] = count;
    if (bi == 0) tile_ptrC[0] = 0;
    int b = bt.nvar;
    for (int v = 0; v < bt.nvar; v++) {
        const StageCaps c = bt.v[v];
        if (na <= c.na && count <= c.nc && nb <= c.nb && pa <= c.pa && pb <= c.pb &&
            (c.prow <= 0 || pa * pb <= (int64_t)c.prow * 256 * na)) {
            b = v;
            break;
        }
    }
    if (b == bt.nvar && bt.rg && pa <= (int64_t)bt.rg_density * na) b = bt.nvar + 1;
    bin[bi] = (uint8_t)b;
    if (stats) {
        int32_t *st = stats + (int64_t)bi * RS_NSTAT;
        st[RS_NA] = (int32_t)min<int64_t>(na, INT_MAX);
        st[RS_PA] = (int32_t)min<int64_t>(pa, INT_MAX);
        st[RS_NB] = (int32_t)min<int64_t>(nb, INT_MAX);
        st[RS_PB] = (int32_t)min<int64_t>(pb, INT_MAX);
        st[RS_COUNT] = (int32_t)min<int64_t>(count, INT_MAX);
        st[RS_SPAN] = cmax >= cmin ? cmax - cmin + 1 : 0;
        st[RS_MAXLEN] = (int32_t)min<int64_t>(f_maxlen, INT_MAX);
        st[RS_SUMSPAN] = (int32_t)min<int64_t>(f_sumspan, INT_MAX);
        st[RS_NONEMPTY] = (int32_t)min<int64_t>(f_nonempty, INT_MAX);
    }
}

// Exact C tile count of each tile-row (as step1_count_kernel,
// src/spgemm_raw_kernels.cuh:23-57) plus the sizes that decide its bin:
// A tiles, A payload, B tiles and B payload reached, C tiles.  Rows reaching
// at most 64 B tiles count distinct columns with a 2048-column LDS bitmap when
// the columns span no more, else with one wave-wide sort; larger rows use an
// LDS bitmap over successive column windows.
__device__ __forceinline__ void
row_stats_one(int32_t bi, const int64_t *__restrict__ tile_ptrA, const int64_t *__restrict__ tile_nnzA,
              const int4 *__restrict__ link, const int32_t *__restrict__ colidxB, const BinTable &bt,
              int64_t *__restrict__ tile_ptrC, uint8_t *__restrict__ bin, int32_t *__restrict__ stats,
              int32_t *__restrict__ hlist, int32_t *__restrict__ hcount, uint32_t *bm, int32_t *s_bsw,
              int32_t *s_prew) {
    const int lane = lane_id();
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];
    const int64_t na = ae - as;
    int32_t cmin = INT_MAX, cmax = -1;
    int64_t nb = 0, pb = 0, count = 0;
    int64_t f_maxlen = 0, f_sumspan = 0, f_nonempty = 0;  // learned-scheduler features (stats only)
    bool done = false;
    if (na <= WAVE) {
        const int4 l = lane < na ? link[as + lane] : make_int4(0, 0, 0, 0);
        const int incl = dpp_incl_sum(l.y);
        nb = last_lane(incl);
        pb = last_lane(dpp_incl_sum(l.z));
        if (nb <= WAVE) {
            if (lane < na) {
                s_prew[lane + 1] = incl;
                s_bsw[lane] = l.x;
            }
            if (lane == 0) s_prew[0] = 0;
            wave_sync();
            int32_t j = 0;
            bool seg_first = false, seg_last = false;  // first / last B tile of its A tile's B tile-row
            if (lane < nb) {
                const int t = owner_of(s_prew, (int)na, lane);
                const int pos = lane - s_prew[t];
                j = colidxB[s_bsw[t] + pos];
                seg_first = pos == 0;
                seg_last = lane == s_prew[t + 1] - 1;
            }
            if (stats) {  // learned-scheduler features from the columns already loaded
                // each A tile's span = last - first + 1 of its B tile-row's columns
                f_sumspan = wave_sum((int64_t)(seg_last ? j : 0) - (seg_first ? (int64_t)j - 1 : 0));
                f_maxlen = last_lane(dpp_incl_max(lane < na ? l.y : 0));
                f_nonempty = __popcll(__ballot(lane < na && l.y > 0));
            }
            if (nb > 0) {
                cmin = -last_lane(dpp_incl_max(lane < nb ? -j : -INT_MAX));
                cmax = last_lane(dpp_incl_max(lane < nb ? j : -1));
            }
            if (nb > 0 && cmax - cmin < WAVE * 32) {
                // distinct columns = set bits of a bitmap over [cmin, cmin + 2048)
                bm[lane] = 0u;
                wave_sync();
                if (lane < nb) atomicOr(&bm[(j - cmin) >> 5], 1u << ((j - cmin) & 31));
                wave_sync();
                count = last_lane(dpp_incl_sum(__popc(bm[lane])));
            } else {
                const uint64_t key = wave_sort64(lane < nb ? (uint64_t)(uint32_t)j : ~0ull);
                const int32_t sj = (int32_t)(uint32_t)key;
                const int32_t prev = __shfl_up(sj, 1, WAVE);
                const bool first = lane < nb && (lane == 0 || prev != sj);
                count = __popcll(__ballot(first));
            }
            done = true;
        }
    }
    if (!done) {
        nb = 0;
        pb = 0;
        for (int64_t i = as + lane; i < ae; i += WAVE) {
            const int4 l = link[i];
            if (l.y > 0) {
                const int32_t f = colidxB[l.x], e = colidxB[l.x + l.y - 1];
                cmin = min(cmin, f);
                cmax = max(cmax, e);
                f_sumspan += e - f + 1;
                f_maxlen = max<int64_t>(f_maxlen, l.y);
                f_nonempty++;
            }
            nb += l.y;
            pb += l.z;
        }
        cmin = wave_min(cmin);
        cmax = wave_max(cmax);
        nb = wave_sum(nb);
        pb = wave_sum(pb);
        if (stats) {
            f_sumspan = wave_sum(f_sumspan);
            f_maxlen = wave_max(f_maxlen);
            f_nonempty = wave_sum(f_nonempty);
        }
        const bool heavy = hlist && nb > HEAVY_NB;  // counted by a workgroup afterwards
        if (heavy && lane == 0) hlist[atomicAdd(hcount, 1)] = bi;
        for (int64_t w0 = cmin; !heavy && w0 <= cmax; w0 += (int64_t)RS_WIN_WORDS * 32) {
            const int32_t wbits = (int32_t)min<int64_t>((int64_t)RS_WIN_WORDS * 32, cmax - w0 + 1);
            const int nwords = (wbits + 31) >> 5;
            for (int i = lane; i < nwords; i += WAVE) bm[i] = 0;
            wave_sync();
            for (int64_t c0 = as; c0 < ae; c0 += WAVE) {
                const int64_t i = c0 + lane;
                int32_t len = 0, bs = 0;
                if (i < ae) {
                    const int4 l = link[i];
                    bs = l.x;
                    len = l.y;
                }
                const int32_t incl = wave_incl_sum(len);
                const int32_t total = __shfl(incl, WAVE - 1, WAVE);
                s_prew[lane] = incl;
                s_bsw[lane] = bs;
                wave_sync();
                for (int32_t t = lane; t < total; t += WAVE) {
                    int lo = 0, hi = WAVE - 1;
                    while (lo < hi) {
                        int mid = (lo + hi) >> 1;
                        if (s_prew[mid] > t) hi = mid; else lo = mid + 1;
                    }
                    const int32_t excl = lo ? s_prew[lo - 1] : 0;
                    const int64_t off = (int64_t)colidxB[s_bsw[lo] + (t - excl)] - w0;
                    if (off >= 0 && off < wbits) atomicOr(&bm[off >> 5], 1u << (off & 31));
                }
                wave_sync();
            }
            int32_t cnt = 0;
            for (int i = lane; i < nwords; i += WAVE) cnt += __popc(bm[i]);
            count += wave_sum(cnt);
            wave_sync();
        }
    }
    if (lane == 0)
        row_result(bi, na, tile_nnzA[ae] - tile_nnzA[as], nb, pb, count, cmin, cmax, f_maxlen, f_sumspan,
                   f_nonempty, bt, tile_ptrC, bin, stats);
}

// One wave per chunk of RS_CHUNK consecutive tile-rows: the rows k_row_light
// left (marked BIN_REST), taken one after the other.  A grid of chunks, not of
// rows, so a product with no such row costs a short launch.
constexpr int RS_CHUNK = 8;
__global__ void __launch_bounds__(RS_WAVES *WAVE)
k_row_stats(const int64_t *__restrict__ tile_ptrA, const int64_t *__restrict__ tile_nnzA, int32_t tilemA,
            const int4 *__restrict__ link, const int32_t *__restrict__ colidxB, BinTable bt,
            int64_t *__restrict__ tile_ptrC, uint8_t *__restrict__ bin, int32_t *__restrict__ stats,
            int32_t *__restrict__ hlist, int32_t *__restrict__ hcount) {
    __shared__ uint32_t s_bm[RS_WAVES][RS_WIN_WORDS];
    __shared__ int32_t s_bs[RS_WAVES][WAVE];
    __shared__ int32_t s_pre[RS_WAVES][WAVE + 1];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t c0 = ((int64_t)blockIdx.x * RS_WAVES + w) * RS_CHUNK;
    const bool marked = lane < RS_CHUNK && c0 + lane < tilemA && bin[c0 + lane] == BIN_REST;
    uint64_t mk = __ballot(marked);
    while (mk) {  // wave-uniform
        const int32_t bi = (int32_t)(c0 + __ffsll((unsigned long long)mk) - 1);
        mk &= mk - 1;
        row_stats_one(bi, tile_ptrA, tile_nnzA, link, colidxB, bt, tile_ptrC, bin, stats, hlist, hcount, s_bm[w],
                      s_bs[w], s_pre[w]);
        wave_sync();
    }
}

// Step-1 count of light tile-rows (at most 32 A tiles, 64 B tiles reached,
// C tile columns spanning less than 1024) at RL_GL = 8 lanes each, 8 a wave:
// the wave-per-row kernel above leaves most of its lanes idle on them.  Lane
// l takes A tiles l + 8 h and B tiles l + 8 h; the passes past the first are
// skipped when no group of the wave needs them.  Each other row's bin is set
// to BIN_REST for that kernel (a mark, not a list: appending to a list
// through one counter serialised on rows that are all heavy).  Results and
// statistics are the ones the wave-per-row kernel computes.
constexpr int RL_NA = 32, RL_NB = 64, RL_WORDS = 32;
constexpr int RL_GL = 8, RL_GROUPS = WAVE / RL_GL;           // lanes a tile-row, tile-rows a wave
constexpr int RL_APL = RL_NA / RL_GL, RL_BPL = RL_NB / RL_GL;  // A and B tiles a lane

template <typename T>
__device__ __forceinline__ T group8_incl_sum(T v) {
    const int l = lane_id() & (RL_GL - 1);
#pragma unroll
    for (int d = 1; d < RL_GL; d <<= 1) {
        const T o = __shfl_up(v, d, RL_GL);
        if (l >= d) v += o;
    }
    return v;
}
__device__ __forceinline__ int group8_max(int v) {
#pragma unroll
    for (int d = RL_GL / 2; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, RL_GL));
    return v;
}

__global__ void __launch_bounds__(WAVE)
k_row_light(const int64_t *__restrict__ tile_ptrA, const int64_t *__restrict__ tile_nnzA, int32_t tilemA,
            const int4 *__restrict__ link, const int32_t *__restrict__ colidxB, BinTable bt,
            int64_t *__restrict__ tile_ptrC, uint8_t *__restrict__ bin, int32_t *__restrict__ stats) {
    __shared__ uint32_t s_bm[RL_GROUPS][RL_WORDS];
    __shared__ int32_t s_bs[RL_GROUPS][RL_NA];
    __shared__ int32_t s_pre[RL_GROUPS][RL_NA + 1];
    const int lane = lane_id(), g = lane / RL_GL, l = lane % RL_GL;
    const int32_t bi = blockIdx.x * RL_GROUPS + g;
    const bool row = bi < tilemA;
    int64_t as = 0, ae = 0;
    if (row) {
        as = tile_ptrA[bi];
        ae = tile_ptrA[bi + 1];
    }
    const int64_t na = ae - as;
    bool light = row && na <= RL_NA;
    // every lane of the group goes through the group operations; `light`
    // (uniform in the group) gates memory traffic
    int4 lk[RL_APL];
#pragma unroll
    for (int h = 0; h < RL_APL; h++) {
        lk[h] = make_int4(0, 0, 0, 0);
        if (light && l + RL_GL * h < na) lk[h] = link[as + l + RL_GL * h];
    }
    // the row's A entries (its bin needs them), loaded with the links
    int64_t nz = 0;
    if (light && l < 2) nz = tile_nnzA[l == 0 ? as : ae];
    const int64_t pa = __shfl(nz, 1, RL_GL) - __shfl(nz, 0, RL_GL);
    // B tiles before each A tile's end (A tile order l, l + 8, ...)
    int incl[RL_APL];
    int nb = 0;
    int64_t pb = 0;
#pragma unroll
    for (int h = 0; h < RL_APL; h++) {
        incl[h] = nb;
        if (h == 0 || __any(light && na > RL_GL * h)) {  // wave-uniform
            incl[h] = group8_incl_sum(lk[h].y) + nb;
            nb = __shfl(incl[h], RL_GL - 1, RL_GL);
            pb += __shfl(group8_incl_sum<int64_t>((int64_t)lk[h].z), RL_GL - 1, RL_GL);
        }
    }
    light = light && nb <= RL_NB;
    if (light) {
#pragma unroll
        for (int h = 0; h < RL_APL; h++) {
            const int t = l + RL_GL * h;
            if (t < na) {
                s_pre[g][t + 1] = incl[h];
                s_bs[g][t] = lk[h].x;
            }
        }
    }
    if (light && l == 0) s_pre[g][0] = 0;
    wave_sync();
    int32_t j[RL_BPL] = {};
    int64_t sumspan = 0;
    int32_t cmn = -INT_MAX, cmx = -1;  // cmn: the negated minimum
#pragma unroll
    for (int h = 0; h < RL_BPL; h++) {
        const int t = l + RL_GL * h;
        if (h > 0 && !__any(light && t < nb)) break;  // no group has B tiles this far
        if (light && t < nb) {
            const int a = owner_of(s_pre[g], (int)na, t);
            const int pos = t - s_pre[g][a];
            j[h] = colidxB[s_bs[g][a] + pos];
            // each A tile's span: last - first + 1 of its B tile-row's columns
            if (t == s_pre[g][a + 1] - 1) sumspan += j[h];
            if (pos == 0) sumspan -= (int64_t)j[h] - 1;
            cmn = max(cmn, -j[h]);
            cmx = max(cmx, j[h]);
        }
    }
    const int32_t cmin = -group8_max(cmn), cmax = group8_max(cmx);
    light = light && (nb == 0 || cmax - cmin < RL_WORDS * 32);
    if (light)
        for (int i = l; i < RL_WORDS; i += RL_GL) s_bm[g][i] = 0u;
    wave_sync();
    if (light)
#pragma unroll
        for (int h = 0; h < RL_BPL; h++)
            if (l + RL_GL * h < nb) atomicOr(&s_bm[g][(j[h] - cmin) >> 5], 1u << ((j[h] - cmin) & 31));
    wave_sync();
    int pc = 0;
    if (light)
#pragma unroll
        for (int i = 0; i < RL_WORDS / RL_GL; i++) pc += __popc(s_bm[g][l * (RL_WORDS / RL_GL) + i]);
    const int cnt = __shfl(group8_incl_sum(pc), RL_GL - 1, RL_GL);
    int64_t f_sumspan = 0, f_maxlen = 0, f_nonempty = 0;
    if (stats) {
        const int64_t ss = group8_incl_sum<int64_t>(sumspan);
        f_sumspan = __shfl(ss, RL_GL - 1, RL_GL);
        int ml = 0, ne = 0;
#pragma unroll
        for (int h = 0; h < RL_APL; h++) {
            const bool in = l + RL_GL * h < na;
            ml = max(ml, in ? lk[h].y : 0);
            ne += (light && in && lk[h].y > 0) ? 1 : 0;
        }
        f_maxlen = group8_max(ml);
        f_nonempty = __shfl(group8_incl_sum(ne), RL_GL - 1, RL_GL);
    }
    if (l == 0 && row) {
        if (light)
            row_result(bi, na, pa, nb, pb, cnt, nb > 0 ? cmin : INT_MAX, nb > 0 ? cmax : -1, f_maxlen, f_sumspan,
                       f_nonempty, bt, tile_ptrC, bin, stats);
        else
            bin[bi] = BIN_REST;
    }
}

// ---------------------------------------------------------------- row binning
constexpr int BIN_THREADS = 256;
constexpr int BIN_ITEMS = 4;
constexpr int BIN_TILE = BIN_THREADS * BIN_ITEMS;

__global__ void __launch_bounds__(BIN_THREADS)
k_bin_count(const uint8_t *__restrict__ bin, int64_t n, int nbins, int64_t nblocks, int64_t *__restrict__ counts) {
    __shared__ int s_cnt[MAX_BINS];
    if (threadIdx.x < MAX_BINS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BIN_ITEMS; i++) {
        const int64_t idx = (int64_t)blockIdx.x * BIN_TILE + i * BIN_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&s_cnt[bin[idx]], 1);
    }
    __syncthreads();
    if (threadIdx.x < nbins) counts[1 + threadIdx.x * nblocks + blockIdx.x] = s_cnt[threadIdx.x];
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[0] = 0;
}

// Stable partition of row ids by bin: rows of bin b, ascending, at
// [start_b, start_{b+1}).  counts holds the inclusive scan of
// [0, (bin, block) counts in bin-major order].
__global__ void __launch_bounds__(BIN_THREADS)
k_bin_scatter(const uint8_t *__restrict__ bin, int64_t n, int nbins, int64_t nblocks,
              const int64_t *__restrict__ counts, int32_t *__restrict__ rows, const int64_t *__restrict__ total0,
              const int64_t *__restrict__ total1, int64_t *__restrict__ hdr) {
    __shared__ int s_wc[BIN_THREADS / WAVE][MAX_BINS];
    __shared__ int64_t s_base[MAX_BINS];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    if (blockIdx.x == 0) {  // the header of the host read (as k_bin_header)
        const int t = threadIdx.x;
        if (t == 0) hdr[0] = total0 ? *total0 : 0;
        if (t == 1) hdr[1] = total1 ? *total1 : 0;
        if (t <= nbins) hdr[2 + t] = counts[(int64_t)t * nblocks];
    }
    if (threadIdx.x < nbins) s_base[threadIdx.x] = counts[threadIdx.x * nblocks + blockIdx.x];
    __syncthreads();
    for (int round = 0; round < BIN_ITEMS; round++) {
        const int64_t idx = (int64_t)blockIdx.x * BIN_TILE + round * BIN_THREADS + threadIdx.x;
        const int b = idx < n ? bin[idx] : 0xFF;
        int rank = 0;
        for (int bb = 0; bb < nbins; bb++) {
            const uint64_t m = __ballot(b == bb);
            if (b == bb) rank = mbcnt(m);
            if (lane == 0) s_wc[w][bb] = __popcll(m);
        }
        __syncthreads();
        if (b < nbins) {
            int64_t off = s_base[b];
            for (int q = 0; q < w; q++) off += s_wc[q][b];
            rows[off + rank] = (int32_t)idx;
        }
        __syncthreads();
        if (threadIdx.x < nbins) {
            int t = 0;
            for (int q = 0; q < BIN_THREADS / WAVE; q++) t += s_wc[q][threadIdx.x];
            s_base[threadIdx.x] += t;
        }
        __syncthreads();
    }
}

// hdr[0] = *total0, hdr[1] = *total1 (0 when absent), hdr[2 + b] = first
// row-list slot of bin b (b <= nbins).
__global__ void k_bin_header(const int64_t *__restrict__ counts, int nbins, int64_t nblocks,
                             const int64_t *__restrict__ total0, const int64_t *__restrict__ total1,
                             int64_t *__restrict__ hdr) {
    const int t = threadIdx.x;
    if (t == 0) hdr[0] = total0 ? *total0 : 0;
    if (t == 1) hdr[1] = total1 ? *total1 : 0;
    if (t <= nbins) hdr[2 + t] = counts[(int64_t)t * nblocks];
}

// ---------------------------------------------------------------- Step-1 tail in two launches
// After the Step-1 count: the exclusive scan of the tile-rows' C tile counts
// (tile_ptrC, the reference's first ExclusiveSum, src/spgemm_raw_v2.cu:131-138),
// the stable partition of the tile-rows by bin, and the header the host reads
// (numtileC, bin offsets) -- two kernels instead of a scan (2), the bin count,
// its scan, the scatter and the publish.  A block takes RT_TILE consecutive
// tile-rows; the apply kernel's blocks add up the earlier blocks' partials
// themselves (at most RT_MAX_BLOCKS of them).
constexpr int RT_THREADS = 256;
constexpr int RT_ITEMS = 4;
constexpr int RT_TILE = RT_THREADS * RT_ITEMS;
constexpr int64_t RT_MAX_BLOCKS = 2048;

// partial[b] = C tiles of block b's rows; bc[bin * nb + b] = its rows in each bin.
__global__ void __launch_bounds__(RT_THREADS)
k_rows_reduce(const int64_t *__restrict__ tile_ptrC, const uint8_t *__restrict__ bin, int64_t n, int nbins,
              int64_t nb, int64_t *__restrict__ partial, int32_t *__restrict__ bc) {
    __shared__ int64_t s_w[RT_THREADS / WAVE];
    __shared__ int32_t s_c[RT_THREADS / WAVE][MAX_BINS];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t base = (int64_t)blockIdx.x * RT_TILE;
    int64_t sum = 0;
    int32_t c[MAX_BINS] = {};
#pragma unroll
    for (int i = 0; i < RT_ITEMS; i++) {
        const int64_t idx = base + i * RT_THREADS + threadIdx.x;
        const int b = idx < n ? bin[idx] : 0xFF;
        sum += idx < n ? tile_ptrC[idx + 1] : 0;
#pragma unroll
        for (int bb = 0; bb < MAX_BINS; bb++) c[bb] += __popcll(__ballot(b == bb));
    }
    sum = wave_sum(sum);
    if (lane == 0) {
        s_w[w] = sum;
#pragma unroll
        for (int bb = 0; bb < MAX_BINS; bb++) s_c[w][bb] = c[bb];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int q = 0; q < RT_THREADS / WAVE; q++) t += s_w[q];
        partial[blockIdx.x] = t;
    }
    if (threadIdx.x < nbins) {
        int32_t t = 0;
        for (int q = 0; q < RT_THREADS / WAVE; q++) t += s_c[q][threadIdx.x];
        bc[(int64_t)threadIdx.x * nb + blockIdx.x] = t;
    }
}

// Block b: its rows' tile_ptrC (inclusive scan of the counts plus the earlier
// blocks' tiles), its rows' places in the bin-ordered row list, and (block 0)
// the header hdr[0] = numtileC, hdr[1] = *total1 (0 when null), hdr[2 + bin] = the bin's first
// slot (bin <= nbins), also published to the host's pinned words (pub, then
// the sequence number seq; pub null: not published).
__global__ void __launch_bounds__(RT_THREADS)
k_rows_apply(int64_t *__restrict__ tile_ptrC, const uint8_t *__restrict__ bin, int64_t n, int nbins, int64_t nb,
             const int64_t *__restrict__ partial, const int32_t *__restrict__ bc, int32_t *__restrict__ rows,
             const int64_t *__restrict__ total1, int64_t *__restrict__ hdr, int32_t *pub, int64_t seq, bool guard,
             BinCaps caps, int64_t cap_tiles, int32_t *__restrict__ overflow, int64_t *__restrict__ numtile_out,
             int nstaged) {
    __shared__ int64_t s_d[RT_TILE + RT_TILE / 32];
    __shared__ int64_t s_w[RT_THREADS / WAVE];
    __shared__ int64_t s_red[RT_THREADS / WAVE][2 * MAX_BINS + 2];
    __shared__ int64_t s_base[MAX_BINS];
    __shared__ int64_t s_hdr[MAX_BINS + 3];
    __shared__ int s_wc[RT_THREADS / WAVE][MAX_BINS];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t blk = blockIdx.x;
    // the earlier blocks' tiles, every block's tiles, each bin's rows (all
    // blocks) and its rows in the earlier blocks
    int64_t carry = 0, total = 0, tot[MAX_BINS] = {}, pre[MAX_BINS] = {};
    for (int64_t j = threadIdx.x; j < nb; j += RT_THREADS) {
        const int64_t pj = partial[j];
        total += pj;
        carry += j < blk ? pj : 0;
#pragma unroll
        for (int bb = 0; bb < MAX_BINS; bb++) {
            const int64_t v = bb < nbins ? bc[(int64_t)bb * nb + j] : 0;
            tot[bb] += v;
            pre[bb] += j < blk ? v : 0;
        }
    }
    carry = wave_sum(carry);
    total = wave_sum(total);
#pragma unroll
    for (int bb = 0; bb < MAX_BINS; bb++) {
        tot[bb] = wave_sum(tot[bb]);
        pre[bb] = wave_sum(pre[bb]);
    }
    if (lane == 0) {
        s_red[w][0] = carry;
        s_red[w][1] = total;
#pragma unroll
        for (int bb = 0; bb < MAX_BINS; bb++) {
            s_red[w][2 + bb] = tot[bb];
            s_red[w][2 + MAX_BINS + bb] = pre[bb];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t r[2 * MAX_BINS + 2] = {};
        for (int q = 0; q < RT_THREADS / WAVE; q++)
            for (int k = 0; k < 2 * MAX_BINS + 2; k++) r[k] += s_red[q][k];
        s_red[0][0] = r[0];  // carry
        int64_t first = 0;  // first slot of bin bb = rows in the bins before it
        s_hdr[0] = r[1];
        s_hdr[1] = total1 ? *total1 : 0;
        for (int bb = 0; bb <= nbins; bb++) {
            s_hdr[2 + bb] = first;
            if (bb < nbins) {
                s_base[bb] = first + r[2 + MAX_BINS + bb];
                first += r[2 + bb];
            }
        }
    }
    __syncthreads();
    if (blk == 0) {  // the header: device copy, then the host's words and sequence number
        const int nw = 2 * (nbins + 3);
        if (guard) {  // past a capacity: the flag, and every bin emptied (no later kernel writes C)
            bool over = s_hdr[0] > cap_tiles;
            for (int b = 0; b < nbins; b++) over = over || s_hdr[3 + b] - s_hdr[2 + b] > caps.v[b];
            if (over && threadIdx.x == 0) atomicOr(overflow, 1);
            if (threadIdx.x == 0 && numtile_out) *numtile_out = s_hdr[0];
            __syncthreads();
            if (over && (int)threadIdx.x <= nbins) s_hdr[2 + threadIdx.x] = 0;
            __syncthreads();
        }
        if ((int)threadIdx.x < nbins + 3) hdr[threadIdx.x] = s_hdr[threadIdx.x];
        if (pub) {
            if ((int)threadIdx.x < nw) pub[threadIdx.x] = reinterpret_cast<const int32_t *>(s_hdr)[threadIdx.x];
            __threadfence_system();
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(reinterpret_cast<int64_t *>(pub + PUB_WORDS), seq, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    // tile_ptrC of this block's rows: counts at [base + 1, base + RT_TILE]
    const int64_t base = blk * RT_TILE;
    auto pad = [](int i) { return i + (i >> 5); };
#pragma unroll
    for (int i = 0; i < RT_ITEMS; i++) {
        const int li = i * RT_THREADS + threadIdx.x;
        const int64_t idx = base + li;
        s_d[pad(li)] = idx < n ? tile_ptrC[idx + 1] : 0;
    }
    __syncthreads();
    int64_t v[RT_ITEMS];
    int64_t run = 0;
#pragma unroll
    for (int i = 0; i < RT_ITEMS; i++) {
        run += s_d[pad(threadIdx.x * RT_ITEMS + i)];
        v[i] = run;
    }
    const int64_t incl = wave_incl_sum(run);
    if (lane == WAVE - 1) s_w[w] = incl;
    __syncthreads();
    int64_t off = s_red[0][0] + incl - run;
    for (int q = 0; q < w; q++) off += s_w[q];
#pragma unroll
    for (int i = 0; i < RT_ITEMS; i++) s_d[pad(threadIdx.x * RT_ITEMS + i)] = v[i] + off;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RT_ITEMS; i++) {
        const int li = i * RT_THREADS + threadIdx.x;
        const int64_t idx = base + li;
        if (idx < n) tile_ptrC[idx + 1] = s_d[pad(li)];
    }
    // stable partition: rows of bin b ascending from s_base[b].  With host
    // sizes, a staged bin (b < nstaged: their launches go through row_list)
    // holding every tile-row needs no list: its kernels take item idx as
    // tile-row idx.  The C-tile and row-by-row bins, and device sizes, always
    // read the list.
    if (!guard) {
        bool whole = false;
        for (int bb = 0; bb < nbins && bb < nstaged; bb++) whole = whole || s_hdr[3 + bb] - s_hdr[2 + bb] == n;
        if (whole) return;
    }
    for (int round = 0; round < RT_ITEMS; round++) {
        const int64_t idx = base + round * RT_THREADS + threadIdx.x;
        const int b = idx < n ? bin[idx] : 0xFF;
        int rank = 0;
        for (int bb = 0; bb < nbins; bb++) {
            const uint64_t m = __ballot(b == bb);
            if (b == bb) rank = mbcnt(m);
            if (lane == 0) s_wc[w][bb] = __popcll(m);
        }
        __syncthreads();
        if (b < nbins) {
            int64_t o = s_base[b];
            for (int q = 0; q < w; q++) o += s_wc[q][b];
            rows[o + rank] = (int32_t)idx;
        }
        __syncthreads();
        if ((int)threadIdx.x < nbins) {
            int t = 0;
            for (int q = 0; q < RT_THREADS / WAVE; q++) t += s_wc[q][threadIdx.x];
            s_base[threadIdx.x] += t;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- wave helpers

// ---------------------------------------------------------------- staged symbolic (fill + masks)
template <typename Var>
struct SymLDS {
    alignas(16) uint32_t cmask[Var::NC * 16];
    alignas(16) uint16_t amask[Var::NA * 16];  // 16-byte stores
    alignas(16) uint16_t bmask[Var::NB * 16];
    int32_t abs_[Var::NA];
    int32_t aseg[Var::NA + 1];
    uint32_t bm[WAVE];   // column bitmap
    int32_t wpre[WAVE];  // set bits before each bitmap word
    uint8_t bown[Var::NB];
    uint8_t bct[Var::NB];
};

// One wave per listed tile-row: C tile columns (ascending, like
// step1_fill_kernel), row masks (step2_symbolic_kernel), per-tile exclusive
// row offsets (step2_build_ptrRowC_kernel) and tile counts.  The distinct
// tile columns come from one wave-wide bitonic sort of the B tiles' columns.
template <typename Var>
__global__ void __launch_bounds__(Var::WAVES *WAVE)
k_sym_staged(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nrows, const int64_t *__restrict__ dlo,
             const int64_t *__restrict__ dhi, const int64_t *__restrict__ tile_ptrA, const uint16_t *__restrict__ maskA,
             const int4 *__restrict__ link, const int32_t *__restrict__ colidxB,
             const uint16_t *__restrict__ csr_maskB, const int64_t *__restrict__ tile_ptrC,
             int32_t *__restrict__ rowidxC, int32_t *__restrict__ colidxC, uint16_t *__restrict__ maskC,
             uint8_t *__restrict__ ptrRowC, int64_t *__restrict__ tile_nnzC, int32_t *__restrict__ overflow,
             int32_t *__restrict__ row_nc) {
    static_assert(Var::NC <= 256 && Var::NA <= 32 && Var::NB <= WAVE, "caps exceed the kernel's index widths");
    __shared__ SymLDS<Var> lds[Var::WAVES];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t idx = (int64_t)blockIdx.x * Var::WAVES + w;
    bin_rows(dlo, dhi, row_begin, nrows);
    if (idx >= nrows) return;
    SymLDS<Var> &s = lds[w];
    const int32_t bi = rows ? rows[row_begin + idx] : (int32_t)(row_begin + idx);
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];
    // exact layout: [cs, ce) are the row's tiles; slot layout: its slots
    const int64_t cs = tile_ptrC[bi], ce = tile_ptrC[bi + 1];
    const int na = (int)(ae - as), cap = (int)(ce - cs);
    if (na > Var::NA || cap > Var::NC) {
        if (lane == 0) {
            atomicOr(overflow, 1);
            if (row_nc) row_nc[bi] = 0;
        }
        return;
    }
    // level 2: A links and masks
    int4 lk = make_int4(0, 0, 0, 0);
    uint4 m0, m1;
    if (lane < na) {
        lk = link[as + lane];
        load_mask16(maskA, as + lane, m0, m1);
    }
    for (int i = lane; i < cap * 16; i += WAVE) s.cmask[i] = 0;
    const int incl = dpp_incl_sum(lk.y);
    const int nB = last_lane(incl);
    if (nB > Var::NB) {
        if (lane == 0) {
            atomicOr(overflow, 1);
            if (row_nc) row_nc[bi] = 0;
        }
        return;
    }
    if (lane < na) {
        s.aseg[lane + 1] = incl;
        s.abs_[lane] = lk.x;
        store_mask16(&s.amask[lane * 16], m0, m1);
    }
    if (lane == 0) s.aseg[0] = 0;
    wave_sync();
    // level 3: B tile columns and masks, one lane per B tile
    int32_t j = 0;
    if (lane < nB) {
        const int t = owner_of(s.aseg, na, lane);
        const int64_t q = (int64_t)s.abs_[t] + (lane - s.aseg[t]);
        j = colidxB[q];
        load_mask16(csr_maskB, q, m0, m1);
        s.bown[lane] = (uint8_t)t;
        store_mask16(&s.bmask[lane * 16], m0, m1);
    }
    // distinct tile columns and their ranks: set bits of a bitmap over the
    // columns' range when it spans at most 2048 columns, else one wave-wide
    // bitonic sort of (column, lane) keys (tile columns are below 2^27)
    int32_t jmin = 0, jmax = -1;
    if (nB > 0) {
        jmin = -last_lane(dpp_incl_max(lane < nB ? -j : -INT_MAX));
        jmax = last_lane(dpp_incl_max(lane < nB ? j : -1));
    }
    const bool bitmap = nB > 0 && jmax - jmin < WAVE * 32;
    int nc, rank = 0, sb = lane;
    bool first = false;
    if (bitmap) {
        s.bm[lane] = 0u;
        wave_sync();
        const int off = j - jmin;
        if (lane < nB) atomicOr(&s.bm[off >> 5], 1u << (off & 31));
        wave_sync();
        const uint32_t word = s.bm[lane];
        const int pc = __popc(word);
        const int incl = dpp_incl_sum(pc);
        nc = last_lane(incl);
        s.wpre[lane] = incl - pc;
        wave_sync();
        if (lane < nB) rank = s.wpre[off >> 5] + __popc(s.bm[off >> 5] & ((1u << (off & 31)) - 1u));
        first = lane < nB;  // every B tile writes its column at its rank (duplicates agree)
    } else {
        const uint64_t key = wave_sort64(lane < nB ? ((uint64_t)(uint32_t)j << 6) | (uint64_t)lane : ~0ull);
        j = (int32_t)(key >> 6);
        sb = (int)(key & 63);
        const int32_t prev = __shfl_up(j, 1, WAVE);
        first = lane < nB && (lane == 0 || prev != j);
        const uint64_t fb = __ballot(first);
        nc = __popcll(fb);
        rank = mbcnt(fb) + (first ? 1 : 0) - 1;
    }
    if (nc > cap) {  // slot layout: the prediction was short for this row
        if (lane == 0) {
            atomicOr(overflow, 1);
            if (row_nc) row_nc[bi] = 0;
        }
        return;
    }
    if (row_nc && lane == 0) row_nc[bi] = nc;
    if (first) {
        colidxC[cs + rank] = j;
        rowidxC[cs + rank] = bi;
    }
    if (lane < nB) s.bct[sb] = (uint8_t)rank;
    wave_sync();
    // masks: item (B tile b, row r) ORs B's rows selected by A's row mask
    for (int it = lane; it < nB * 16; it += WAVE) {
        const int b = it >> 4, r = it & 15;
        uint32_t ma = s.amask[s.bown[b] * 16 + r];
        uint32_t rm = 0;
        while (ma) {
            const int kk = __clz(ma) - 16;
            ma ^= 0x8000u >> kk;
            rm |= s.bmask[b * 16 + kk];
        }
        if (rm) atomicOr(&s.cmask[s.bct[b] * 16 + r], rm);
    }
    wave_sync();
    // write: lane ct writes C tile ct whole -- its 16 row masks (two 16-byte
    // stores), their exclusive row offsets (one) and its count
    for (int ct = lane; ct < nc; ct += WAVE) {
        uint32_t mw[8], pw[4] = {0u, 0u, 0u, 0u}, acc = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 m = reinterpret_cast<const uint4 *>(&s.cmask[ct * 16])[q];
            mw[2 * q] = (m.x & 0xFFFFu) | (m.y << 16);
            mw[2 * q + 1] = (m.z & 0xFFFFu) | (m.w << 16);
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            pw[i >> 2] |= acc << (8 * (i & 3));
            acc += __popc((mw[i >> 1] >> (16 * (i & 1))) & 0xFFFFu);
        }
        uint4 *mo = reinterpret_cast<uint4 *>(maskC + (cs + ct) * 16);
        mo[0] = make_uint4(mw[0], mw[1], mw[2], mw[3]);
        mo[1] = make_uint4(mw[4], mw[5], mw[6], mw[7]);
        reinterpret_cast<uint4 *>(ptrRowC + (cs + ct) * 16)[0] = make_uint4(pw[0], pw[1], pw[2], pw[3]);
        tile_nnzC[cs + ct + 1] = acc;
    }
}

// ---------------------------------------------------------------- staged numeric
template <typename Var>
struct NumLDS {
    int64_t bgoff[Var::NB];
    int64_t coff[Var::NC + 1];
    int32_t cj[Var::NC];
    int32_t abs_[Var::NA];
    int32_t aoff[Var::NA + 1];
    int32_t aseg[Var::NA + 1];
    uint32_t tmask[Var::NC];
    int16_t pstart[Var::NC + 1];
    uint16_t desc[DESC_CAP];
    alignas(16) uint16_t cmask[Var::NC * 16];  // 16-byte stores
    alignas(16) uint16_t amask[Var::NA * 16];
    alignas(16) uint16_t bcmask[Var::NB * 16];  // B column masks
    alignas(16) uint8_t aptr[Var::NA * 16];
    alignas(16) uint8_t bcptr[Var::NB * 16];    // B column offsets
    uint8_t bown[Var::NB];
    uint16_t plist[Var::NB];
};

template <typename V>
__device__ __forceinline__ V fma_v(V a, V b, V c);
template <>
__device__ __forceinline__ double fma_v<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_v<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// A's and B's tile headers are staged in LDS; the values of both are read
// from memory in the output loop, four products of an output at a time
// (values reused across outputs come from the caches), which keeps LDS per
// wave small and occupancy high: with A's values staged too (4 KB a wave
// for 512 FP64 values) a 27-point stencil times an aggregation ran 1.33x
// slower.  B's tiles come column by column (column masks, offsets, values),
// so for output (r, c) and pair (A tile, B tile) the contributing k are
// exactly the set bits of (A row mask r) & (B column mask c): the loop runs
// once per product.
template <typename V, typename Var>
__global__ void __launch_bounds__(Var::WAVES *WAVE)
k_num_staged(const int32_t *__restrict__ rows, int64_t row_begin, int64_t nrows, const int64_t *__restrict__ dlo,
             const int64_t *__restrict__ dhi, const int64_t *__restrict__ tile_ptrA,
             const int64_t *__restrict__ tile_nnzA,
             const uint8_t *__restrict__ ptrA, const uint16_t *__restrict__ maskA, const V *__restrict__ valA,
             const int4 *__restrict__ link, const int32_t *__restrict__ colidxB,
             const int64_t *__restrict__ csr_nnzB, const uint8_t *__restrict__ cptrB,
             const uint16_t *__restrict__ cmaskB, const V *__restrict__ cvalB,
             const int64_t *__restrict__ tile_ptrC, const int32_t *__restrict__ colidxC,
             const int64_t *__restrict__ tile_nnzC, const uint16_t *__restrict__ maskC, V *__restrict__ valC,
             uint8_t *__restrict__ colC, int32_t *__restrict__ overflow, const int32_t *__restrict__ row_nc,
             int64_t val_cap, int64_t val_done) {
    static_assert(Var::NC < WAVE && Var::NA <= 32 && Var::NB <= WAVE, "caps exceed the kernel's index widths");
    __shared__ NumLDS<Var> lds[Var::WAVES];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t idx = (int64_t)blockIdx.x * Var::WAVES + w;
    bin_rows(dlo, dhi, row_begin, nrows);
    if (idx >= nrows) return;
    NumLDS<Var> &s = lds[w];
    const int32_t bi = rows ? rows[row_begin + idx] : (int32_t)(row_begin + idx);
    // level 1: tile-row pointers (wave-uniform scalars)
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];
    int64_t cs, ce;
    c_row_range(tile_ptrC, row_nc, bi, cs, ce);
    if (val_done >= 0 && tile_nnzC[ce] <= val_done) return;  // a re-run: this tile-row's values are written
    const int na = (int)(ae - as), nc = (int)(ce - cs);
    if (na > Var::NA || nc > Var::NC) {
        if (lane == 0) atomicOr(overflow, 1);
        return;
    }
    // level 2: A headers and links, C headers (all loads issued, then stored)
    const int64_t abase = tile_nnzA[as];
    const int pA = (int)(tile_nnzA[ae] - abase);
    int32_t anz = 0, cjv = 0;
    int64_t co = 0;
    int4 lk = make_int4(0, 0, 0, 0);
    uint4 ap, am0, am1, cm0, cm1;
    if (lane < na) {
        anz = (int32_t)(tile_nnzA[as + lane] - abase);
        lk = link[as + lane];
        ap = load_ptr16(ptrA, as + lane);
        load_mask16(maskA, as + lane, am0, am1);
    }
    if (lane <= nc) co = tile_nnzC[cs + lane];
    if (lane < nc) {
        cjv = colidxC[cs + lane];
        load_mask16(maskC, cs + lane, cm0, cm1);
    }
    const int incl = dpp_incl_sum(lk.y);
    const int nB = last_lane(incl);
    const int64_t cend = __builtin_amdgcn_readlane((int)(co >> 32), nc) * 4294967296ll +
                         (uint32_t)__builtin_amdgcn_readlane((int)co, nc);  // end of the row's values
    if (nB > Var::NB || cend > val_cap) {
        if (lane == 0) atomicOr(overflow, 1);
        return;
    }
    if (lane < na) {
        s.aoff[lane] = anz;
        s.aseg[lane + 1] = incl;
        s.abs_[lane] = lk.x;
        reinterpret_cast<uint4 *>(s.aptr)[lane] = ap;
        store_mask16(&s.amask[lane * 16], am0, am1);
    }
    if (lane == 0) {
        s.aoff[na] = pA;
        s.aseg[0] = 0;
    }
    if (lane <= nc) s.coff[lane] = co;
    if (lane < nc) {
        s.cj[lane] = cjv;
        s.tmask[lane] = 0;
        store_mask16(&s.cmask[lane * 16], cm0, cm1);
    }
    wave_sync();
    // level 3: B headers, one lane per B tile
    int bt = 0, ct_b = 0;
    if (lane < nB) {
        bt = owner_of(s.aseg, na, lane);
        const int64_t q = (int64_t)s.abs_[bt] + (lane - s.aseg[bt]);
        const int32_t bjv = colidxB[q];
        const int64_t bg = csr_nnzB[q];
        uint4 bm0, bm1;
        const uint4 bp = load_ptr16(cptrB, q);
        load_mask16(cmaskB, q, bm0, bm1);
        int lo = 0, hi = nc;  // C tile of this B tile
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s.cj[mid] < bjv) lo = mid + 1; else hi = mid;
        }
        ct_b = lo;
        s.bgoff[lane] = bg;
        s.bown[lane] = (uint8_t)bt;
        reinterpret_cast<uint4 *>(s.bcptr)[lane] = bp;
        store_mask16(&s.bcmask[lane * 16], bm0, bm1);
        atomicOr(&s.tmask[ct_b], 1u << bt);
    }
    wave_sync();
    // pair table: the B tiles of C tile ct in ascending k = set bits of tmask[ct];
    // an entry holds the B tile and its A tile (b | t << 8)
    const int pinc = dpp_incl_sum(lane < nc ? __popc(s.tmask[lane]) : 0);
    if (lane < nc) s.pstart[lane + 1] = (int16_t)pinc;
    if (lane == 0) s.pstart[0] = 0;
    wave_sync();
    if (lane < nB) s.plist[s.pstart[ct_b] + __popc(s.tmask[ct_b] & ((1u << bt) - 1u))] = (uint16_t)(lane | (bt << 8));
    wave_sync();
    // outputs in chunks of at most DESC_CAP: the descriptors (C tile, row,
    // column) of a chunk are written 4 C tiles (64 items, lane = item) at a
    // time, then lane = output.  A whole tile-row usually fits one chunk.
    for (int ct0 = 0; ct0 < nc;) {
        int ct1 = ct0, nout = 0;
        while (ct1 < nc) {
            int span = min(4, nc - ct1);
            int blk = (int)(s.coff[ct1 + span] - s.coff[ct1]);
            if (nout + blk > DESC_CAP) {
                if (nout > 0) break;
                while (span > 1 && blk > DESC_CAP) {  // one tile holds at most 256 = DESC_CAP outputs
                    span >>= 1;
                    blk = (int)(s.coff[ct1 + span] - s.coff[ct1]);
                }
            }
            const int item = ct1 * 16 + lane;
            const uint32_t mc = (lane < span * 16) ? s.cmask[item] : 0u;
            const int pc = __popc(mc);
            const int oincl = dpp_incl_sum(pc);
            int o = nout + oincl - pc;
            uint32_t m = mc;
            while (m) {
                const int c = __clz(m) - 16;
                m ^= 0x8000u >> c;
                s.desc[o++] = (uint16_t)((item << 4) | c);  // (C tile, row, column)
            }
            nout += blk;
            ct1 += span;
        }
        wave_sync();
        const int64_t obase = s.coff[ct0];
        // results stay in registers until the chunk is done: the stores issue
        // together after the last output (no store in flight while computing)
        V res[DESC_CAP / WAVE];
        uint8_t cb[DESC_CAP / WAVE];
#pragma unroll
        for (int it = 0; it < DESC_CAP / WAVE; it++) {
            const int o = lane + it * WAVE;
            res[it] = V(0);
            cb[it] = 0;
            if (it * WAVE < nout && o < nout) {
                const int d = s.desc[o];
                const int ct = d >> 8, r = (d >> 4) & 15, c = d & 15;
                V acc = V(0);
                int p = s.pstart[ct];
                const int pe = s.pstart[ct + 1];
                uint32_t ma = 0, cm = 0, m = 0;  // the current pair's masks; m: its k not yet taken
                int ea = 0;
                const V *bv = cvalB;
                // the output's products NS_K at a time: their A and B values
                // are loaded together, then added in order (pairs ascending,
                // k ascending within a pair)
                do {
                    int ai[NS_K];
                    const V *bp[NS_K];
                    int nk = 0;
#pragma unroll
                    for (int k = 0; k < NS_K; k++) {
                        while (!m && p < pe) {
                            const int e = s.plist[p++];
                            const int b = e & 0xFF, t = e >> 8;
                            ma = s.amask[t * 16 + r];
                            cm = s.bcmask[b * 16 + c];
                            m = ma & cm;  // k of the products, ascending from the top bit
                            ea = s.aoff[t] + s.aptr[t * 16 + r];
                            bv = cvalB + (s.bgoff[b] + s.bcptr[b * 16 + c]);
                        }
                        ai[k] = 0;
                        bp[k] = bv;
                        if (m) {
                            // top set bit hb: k = 15 - hb; the entries before it
                            // in A's row and B's column are the set bits above hb
                            const int hb = 31 - __builtin_clz(m);
                            m ^= 1u << hb;
                            ai[k] = ea + __popc(ma >> (hb + 1));
                            bp[k] = bv + __popc(cm >> (hb + 1));
                            nk = k + 1;
                        }
                    }
                    V aq[NS_K], bq[NS_K];
#pragma unroll
                    for (int k = 0; k < NS_K; k++) {
                        aq[k] = k < nk ? valA[abase + ai[k]] : V(0);
                        bq[k] = k < nk ? *bp[k] : V(0);
                    }
#pragma unroll
                    for (int k = 0; k < NS_K; k++)
                        if (k < nk) acc = fma_v<V>(aq[k], bq[k], acc);
                } while (m || p < pe);
                res[it] = acc;
                cb[it] = (uint8_t)((r << 4) | c);
            }
        }
#pragma unroll
        for (int it = 0; it < DESC_CAP / WAVE; it++) {
            const int o = lane + it * WAVE;
            if (o < nout) {
                valC[obase + o] = res[it];
                colC[obase + o] = cb[it];
            }
        }
        wave_sync();
        ct0 = ct1;
    }
}

// ---------------------------------------------------------------- host side
static bool binned_variant(const tsg_options *opts) { return !opts || opts->numeric_variant == 0; }

bool staged_ok(const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts) {
    if (!binned_variant(opts)) return false;  // C-tile kernels everywhere
    return B->csr_tile_nnz && B->csr_mask && B->csr_col_mask && B->csr_col_Ptr && B->csr_col_Val && B->row_nnz_ptr &&
           !(A->flags & 1) && !(B->flags & 1);
}

bool rg_ok(const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts) {
    if (!binned_variant(opts)) return false;  // C-tile kernels everywhere
    return A->row_ptr && A->row_col && A->row_val && B->row_ptr && B->row_col && B->row_val && !(A->flags & 1) &&
           !(B->flags & 1);
}

BinTable bin_table_for(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_options *opts) {
    TuneCfg cfg;
    tuned_config(ctx, A, B, opts, &cfg);
    // the row-by-row bin needs a per-workgroup cursor area far larger than the
    // no-alloc scratch: inside such a call heavy rows take the C-tile kernels
    // (every path computes the same bits)
    BinTable t = bin_table(staged_ok(A, B, opts), rg_ok(A, B, opts) && !ctx->arena, cfg.rg_density, cfg.var_mask,
                           num_rows_ok(A, B));
    t.any_order = (opts && opts->sum_order == 1) ? 1 : 0;
    return t;
}

BinTable bin_table(bool ok, bool rg, int rg_density, int var_mask, bool row_lanes) {
    BinTable t{};
    t.nvar = 0;
    t.rg = (rg && rg_density > 0) ? 1 : 0;
    t.rg_density = rg_density;
    t.rl_any = (var_mask & 16) ? 1 : 0;
    if (ok) {
        // bins in first-fit order R, S, W, M; var_mask bits 0-2 = S, W, M, bit 3 = R
        const StageCaps none{-1, -1, -1, -1, -1};
        t.v[0] = (row_lanes && (var_mask & 8)) ? CAPS_R : none;
        t.v[1] = (var_mask & 1) ? caps_of<VarS>() : none;
        t.v[2] = (var_mask & 2) ? caps_of<VarW>() : none;
        t.v[3] = (var_mask & 4) ? caps_of<VarM>() : none;
        t.nvar = NUM_STAGED;
    }
    return t;
}

void free_heavy_bitmaps(Context *ctx, Plan &p) {
    pool_free(ctx, p.hb.bwords);
    pool_free(ctx, p.hb.boff);
    pool_free(ctx, p.hb.words);
    pool_free(ctx, p.hb.top);
    pool_free(ctx, p.hb.off);
    p.hb = HeavyBitmaps{};
}

void free_plan(Context *ctx, Plan &p) {
    free_heavy_bitmaps(ctx, p);
    void *ps[] = {p.rows, p.bin, p.counts, p.hdr, p.flags, p.link, p.lpay, p.rgw, p.rgn, p.rgc};
    for (void *q : ps) pool_free(ctx, q);
    p = Plan{};
}

static unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

int plan_setup(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const BinTable &bt, Plan &p,
               bool zero_flags) {
    const int64_t tilem = A->tilem;
    const int64_t nb = (tilem + BIN_TILE - 1) / BIN_TILE;
    const bool ok = bt.nvar > 0;
    p.nvar = bt.nvar;
    p.rg = bt.rg != 0;
    p.rl_any = bt.rl_any != 0;
    p.any_order = bt.any_order != 0;
    p.nbins = bt.nvar + 1 + (p.rg ? 1 : 0);
    TSG_CHECK(palloc(ctx, &p.rows, (size_t)tilem));
    TSG_CHECK(palloc(ctx, &p.bin, (size_t)tilem));
    TSG_CHECK(palloc(ctx, &p.counts, (size_t)(p.nbins * nb + 1)));
    TSG_CHECK(palloc(ctx, &p.hdr, MAX_BINS + 3));
    TSG_CHECK(pool_alloc(ctx, reinterpret_cast<void **>(&p.flags), sizeof(PlanFlags)));
    TSG_CHECK(palloc(ctx, &p.link, (size_t)A->numtile));
    TSG_CHECK(palloc(ctx, &p.lpay, (size_t)A->numtile));
    if (p.rg) {
        TSG_CHECK(palloc(ctx, &p.rgw, (size_t)tilem + 4));  // offsets, 2 work counters, numeric item size
        TSG_CHECK(palloc(ctx, &p.rgn, (size_t)tilem + 1));
    }
    if (zero_flags) TSG_CHECK(zero_small(ctx, p.overflow(), PLAN_FLAG_WORDS, nullptr, 0));
    if (A->numtile > 0) {
        k_ab_link<<<nblk(A->numtile, 256), 256, 0, ctx->stream>>>(A->numtile, A->tile_columnidx, B->tile_ptr,
                                                                   ok ? B->row_nnz_ptr : nullptr, p.link, p.lpay,
                                                                   p.overflow(), PLAN_FLAG_WORDS);
        TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

int plan_bin(Context *ctx, Plan &p, int32_t tilem, const int64_t *total0, const int64_t *total1) {
    hipStream_t s = ctx->stream;
    const int64_t nb = ((int64_t)tilem + BIN_TILE - 1) / BIN_TILE;
    if (tilem > 0) {
        k_bin_count<<<(unsigned)nb, BIN_THREADS, 0, s>>>(p.bin, tilem, p.nbins, nb, p.counts);
        TSG_HIP(hipGetLastError());
        TSG_CHECK(scan_inclusive_i64(ctx, p.counts, p.nbins * nb + 1));
        k_bin_scatter<<<(unsigned)nb, BIN_THREADS, 0, s>>>(p.bin, tilem, p.nbins, nb, p.counts, p.rows, total0,
                                                            total1, p.hdr);
    } else {
        TSG_HIP(hipMemsetAsync(p.counts, 0, sizeof(int64_t), s));
        k_bin_header<<<1, 64, 0, s>>>(p.counts, p.nbins, nb, total0, total1, p.hdr);
    }
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int plan_read(Context *ctx, Plan &p, int64_t *total0, int64_t *total1) {
    TSG_CHECK(read_words(ctx, p.hdr, 2 * (p.nbins + 3), ctx->pinned));
    if (total0) *total0 = ctx->pinned[0];
    if (total1) *total1 = ctx->pinned[1];
    for (int b = 0; b <= p.nbins; b++) p.bin_off[b] = ctx->pinned[2 + b];
    return TSG_OK;
}

// A->B links, Step 1 count, statistics, exclusive scan of tile_ptrC.
int plan_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
               const tsg_options *opts, Plan &p, int32_t *stats_out, bool scan) {
    const BinTable bt = bin_table_for(ctx, A, B, opts);
    const int32_t tilem = A->tilem;
    TSG_CHECK(plan_setup(ctx, A, B, bt, p, tilem == 0));
    if (tilem == 0) {
        TSG_HIP(hipMemsetAsync(C->tile_ptr, 0, sizeof(int64_t), ctx->stream));
    } else {
        // heavy rows (counted by a workgroup each); hc = {count, pad, 64-bit work counter}
        int32_t *hlist = nullptr, *hc = &p.flags->heavy_rows;
        // a heavy row reaches more than HEAVY_NB B tiles, so one operand has a
        // tile-row of more than HEAVY_LIST tiles: none when tsg_csr2tile
        // checked both (flags bit 2) and found none (bit 1)
        const bool no_heavy = (A->flags & 4) && (B->flags & 4) && !((A->flags | B->flags) & 2);
        int rc = no_heavy ? TSG_OK : palloc(ctx, &hlist, (size_t)tilem);
        // keep union bitmaps only where long tile-rows are known (without them
        // the fill just rebuilds every bitmap)
        if (!rc && !ctx->arena && B->numtile > HEAVY_NB && ((A->flags | B->flags) & 2)) {
            p.hb.cap = std::min<int64_t>(HEAVY_BITMAP_WORDS, (int64_t)tilem * ((int64_t)B->tilen / 32 + 1));
            rc = palloc(ctx, &p.hb.words, (size_t)p.hb.cap);
            if (!rc) rc = palloc(ctx, &p.hb.off, (size_t)tilem);
            if (!rc) rc = palloc(ctx, reinterpret_cast<int64_t **>(&p.hb.top), 1);
            if (!rc && (B->flags & 2)) rc = build_brow_bitmaps(ctx, B, p.hb);  // long B tile-rows exist
        }
        // k_ab_link cleared the flags (it runs when A has tiles)
        if (!rc)
            rc = zero_small(ctx, A->numtile > 0 ? nullptr : p.overflow(), PLAN_FLAG_WORDS,
                            reinterpret_cast<int64_t *>(p.hb.top), p.hb.top ? 1 : 0);
        // light tile-rows at 8 lanes each; the rest (marked) a wave each
        if (!rc) {
            k_row_light<<<(unsigned)((tilem + RL_GROUPS - 1) / RL_GROUPS), WAVE, 0, ctx->stream>>>(
                A->tile_ptr, A->tile_nnz, A->tilem, p.link, B->tile_columnidx, bt, C->tile_ptr, p.bin, stats_out);
            if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_row_light launch");
        }
        if (!rc) {
            k_row_stats<<<nblk((tilem + RS_CHUNK - 1) / RS_CHUNK, RS_WAVES), RS_WAVES * WAVE, 0, ctx->stream>>>(
                A->tile_ptr, A->tile_nnz, A->tilem, p.link, B->tile_columnidx, bt, C->tile_ptr, p.bin, stats_out,
                hlist, hc);
            if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_row_stats launch");
        }
        if (!rc && hlist) rc = launch_heavy_count(ctx, A, B, C->tile_ptr, hlist, hc, stats_out, p.hb);
        pool_free(ctx, hlist);
        TSG_CHECK(rc);
        if (scan) TSG_CHECK(scan_inclusive_i64(ctx, C->tile_ptr, (int64_t)tilem + 1));
    }
    return TSG_OK;
}

bool plan_tail_fused(int32_t tilem) {
    return tilem > 0 && ((int64_t)tilem + RT_TILE - 1) / RT_TILE <= RT_MAX_BLOCKS;
}

// The fused tail's two launches; with publish, block 0 of the second also
// sends the header to the host (seq: its number).
static int rows_tail_launch(Context *ctx, Plan &p, int32_t tilem, int64_t *tile_ptrC, const int64_t *total1,
                            bool publish, int64_t *seq, const BinCaps *caps = nullptr, int64_t cap_tiles = 0,
                            int64_t *numtile_out = nullptr) {
    hipStream_t s = ctx->stream;
    const int64_t nb = ((int64_t)tilem + RT_TILE - 1) / RT_TILE;
    int64_t *partial = nullptr;
    int32_t *bc = nullptr;
    int rc = palloc(ctx, &partial, (size_t)nb);
    if (!rc) rc = palloc(ctx, &bc, (size_t)(nb * p.nbins));
    if (!rc) {
        k_rows_reduce<<<(unsigned)nb, RT_THREADS, 0, s>>>(tile_ptrC, p.bin, tilem, p.nbins, nb, partial, bc);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_rows_reduce launch");
    }
    if (!rc) {
        *seq = publish ? publish_seq(ctx) : 0;
        k_rows_apply<<<(unsigned)nb, RT_THREADS, 0, s>>>(tile_ptrC, p.bin, tilem, p.nbins, nb, partial, bc, p.rows,
                                                          total1, p.hdr, publish ? ctx->pub_dev : nullptr, *seq,
                                                          caps != nullptr, caps ? *caps : BinCaps{}, cap_tiles,
                                                          p.overflow(), numtile_out, p.nvar);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_rows_apply launch");
    }
    // stream-ordered reuse: later work on this stream runs after the kernels above
    pool_free(ctx, partial);
    pool_free(ctx, bc);
    return rc;
}

__global__ void k_device_guard(int64_t *__restrict__ hdr, int nbins, int64_t cap_tiles, BinCaps caps,
                               int32_t *__restrict__ overflow);

int plan_tail_device(Context *ctx, Plan &p, int32_t tilem, int64_t *tile_ptrC, const int64_t *total1,
                     const BinCaps *caps, int64_t cap_tiles, int64_t *numtile_out) {
    if (!plan_tail_fused(tilem)) {
        TSG_CHECK(scan_inclusive_i64(ctx, tile_ptrC, (int64_t)tilem + 1));
        TSG_CHECK(plan_bin(ctx, p, tilem, tile_ptrC + tilem, total1));
        if (caps) {
            k_device_guard<<<1, 64, 0, ctx->stream>>>(p.hdr, p.nbins, cap_tiles, *caps, p.overflow());
            TSG_HIP(hipGetLastError());
            if (numtile_out)
                TSG_HIP(hipMemcpyAsync(numtile_out, p.hdr, sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream));
        }
        return TSG_OK;
    }
    int64_t seq = 0;
    return rows_tail_launch(ctx, p, tilem, tile_ptrC, total1, false, &seq, caps, cap_tiles, numtile_out);
}

int plan_tail(Context *ctx, Plan &p, int32_t tilem, int64_t *tile_ptrC, const int64_t *total1, int64_t *numtile,
              int64_t *t1) {
    if (!plan_tail_fused(tilem)) {
        TSG_CHECK(scan_inclusive_i64(ctx, tile_ptrC, (int64_t)tilem + 1));
        TSG_CHECK(plan_bin(ctx, p, tilem, tile_ptrC + tilem, total1));
        return plan_read(ctx, p, numtile, t1);
    }
    int64_t seq = 0;
    TSG_CHECK(rows_tail_launch(ctx, p, tilem, tile_ptrC, total1, true, &seq));
    TSG_CHECK(publish_wait(ctx, seq, 2 * (p.nbins + 3), ctx->pinned));
    if (numtile) *numtile = ctx->pinned[0];
    if (t1) *t1 = ctx->pinned[1];
    for (int b = 0; b <= p.nbins; b++) p.bin_off[b] = ctx->pinned[2 + b];
    return TSG_OK;
}

// plan_count, binning and one host read of (numtileC, bin offsets).
static int plan_rows(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                     const tsg_options *opts, Plan &p) {
    TSG_CHECK(plan_count(ctx, A, B, C, opts, p, nullptr, !plan_tail_fused(A->tilem)));
    return plan_tail(ctx, p, A->tilem, C->tile_ptr, nullptr, &C->numtile, nullptr);
}

// Staged kernels over the rows of bins [bin, bin_end) (adjacent in the row list).
template <typename Var>
static int launch_sym_staged(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                             const tsg_tile_matrix *C, const Plan &p, int bin, const SlotLayout &sl,
                             int bin_end = -1) {
    const BinRange br = bin_range(p, bin, bin_end < 0 ? bin + 1 : bin_end);
    if (br.count <= 0) return TSG_OK;
    k_sym_staged<Var><<<nblk(br.count, Var::WAVES), Var::WAVES * WAVE, 0, ctx->stream>>>(
        row_list(p.rows, br, C->tilem), br.begin, br.count, br.lo, br.hi,
        A->tile_ptr, A->mask, p.link, B->tile_columnidx, B->csr_mask, C->tile_ptr,
        C->tile_rowidx, C->tile_columnidx, C->mask, C->tile_csr_Ptr, C->tile_nnz, p.overflow(), sl.row_nc);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

template <typename V, typename Var>
static int launch_num_staged(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                             const tsg_tile_matrix *C, const Plan &p, int bin, const SlotLayout &sl,
                             int bin_end = -1) {
    const BinRange br = bin_range(p, bin, bin_end < 0 ? bin + 1 : bin_end);
    if (br.count <= 0) return TSG_OK;
    k_num_staged<V, Var><<<nblk(br.count, Var::WAVES), Var::WAVES * WAVE, 0, ctx->stream>>>(
        row_list(p.rows, br, C->tilem), br.begin, br.count, br.lo, br.hi,
        A->tile_ptr, A->tile_nnz, A->tile_csr_Ptr, A->mask, (const V *)A->tile_csr_Val,
        p.link, B->tile_columnidx, B->csr_tile_nnz, B->csr_col_Ptr, B->csr_col_mask,
        (const V *)B->csr_col_Val, C->tile_ptr, C->tile_columnidx, C->tile_nnz, C->mask, (V *)C->tile_csr_Val,
        C->tile_csr_Col, p.overflow(), sl.row_nc, sl.val_cap, sl.val_done);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

// Whether bin R's rows run the row-lane kernels: many of them (a launch of
// their own pays), var_mask bit 4, or a captured graph (sizes unknown); else
// they join S's launches (they fit VarS's caps).
static bool rows_own(const Plan &p) {
    return p.d_bins || p.rl_any || p.bin_off[1] - p.bin_off[0] >= ROW_LANES_MIN;
}

// The rows of bins [b, e): host offsets, or in a captured graph the device
// offsets with the bins' capacities as the count the launches cover.
BinRange bin_range(const Plan &p, int b, int e) {
    BinRange r;
    if (p.d_bins) {
        for (int q = b; q < e; q++) r.count += p.bin_cap[q];
        r.lo = p.d_bins + b;
        r.hi = p.d_bins + e;
        r.all = p.all_bin >= b && p.all_bin < e;
    } else {
        r.begin = p.bin_off[b];
        r.count = p.bin_off[e] - p.bin_off[b];
    }
    return r;
}

// Structure fill + symbolic for every bin.

int run_symbolic(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                 Plan &p, const SlotLayout &sl) {
    const int fb = p.nvar;  // C-tile kernels
    if (p.nvar > 0) {
        const bool rn = rows_own(p);
        if (rn) TSG_CHECK(launch_sym_rows(ctx, A, B, C, p.rows, bin_range(p, 0, 1), p.link, sl));
        TSG_CHECK(launch_sym_staged<VarS>(ctx, A, B, C, p, rn ? 1 : 0, sl, 2));
        TSG_CHECK(launch_sym_staged<VarW>(ctx, A, B, C, p, 2, sl));
        TSG_CHECK(launch_sym_staged<VarM>(ctx, A, B, C, p, 3, sl));
    }
    const BinRange f = bin_range(p, fb, fb + 1);
    TSG_CHECK(launch_structure_fill_list(ctx, A, B, C, p.rows, f, sl, p.hb));
    TSG_CHECK(launch_symbolic_list(ctx, A, B, p.rows, f, C, sl));
    if (p.rg) {  // row-by-row kernels: Step 1 fill as the C-tile rows, then masks from the products
        const BinRange r = bin_range(p, fb + 1, fb + 2);
        TSG_CHECK(launch_structure_fill_list(ctx, A, B, C, p.rows, r, sl, p.hb));
        TSG_CHECK(launch_rg_symbolic(ctx, A, B, C, p.rows, r, p.rgw, &p.rgc, sl));
    }
    return TSG_OK;
}

// The numeric kernels run_numeric launches for host-known bins (tsg_stats.numeric_kernels).
int numeric_kernel_mask(const Plan &p) {
    if (p.d_bins) return 0;
    auto rows = [&](int b, int e) { return p.bin_off[e] - p.bin_off[b]; };
    int m = 0;
    const int fb = p.nvar;
    if (p.nvar > 0) {
        const bool rn = rows_own(p);
        if (rn && rows(0, 1) > 0) m |= 1;
        if (rows(rn ? 1 : 0, 2) > 0) m |= 2;
        if (rows(2, 3) > 0) m |= 4;
        if (rows(3, 4) > 0) m |= 8;
    }
    if (rows(fb, fb + 1) > 0) m |= 16;
    if (p.rg && rows(fb + 1, fb + 2) > 0) m |= 32;
    return m;
}

int run_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                const Plan &p, const SlotLayout &sl) {
    const int fb = p.nvar;
    if (p.nvar > 0) {
        const bool rn = rows_own(p);
        if (rn) TSG_CHECK(launch_num_rows(ctx, A, B, C, p.rows, bin_range(p, 0, 1), sl));
        const int s0 = rn ? 1 : 0;
        if (A->dtype == TSG_F32) {
            TSG_CHECK((launch_num_staged<float, VarS>(ctx, A, B, C, p, s0, sl, 2)));
            TSG_CHECK((launch_num_staged<float, VarW>(ctx, A, B, C, p, 2, sl)));
            TSG_CHECK((launch_num_staged<float, VarM>(ctx, A, B, C, p, 3, sl)));
        } else {
            TSG_CHECK((launch_num_staged<double, VarS>(ctx, A, B, C, p, s0, sl, 2)));
            TSG_CHECK((launch_num_staged<double, VarW>(ctx, A, B, C, p, 2, sl)));
            TSG_CHECK((launch_num_staged<double, VarM>(ctx, A, B, C, p, 3, sl)));
        }
    }
    TSG_CHECK(launch_numeric_list(ctx, A, B, C, p.rows, bin_range(p, fb, fb + 1), sl));
    if (p.rg)
        TSG_CHECK(launch_rg_numeric(ctx, A, B, C, p.rows, bin_range(p, fb + 1, fb + 2), p.rgw, p.rgn, p.rgc, sl,
                                    p.any_order));
    return TSG_OK;
}

// Numeric set-up that needs C's scanned value offsets and must stay out of
// the numeric launches (which a plan captures in a graph): the row-by-row
// bin's work items.
int prepare_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *C, Plan &p, const SlotLayout &sl) {
    if (!p.rg) return TSG_OK;
    const int fb = p.nvar;
    return prepare_rg_numeric(ctx, A, C, p.rows, bin_range(p, fb + 1, fb + 2), p.rgw, p.rgn, &p.rgc,
                              p.flags ? &p.flags->nnz : nullptr, sl);
}

void free_c(Context *ctx, tsg_tile_matrix *C) {
    void *ps[] = {C->tile_ptr, C->tile_columnidx, C->tile_rowidx, C->tile_nnz, C->tile_csr_Val, C->tile_csr_Col,
                  C->tile_csr_Ptr, C->mask};
    for (void *q : ps) pool_free(ctx, q);
    memset(C, 0, sizeof *C);
}

// The exact binned pipeline: the reference's three steps (exact counts and
// host reads of numtileC and nnzC, src/spgemm_raw_v2.cu:141, :211), with rows
// binned to the staged kernels or the C-tile kernels.  The staged kernels'
// capacity flag is read with nnzC's sync.
int spgemm_binned(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                  const tsg_options *opts, tsg_stats *st, Plan *keep) {
    const bool timing = opts && opts->timing;
    hipEvent_t ev[4] = {};
    if (timing)
        for (auto &e : ev) TSG_HIP(hipEventCreate(&e));
    memset(C, 0, sizeof *C);
    C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
    C->dtype = A->dtype; C->location = TSG_DEVICE;
    Plan p;
    hipStream_t s = ctx->stream;
    int rc = TSG_OK;
    int syncs = 0;
    if (timing) (void)hipEventRecord(ev[0], s);
    const FixedOut *fo = ctx->fixed_out;  // caller-owned C arrays (no-alloc entry point)
    if (fo) C->tile_ptr = fo->c.tile_ptr;
    else rc = palloc(ctx, &C->tile_ptr, (size_t)C->tilem + 1);
    if (!rc) { rc = plan_rows(ctx, A, B, C, opts, p); syncs++; }
    if (!rc && C->numtile > INT32_MAX) rc = fail(TSG_ERR_OVERFLOW, "numtileC exceeds int32 tile ids");
    if (!rc && fo) {
        if (C->numtile > fo->cap_tiles) {
            rc = fail(TSG_ERR_OVERFLOW, "C has %lld tiles; the caller's tile arrays hold %lld", (long long)C->numtile,
                      (long long)fo->cap_tiles);
        } else {
            C->tile_rowidx = fo->c.tile_rowidx;
            C->tile_columnidx = fo->c.tile_columnidx;
            C->mask = fo->c.mask;
            C->tile_csr_Ptr = fo->c.tile_csr_Ptr;
            C->tile_nnz = fo->c.tile_nnz;
        }
    } else {
        if (!rc) rc = palloc(ctx, &C->tile_rowidx, (size_t)C->numtile);
        if (!rc) rc = palloc(ctx, &C->tile_columnidx, (size_t)C->numtile);
        if (!rc) rc = palloc(ctx, &C->mask, (size_t)C->numtile * 16);
        if (!rc) rc = palloc(ctx, &C->tile_csr_Ptr, (size_t)C->numtile * 16);
        if (!rc) rc = palloc(ctx, &C->tile_nnz, (size_t)C->numtile + 1);
    }
    if (timing && !rc) (void)hipEventRecord(ev[1], s);
    SlotLayout exact;
    exact.overflow = p.overflow();
    if (!rc) rc = run_symbolic(ctx, A, B, C, p, exact);
    free_heavy_bitmaps(ctx, p);  // the fill has used them
    // tile_nnz[0] = 0 by the scan; nnzC also lands next to the capacity flag,
    // and the scan publishes both for the host
    int64_t seq = 0;
    if (!rc) rc = scan_publish_i64(ctx, C->tile_nnz, C->numtile + 1, true, &p.flags->nnz, &p.flags->cap, &seq);
    if (!rc) rc = prepare_numeric(ctx, A, C, p, exact);
    if (!rc) {
        // the capacity flag and nnzC in one read
        rc = publish_wait(ctx, seq, 4, ctx->pinned);
        syncs++;
        C->nnz = ctx->pinned[1];
        if (!rc && *reinterpret_cast<int32_t *>(ctx->pinned))
            rc = fail(TSG_ERR_OVERFLOW, "a staged row exceeded its variant's capacity");
    }
    if (timing && !rc) (void)hipEventRecord(ev[2], s);
    if (!rc && fo) {
        if (C->nnz > fo->cap_nnz) {
            rc = fail(TSG_ERR_OVERFLOW, "C has %lld nonzeros; the caller's value arrays hold %lld", (long long)C->nnz,
                      (long long)fo->cap_nnz);
        } else {
            C->tile_csr_Val = fo->c.tile_csr_Val;
            C->tile_csr_Col = fo->c.tile_csr_Col;
        }
    } else {
        if (!rc) rc = pool_alloc(ctx, &C->tile_csr_Val, (size_t)C->nnz * vsize(C->dtype));
        if (!rc) rc = palloc(ctx, &C->tile_csr_Col, (size_t)C->nnz);
    }
    if (!rc && C->nnz > 0) rc = run_numeric(ctx, A, B, C, p, exact);
    if (timing && !rc) (void)hipEventRecord(ev[3], s);
    if (st) {
        st->numtileC = C->numtile;
        st->nnzC = C->nnz;
        st->host_syncs = syncs;
        st->numeric_kernels = C->nnz > 0 ? numeric_kernel_mask(p) : 0;
    }
    if (timing) {
        if (!rc && st) {
            float a = 0, b = 0, c3 = 0;
            (void)hipEventSynchronize(ev[3]);
            (void)hipEventElapsedTime(&a, ev[0], ev[1]);
            (void)hipEventElapsedTime(&b, ev[1], ev[2]);
            (void)hipEventElapsedTime(&c3, ev[2], ev[3]);
            st->ms_step1 = a; st->ms_step2 = b; st->ms_step3 = c3; st->ms_total = a + b + c3;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    if (keep && !rc) {
        *keep = p;
        p = Plan{};
    }
    free_plan(ctx, p);
    if (rc) free_c(ctx, C);
    return rc;
}

// Capacity check of the device-sized pipeline, after binning: numtileC
// (hdr[0]) against C's tile arrays and every bin's row count against the rows
// its launches cover.  Past any of them the flag is set and every bin emptied,
// so no later kernel writes C (the caller then computes the batch exactly).
__global__ void k_device_guard(int64_t *__restrict__ hdr, int nbins, int64_t cap_tiles, BinCaps caps,
                               int32_t *__restrict__ overflow) {
    if (threadIdx.x != 0) return;
    bool over = hdr[0] > cap_tiles;
    for (int b = 0; b < nbins; b++) over = over || hdr[3 + b] - hdr[2 + b] > caps.v[b];
    if (!over) return;
    atomicOr(overflow, 1);
    for (int b = 0; b <= nbins; b++) hdr[2 + b] = 0;
}

int spgemm_binned_device(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                         const tsg_options *opts, const DeviceCaps &caps, Plan &p, hipEvent_t *ev,
                         int64_t *flags_seq) {
    if (!ctx->device_sizes) return fail(TSG_ERR_ARG, "device-sized pipeline outside a graph build");
    if (A->tilem <= 0) return fail(TSG_ERR_ARG, "a device-sized product needs tile-rows");
    if (!binned_variant(opts) || !staged_ok(A, B, opts))
        return fail(TSG_ERR_ARG, "a device-sized product runs the binned kernels: numeric_variant 0 and B's "
                                 "column-major tile copies");
    memset(C, 0, sizeof *C);
    C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
    C->dtype = A->dtype; C->location = TSG_DEVICE;
    p = Plan{};
    const int32_t tilem = A->tilem;
    TSG_CHECK(palloc(ctx, &C->tile_ptr, (size_t)tilem + 1));
    TSG_CHECK(plan_count(ctx, A, B, C, opts, p, nullptr, !plan_tail_fused(tilem)));
    if (p.nbins > MAX_BINS) return fail(TSG_ERR_ARG, "bin table");
    BinCaps bc{};
    int open = 0, last_open = -1;
    for (int b = 0; b < p.nbins; b++) {
        bc.v[b] = std::min<int64_t>(caps.bins[b], tilem);
        p.bin_cap[b] = bc.v[b];
        if (bc.v[b] > 0) {
            open++;
            last_open = b;
        }
    }
    if (open == 1 && bc.v[last_open] >= tilem) p.all_bin = last_open;
    TSG_CHECK(plan_tail_device(ctx, p, tilem, C->tile_ptr, nullptr, &bc, caps.tiles,
                               reinterpret_cast<int64_t *>(&p.flags->pred_tiles)));
    p.d_bins = p.hdr + 2;
    C->numtile = caps.tiles;
    TSG_CHECK(palloc(ctx, &C->tile_rowidx, (size_t)caps.tiles));
    TSG_CHECK(palloc(ctx, &C->tile_columnidx, (size_t)caps.tiles));
    TSG_CHECK(palloc(ctx, &C->mask, (size_t)caps.tiles * 16));
    TSG_CHECK(palloc(ctx, &C->tile_csr_Ptr, (size_t)caps.tiles * 16));
    TSG_CHECK(palloc(ctx, &C->tile_nnz, (size_t)caps.tiles + 1));
    if (ev) (void)hipEventRecord(ev[0], ctx->stream);
    SlotLayout sl;
    sl.overflow = p.overflow();
    sl.val_cap = caps.nnz;
    TSG_CHECK(run_symbolic(ctx, A, B, C, p, sl));
    // C's value offsets over its numtileC tiles (hdr[0] = tile_ptrC[tilem])
    TSG_CHECK(scan_inclusive_i64(ctx, C->tile_nnz, caps.tiles + 1, true, &p.flags->nnz, C->tile_ptr + tilem));
    TSG_CHECK(prepare_numeric(ctx, A, C, p, sl));
    if (ev) (void)hipEventRecord(ev[1], ctx->stream);
    // every capacity check is done by now (numeric kernels write no value past
    // caps.nnz, and nnz(C) is known): the flags can go to the host while the
    // numeric step runs
    if (flags_seq) TSG_CHECK(publish_start(ctx, p.flags, PLAN_FLAG_WORDS, flags_seq));
    C->nnz = caps.nnz;
    TSG_CHECK(pool_alloc(ctx, &C->tile_csr_Val, (size_t)std::max<int64_t>(caps.nnz, 1) * vsize(C->dtype)));
    TSG_CHECK(palloc(ctx, &C->tile_csr_Col, (size_t)std::max<int64_t>(caps.nnz, 1)));
    return run_numeric(ctx, A, B, C, p, sl);
}

int replay_numeric(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, const tsg_tile_matrix *C,
                   NumericReplay &r) {
    if (C->nnz == 0) return TSG_OK;
    if (r.path == 0) {
        SlotLayout exact;
        exact.overflow = r.p.overflow();
        return run_numeric(ctx, A, B, C, r.p, exact);
    }
    tsg_tile_matrix Cm = *C;
    if (r.path == 2) return launch_numeric_mfma(ctx, A, B, &Cm, &r.o);
    return launch_numeric(ctx, A, B, &Cm, &r.o);
}

}  // namespace tsg
