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
#include <algorithm>
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
            int64_t nrows, const int64_t *__restrict__ dlo, const int64_t *__restrict__ dhi,
            int32_t *__restrict__ row_nc, int32_t *__restrict__ overflow,
            int32_t *__restrict__ hlist, int32_t *__restrict__ hcount) {
    __shared__ uint32_t s_bm[ST_WAVES][ST_WIN_WORDS];
    __shared__ int64_t s_bs[ST_WAVES][WAVE];
    __shared__ int32_t s_pre[ST_WAVES][WAVE];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t idx = (int64_t)blockIdx.x * ST_WAVES + w;
    const int32_t bi = rows ? bin_row(rows, row_begin, nrows, dlo, dhi, idx) : (idx < tilemA ? (int32_t)idx : -1);
    if (bi < 0) return;  // whole wave leaves; no block barrier below
    uint32_t *bm = s_bm[w];
    const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];

    int32_t cmin = INT_MAX, cmax = -1;
    int64_t nb = 0;
    for (int64_t i = as + lane; i < ae; i += WAVE) {
        int32_t k = colidxA[i];
        int64_t bs = tile_ptrB[k], be = tile_ptrB[k + 1];
        nb += be - bs;
        if (be > bs) {
            cmin = min(cmin, colidxB[bs]);
            cmax = max(cmax, colidxB[be - 1]);
        }
    }
    cmin = wave_min(cmin);
    cmax = wave_max(cmax);
    if (hlist && wave_sum(nb) > HEAVY_NB) {  // a workgroup kernel takes this row
        if (lane == 0) hlist[atomicAdd(hcount, 1)] = bi;
        return;
    }

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
        nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int launch_structure_fill(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                          const int64_t *tile_ptrC, int32_t *rowidxC, int32_t *colidxC) {
    if (A->tilem == 0) return TSG_OK;
    unsigned grid = (unsigned)((A->tilem + ST_WAVES - 1) / ST_WAVES);
    k_structure<true><<<grid, ST_WAVES * WAVE, 0, ctx->stream>>>(
        A->tile_ptr, A->tile_columnidx, A->tilem, B->tile_ptr, B->tile_columnidx,
        const_cast<int64_t *>(tile_ptrC), rowidxC, colidxC, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr,
        nullptr, nullptr);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

// ---------------------------------------------------------------- heavy tile-rows
// A tile-row reaching more than HEAVY_NB B tiles gets a whole workgroup: its
// A tiles in blocks of 256 (a thread per A tile walks a short B tile-row list
// itself; longer lists go to a queue the whole block walks, coalesced), into
// an LDS bitmap of HS_WIN_WORDS * 32 tile columns per pass.  Rows come from a
// device list; a persistent grid takes them from a counter.
constexpr int HS_THREADS = 256;
constexpr int HS_WIN_WORDS = 2048;  // 65,536 tile columns per pass
constexpr int HS_SHORT = HEAVY_LIST;  // B tile-row lists walked by one thread

__device__ __forceinline__ int hs_block_sum(int v, int *s_w) {
    v = wave_sum(v);
    __syncthreads();
    if (lane_id() == 0) s_w[threadIdx.x / WAVE] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < HS_THREADS / WAVE; w++) t += s_w[w];
    return t;
}

template <bool FILL>
__global__ void __launch_bounds__(HS_THREADS)
k_heavy_structure(const int32_t *__restrict__ hlist, const int32_t *__restrict__ hcount,
                  unsigned long long *__restrict__ ctr, const int64_t *__restrict__ tile_ptrA,
                  const int32_t *__restrict__ colidxA, const int64_t *__restrict__ tile_ptrB,
                  const int32_t *__restrict__ colidxB, int64_t *__restrict__ tile_ptrC,
                  int32_t *__restrict__ rowidxC, int32_t *__restrict__ colidxC, int32_t *__restrict__ row_nc,
                  int32_t *__restrict__ overflow, int32_t *__restrict__ stats, HeavyBitmaps hb) {
    __shared__ uint32_t bm[HS_WIN_WORDS];
    __shared__ int64_t s_off;
    __shared__ int64_t s_bo[HS_THREADS];  // queued B tile-row bitmaps: first word, word count, first column word
    __shared__ int32_t s_bw[HS_THREADS];
    __shared__ int32_t s_bl[HS_THREADS];
    __shared__ int s_nb;
    __shared__ int32_t s_tpre[HS_THREADS];  // fill: first window entry of each thread's words
    __shared__ int64_t s_qbs[HS_THREADS];
    __shared__ int32_t s_qlen[HS_THREADS];
    __shared__ int s_w[HS_THREADS / WAVE];
    __shared__ int s_nq, s_h;
    __shared__ int32_t s_min, s_max;
    const int tid = threadIdx.x;
    const int nh = *hcount;
    if ((int64_t)blockIdx.x >= nh) return;  // more workgroups than heavy rows: no counter traffic
    for (;;) {
        if (tid == 0) {
            s_h = (int)atomicAdd(ctr, 1ull);
            s_nq = 0;
            s_nb = 0;
            s_min = INT_MAX;
            s_max = -1;
        }
        __syncthreads();
        const int h = s_h;
        if (h >= nh) break;  // uniform
        const int32_t bi = hlist[h];
        const int64_t as = tile_ptrA[bi], ae = tile_ptrA[bi + 1];
        {
            int32_t mn = INT_MAX, mx = -1;
            for (int64_t i = as + tid; i < ae; i += HS_THREADS) {
                const int32_t k = colidxA[i];
                const int64_t bs = tile_ptrB[k], be = tile_ptrB[k + 1];
                if (be > bs) {
                    mn = min(mn, colidxB[bs]);
                    mx = max(mx, colidxB[be - 1]);
                }
            }
            mn = wave_min(mn);
            mx = wave_max(mx);
            if (lane_id() == 0) {
                atomicMin(&s_min, mn);
                atomicMax(&s_max, mx);
            }
        }
        __syncthreads();
        // windows start on a word boundary (B tile-row bitmaps are word-aligned)
        const int32_t cmin = s_min & ~31, cmax = s_max;
        if (tid == 0) {  // where this row's union bitmap is kept (count) or found (fill)
            int64_t o = -1;
            if (!FILL && hb.off) {
                const int64_t words = cmax >= cmin ? (int64_t)(cmax - cmin) / 32 + 1 : 0;
                if (words > 0) {
                    o = (int64_t)atomicAdd(hb.top, (unsigned long long)words);
                    if (o + words > hb.cap) o = -1;
                }
                hb.off[bi] = o;
            } else if (FILL && hb.off) {
                o = hb.off[bi];
            }
            s_off = o;
        }
        __syncthreads();
        const int64_t boff = s_off;
        int64_t out = FILL ? tile_ptrC[bi] : 0;
        const int64_t slot_end = (FILL && row_nc) ? tile_ptrC[bi + 1] : INT64_MAX;
        int64_t count = 0;
        for (int64_t w0 = cmin; w0 <= cmax; w0 += (int64_t)HS_WIN_WORDS * 32) {  // uniform
            const int64_t wbits = min<int64_t>((int64_t)HS_WIN_WORDS * 32, cmax - w0 + 1);
            const int nwords = (int)((wbits + 31) >> 5);
            const int64_t wbase = boff + ((w0 - cmin) >> 5);  // this window's first kept word
            if (FILL && boff >= 0) {  // the count left this row's bitmap
                for (int i = tid; i < HS_WIN_WORDS; i += HS_THREADS) bm[i] = i < nwords ? hb.words[wbase + i] : 0u;
                __syncthreads();
            } else {
                for (int i = tid; i < HS_WIN_WORDS; i += HS_THREADS) bm[i] = 0u;
                __syncthreads();
                for (int64_t base = as; base < ae; base += HS_THREADS) {  // uniform trip count
                    const int64_t i = base + tid;
                    if (i < ae) {
                        const int32_t k = colidxA[i];
                        const int64_t bs = tile_ptrB[k];
                        const int32_t len = (int32_t)(tile_ptrB[k + 1] - bs);
                        int64_t bo = 0, bw = 0;
                        if (hb.boff) {
                            bo = hb.boff[k];
                            bw = hb.boff[k + 1] - bo;
                        }
                        if (bw > 0) {  // OR the row's bitmap words
                            const int qi = atomicAdd(&s_nb, 1);
                            s_bo[qi] = bo;
                            s_bw[qi] = (int32_t)bw;
                            s_bl[qi] = colidxB[bs] >> 5;
                        } else if (len > HS_SHORT) {
                            const int qi = atomicAdd(&s_nq, 1);
                            s_qbs[qi] = bs;
                            s_qlen[qi] = len;
                        } else {
                            for (int32_t q = 0; q < len; q++) {
                                const int64_t off = (int64_t)colidxB[bs + q] - w0;
                                if (off >= 0 && off < wbits) atomicOr(&bm[off >> 5], 1u << (off & 31));
                            }
                        }
                    }
                    __syncthreads();
                    const int nq = s_nq;
                    for (int qi = 0; qi < nq; qi++) {  // long lists: the whole block, coalesced
                        const int64_t bs = s_qbs[qi];
                        const int32_t len = s_qlen[qi];
                        for (int32_t q = tid; q < len; q += HS_THREADS) {
                            const int64_t off = (int64_t)colidxB[bs + q] - w0;
                            if (off >= 0 && off < wbits) atomicOr(&bm[off >> 5], 1u << (off & 31));
                        }
                    }
                    const int nbq = s_nb;
                    const int64_t wlo = w0 >> 5;
                    for (int qi = 0; qi < nbq; qi++) {  // B tile-row bitmaps: one LDS atomic per word
                        const int64_t bo = s_bo[qi];
                        const int32_t bw = s_bw[qi];
                        const int64_t bl = s_bl[qi];
                        for (int32_t q = tid; q < bw; q += HS_THREADS) {
                            const int64_t wi = bl + q - wlo;
                            if (wi >= 0 && wi < nwords) {
                                const uint32_t v = hb.bwords[bo + q];
                                if (v) atomicOr(&bm[wi], v);
                            }
                        }
                    }
                    __syncthreads();
                    if (tid == 0) {
                        s_nq = 0;
                        s_nb = 0;
                    }
                    __syncthreads();
                }
                if (!FILL && boff >= 0)  // keep the window's words for the fill
                    for (int i = tid; i < nwords; i += HS_THREADS) hb.words[wbase + i] = bm[i];
            }
            // each thread owns HS_WIN_WORDS / HS_THREADS consecutive words
            constexpr int WPT = HS_WIN_WORDS / HS_THREADS;
            int cnt = 0;
#pragma unroll
            for (int q = 0; q < WPT; q++) cnt += __popc(bm[tid * WPT + q]);
            if (!FILL) {
                count += hs_block_sum(cnt, s_w);
            } else {
                // exclusive scan of the per-thread counts: waves, then wave totals
                const int incl = wave_incl_sum(cnt);
                __syncthreads();
                if (lane_id() == WAVE - 1) s_w[tid / WAVE] = incl;
                __syncthreads();
                int before = 0, tot = 0;
                for (int w = 0; w < HS_THREADS / WAVE; w++) {
                    if (w < tid / WAVE) before += s_w[w];
                    tot += s_w[w];
                }
                // entry e of the window (ascending column) is written by thread
                // e % HS_THREADS, so the stores are coalesced: the owner of its
                // words is found by a binary search of the threads' starts, the
                // word by the owner's word counts, the bit by halving selects
                s_tpre[tid] = before + incl - cnt;
                __syncthreads();
                for (int e = tid; e < tot; e += HS_THREADS) {
                    int lo = 0, hi = HS_THREADS - 1;  // last thread starting at or before e
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_tpre[mid] <= e) lo = mid; else hi = mid - 1;
                    }
                    int r = e - s_tpre[lo], wd = lo * WPT;
                    uint32_t x = bm[wd];
                    for (int q = 1; q < WPT; q++) {  // stops at the word holding entry r
                        const int pc = __popc(x);
                        if (r < pc) break;
                        r -= pc;
                        x = bm[++wd];
                    }
                    int bit = 0;  // the r-th lowest set bit of x
#pragma unroll
                    for (int half = 16; half > 0; half >>= 1) {
                        const uint32_t low = x & ((1u << half) - 1u);
                        const int c = __popc(low);
                        if (r >= c) {
                            r -= c;
                            x >>= half;
                            bit += half;
                        } else {
                            x = low;
                        }
                    }
                    const int64_t pos = out + e;
                    if (pos < slot_end) {
                        rowidxC[pos] = bi;
                        colidxC[pos] = (int32_t)(w0 + (int64_t)wd * 32 + bit);
                    }
                }
                out += tot;
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (!FILL) {
                tile_ptrC[bi + 1] = count;
                if (stats) stats[(int64_t)bi * RS_NSTAT + RS_COUNT] = (int32_t)min<int64_t>(count, INT_MAX);
            } else if (row_nc) {
                const int64_t n = min<int64_t>(out, slot_end) - tile_ptrC[bi];
                row_nc[bi] = (int32_t)min<int64_t>(n, INT32_MAX);
                if (out > slot_end) atomicOr(overflow, 1);
            }
        }
        __syncthreads();
    }
}

// Words of B tile-row k's column bitmap when it pays (more than HS_SHORT
// tiles, at least two per word, at most one LDS window), else 0, into
// boff[k + 1]; boff[0] = 0 (a scan then gives each row's first word).
__global__ void __launch_bounds__(256) k_brow_words(int32_t tilemB, const int64_t *__restrict__ tile_ptrB,
                                                    const int32_t *__restrict__ colidxB, int64_t *__restrict__ boff) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k == 0) boff[0] = 0;
    if (k >= tilemB) return;
    const int64_t bs = tile_ptrB[k], len = tile_ptrB[k + 1] - bs;
    int64_t words = 0;
    if (len > HS_SHORT) {
        const int64_t w = (int64_t)(colidxB[bs + len - 1] >> 5) - (colidxB[bs] >> 5) + 1;
        if (2 * w <= len && w <= HS_WIN_WORDS) words = w;
    }
    boff[k + 1] = words;
}

// A workgroup per B tile-row with a bitmap: set its bits in LDS, store the words.
__global__ void __launch_bounds__(HS_THREADS)
k_brow_fill(int32_t tilemB, const int64_t *__restrict__ tile_ptrB, const int32_t *__restrict__ colidxB,
            const int64_t *__restrict__ boff, uint32_t *__restrict__ bwords) {
    __shared__ uint32_t bm[HS_WIN_WORDS];
    const int tid = threadIdx.x;
    for (int64_t k = blockIdx.x; k < tilemB; k += gridDim.x) {  // uniform
        const int64_t o = boff[k];
        const int32_t w = (int32_t)(boff[k + 1] - o);
        if (w == 0) continue;
        const int64_t bs = tile_ptrB[k], be = tile_ptrB[k + 1];
        const int32_t lo = colidxB[bs] >> 5;
        for (int i = tid; i < w; i += HS_THREADS) bm[i] = 0u;
        __syncthreads();
        for (int64_t q = bs + tid; q < be; q += HS_THREADS) {
            const int32_t c = colidxB[q];
            atomicOr(&bm[(c >> 5) - lo], 1u << (c & 31));
        }
        __syncthreads();
        for (int i = tid; i < w; i += HS_THREADS) bwords[o + i] = bm[i];
        __syncthreads();
    }
}

int build_brow_bitmaps(Context *ctx, const tsg_tile_matrix *B, HeavyBitmaps &hb) {
    const int32_t tilemB = B->tilem;
    TSG_CHECK(palloc(ctx, &hb.boff, (size_t)tilemB + 1));
    TSG_CHECK(palloc(ctx, &hb.bwords, (size_t)(B->numtile / 2 + 1)));
    k_brow_words<<<(unsigned)(((int64_t)tilemB + 256) / 256), 256, 0, ctx->stream>>>(tilemB, B->tile_ptr,
                                                                                     B->tile_columnidx, hb.boff);
    TSG_HIP(hipGetLastError());
    TSG_CHECK(scan_inclusive_i64(ctx, hb.boff, (int64_t)tilemB + 1));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tilemB, ctx->prop.multiProcessorCount * 8));
    k_brow_fill<<<grid, HS_THREADS, 0, ctx->stream>>>(tilemB, B->tile_ptr, B->tile_columnidx, hb.boff, hb.bwords);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

// Exact C tile counts of the heavy rows listed by k_row_stats (written into
// tile_ptrC[bi + 1] before its scan).  hc = {row count, work counter}.
int launch_heavy_count(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, int64_t *tile_ptrC,
                       const int32_t *hlist, int32_t *hc, int32_t *stats, const HeavyBitmaps &hb) {
    k_heavy_structure<false><<<(unsigned)std::max(1, ctx->prop.multiProcessorCount * 8), HS_THREADS, 0, ctx->stream>>>(
        hlist, hc, reinterpret_cast<unsigned long long *>(hc + 2), A->tile_ptr, A->tile_columnidx, B->tile_ptr,
        B->tile_columnidx, tile_ptrC, nullptr, nullptr, nullptr, nullptr, stats, hb);
    TSG_HIP(hipGetLastError());
    return TSG_OK;
}

int launch_structure_fill_list(Context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                               const tsg_tile_matrix *C, const int32_t *rows, const BinRange &br,
                               const SlotLayout &sl, const HeavyBitmaps &hb) {
    const int64_t count = br.count;
    if (count <= 0) return TSG_OK;
    int32_t *hlist = nullptr, *hc = nullptr;  // heavy rows; hc = {count, pad, work counter (8 bytes)}
    TSG_CHECK(palloc(ctx, &hlist, (size_t)count));
    int rc = palloc(ctx, &hc, 4);
    if (!rc) rc = zero_small(ctx, hc, 4, nullptr, 0);  // a kernel: capturable
    if (!rc) {
        unsigned grid = (unsigned)((count + ST_WAVES - 1) / ST_WAVES);
        k_structure<true><<<grid, ST_WAVES * WAVE, 0, ctx->stream>>>(
            A->tile_ptr, A->tile_columnidx, A->tilem, B->tile_ptr, B->tile_columnidx, C->tile_ptr, C->tile_rowidx,
            C->tile_columnidx, rows, br.begin, count, br.lo, br.hi, sl.row_nc, sl.overflow, hlist, hc);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_structure launch");
    }
    if (!rc) {
        k_heavy_structure<true><<<(unsigned)std::max(1, ctx->prop.multiProcessorCount * 8), HS_THREADS, 0,
                                   ctx->stream>>>(
            hlist, hc, reinterpret_cast<unsigned long long *>(hc + 2), A->tile_ptr, A->tile_columnidx, B->tile_ptr,
            B->tile_columnidx, C->tile_ptr, C->tile_rowidx, C->tile_columnidx, sl.row_nc, sl.overflow, nullptr, hb);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_heavy_structure launch");
    }
    pool_free(ctx, hlist);
    pool_free(ctx, hc);
    return rc;
}

}  // namespace tsg

// ---------------------------------------------------------------- tile bitmaps
namespace tsg {

// One wave per list (a tile-row of A, or a tile column of B in CSC-tile
// order): bit 31 - k % 32 of word k / 32 for every tile index k it holds.
__global__ void __launch_bounds__(256) k_tile_bitmaps(int32_t nlists, const int64_t *__restrict__ ptr,
                                                      const int32_t *__restrict__ idx, int32_t bm_len,
                                                      uint32_t *__restrict__ bm) {
    const int64_t l = ((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE;
    if (l >= nlists) return;
    for (int64_t j = ptr[l] + lane_id(); j < ptr[l + 1]; j += WAVE) {
        const int32_t k = idx[j];
        atomicOr(&bm[l * bm_len + (k >> 5)], 1u << (31 - (k & 31)));
    }
}

// Bitmaps of n lists into the caller's host array out[n * bm_len].
static int tile_bitmaps_to_host(Context *ctx, int32_t n, const int64_t *ptr, const int32_t *idx, int32_t bm_len,
                                uint32_t *out) {
    const size_t words = (size_t)n * (size_t)bm_len;
    if (words == 0) return TSG_OK;
    uint32_t *d = nullptr;
    TSG_CHECK(palloc(ctx, &d, words));
    int rc = TSG_OK;
    if (hipMemsetAsync(d, 0, words * sizeof(uint32_t), ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "memset");
    if (!rc) {
        k_tile_bitmaps<<<(unsigned)(((int64_t)n * WAVE + 255) / 256), 256, 0, ctx->stream>>>(n, ptr, idx, bm_len, d);
        if (hipGetLastError() != hipSuccess) rc = fail(TSG_ERR_HIP, "k_tile_bitmaps launch");
    }
    if (!rc && (hipMemcpyAsync(out, d, words * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = fail(TSG_ERR_HIP, "bitmap download failed");
    pool_free(ctx, d);
    return rc;
}

}  // namespace tsg

extern "C" int tsg_build_bitmasks(tsg_context *c, const tsg_tile_matrix *A, const tsg_tile_matrix *B, uint32_t *bmA,
                                  uint32_t *bmB, int32_t *bm_len) {
    using namespace tsg;
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !bm_len) return fail(TSG_ERR_ARG, "null argument");
    TSG_CHECK(check_operands(A, B));
    (void)hipSetDevice(ctx->device);
    *bm_len = (A->tilen + 31) / 32;
    if (bmA) TSG_CHECK(tile_bitmaps_to_host(ctx, A->tilem, A->tile_ptr, A->tile_columnidx, *bm_len, bmA));
    if (bmB) TSG_CHECK(tile_bitmaps_to_host(ctx, B->tilen, B->csc_tile_ptr, B->csc_tile_rowidx, *bm_len, bmB));
    return TSG_OK;
}
