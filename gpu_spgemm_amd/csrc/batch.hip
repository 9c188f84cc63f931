// batch.hip -- many independent products C_i = A_i * B_i in one pipeline run.
//
// No reference counterpart (the reference computes one product per process,
// src/spgemm_raw_v2.cu:61-372).  A batch is one product of block-diagonal
// operands: diag(A_1..A_n) * diag(B_1..B_n) = diag(C_1..C_n).  The tile
// arrays of every A_i and B_i are concatenated into one A and one B with
// shifted tile-row, tile-column and payload offsets (each problem starts on
// a tile boundary, so partial edge tiles stay inside their own block), the
// binned pipeline runs once over all tile-rows of all problems (one launch
// per kernel variant for the whole batch, rows of every problem binned
// together), and one split launch copies every C_i's tile offsets and
// indices out with the batch offsets removed; its masks, row offsets, values
// and col bytes stay where the product wrote them (pool views into C's
// arrays, recycled when the last of them is freed).  Host syncs per batch:
// those of one product plus one read of the per-problem boundaries, instead
// of those of n products.
//
// Results are bit-identical to n separate calls: every kernel computes each
// C tile from the same A and B tiles in the same order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rowstaged.h"

namespace tsg {

__global__ void __launch_bounds__(256) k_cat(const CatJob *__restrict__ jobs) {
    const CatJob j = jobs[blockIdx.y];
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < j.count; i += stride) {
        switch (j.kind) {
        case 0: static_cast<uint8_t *>(j.dst)[i] = static_cast<const uint8_t *>(j.src)[i]; break;
        case 1: static_cast<uint4 *>(j.dst)[i] = static_cast<const uint4 *>(j.src)[i]; break;
        case 2: static_cast<int32_t *>(j.dst)[i] = static_cast<const int32_t *>(j.src)[i] + (int32_t)j.add; break;
        case 3: static_cast<int64_t *>(j.dst)[i] = static_cast<const int64_t *>(j.src)[i] + j.add; break;
        case 4: static_cast<uint64_t *>(j.dst)[i] = static_cast<const uint64_t *>(j.src)[i]; break;
        case 6: static_cast<int64_t *>(j.dst)[i] = j.add; break;
        default: static_cast<uint32_t *>(j.dst)[i] = static_cast<const uint32_t *>(j.src)[i]; break;
        }
    }
}

// bounds[2i] = C tile offset of problem i's first tile-row, bounds[2i+1] = its
// first value (i = 0..n, row_off[n] = total tile-rows).
__global__ void __launch_bounds__(256) k_bounds(int32_t n, const int64_t *__restrict__ row_off,
                                                const int64_t *__restrict__ tile_ptr,
                                                const int64_t *__restrict__ tile_nnz, int64_t *__restrict__ bounds) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    const int64_t t = tile_ptr[row_off[i]];
    bounds[2 * i] = t;
    bounds[2 * i + 1] = tile_nnz[t];
}

namespace {

struct JobList {
    std::vector<CatJob> jobs;
    void add(const void *src, void *dst, int64_t count, int kind, int64_t add = 0) {
        if (count > 0 && src) jobs.push_back(CatJob{src, dst, count, add, kind, 0});
    }
    // count int64 copies of `value`
    void fill(int64_t *dst, int64_t count, int64_t value) {
        if (count > 0) jobs.push_back(CatJob{nullptr, dst, count, value, 6, 0});
    }
};

int64_t max_count(const JobList &jl) {
    int64_t maxc = 0;
    for (const CatJob &j : jl.jobs) maxc = std::max(maxc, j.count);
    return maxc;
}

// Device copy of a job table (the host vector must live until the stream syncs).
int upload_jobs(Context *ctx, const JobList &jl, CatJob **out) {
    *out = nullptr;
    if (jl.jobs.empty()) return TSG_OK;
    TSG_CHECK(palloc(ctx, out, jl.jobs.size()));
    if (hipMemcpyAsync(*out, jl.jobs.data(), sizeof(CatJob) * jl.jobs.size(), hipMemcpyHostToDevice, ctx->stream) !=
        hipSuccess)
        return fail(TSG_ERR_HIP, "job table upload failed");
    return TSG_OK;
}

int run_jobs(Context *ctx, const JobList &jl) {
    if (jl.jobs.empty()) return TSG_OK;
    CatJob *dj = nullptr;
    int rc = upload_jobs(ctx, jl, &dj);
    if (!rc) rc = launch_jobs(ctx, dj, (int64_t)jl.jobs.size(), max_count(jl));
    // the table must outlive the copies that read it
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(TSG_ERR_HIP, "batch copy failed");
    pool_free(ctx, dj);
    return rc;
}

}  // namespace

int launch_jobs(Context *ctx, const CatJob *dj, int64_t n, int64_t maxc) {
    for (int64_t j0 = 0; j0 < n; j0 += 65535) {
        const unsigned ny = (unsigned)std::min<int64_t>(65535, n - j0);
        const unsigned nx = (unsigned)std::min<int64_t>(std::max<int64_t>((maxc + 255) / 256, 1), 64);
        k_cat<<<dim3(nx, ny), 256, 0, ctx->stream>>>(dj + j0);
        TSG_HIP(hipGetLastError());
    }
    return TSG_OK;
}

void free_batch_keep(Context *ctx, BatchKeep &k) {
    tsg_tile_matrix *ms[] = {&k.BA, &k.BB, &k.BC};
    for (tsg_tile_matrix *M : ms) {
        void *ps[] = {M->tile_ptr, M->tile_columnidx, M->tile_rowidx, M->tile_nnz, M->tile_csr_Val, M->tile_csr_Col,
                      M->tile_csr_Ptr, M->mask, M->csc_tile_ptr, M->csc_tile_rowidx, M->csr_tile_nnz, M->csr_mask,
                      M->csr_col_mask, M->csr_col_Ptr, M->csr_col_Val, M->row_nnz_ptr, M->row_ptr, M->row_col,
                      M->row_val, M->val_perm, M->val_perm_col, M->tile_csc_Ptr, M->tile_csc_Row, M->tile_csc_Val};
        for (void *p : ps) pool_free(ctx, p);
    }
    free_plan(ctx, k.nr.p);
    pool_free(ctx, k.vin);
    pool_free(ctx, k.vout);
    k = BatchKeep{};
}

int spgemm_batched(Context *ctx, int32_t n, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *stats, BatchKeep *keep) {
    if (n < 0) return fail(TSG_ERR_ARG, "negative batch size");
    if (n == 0) return TSG_OK;
    const int32_t dtype = A[0].dtype;
    // offsets of every problem in the block-diagonal operands
    std::vector<int64_t> R(n + 1, 0), K(n + 1, 0), Cc(n + 1, 0), at(n + 1, 0), an(n + 1, 0), bt(n + 1, 0),
        bn(n + 1, 0);
    bool staged = true, brow = true, scal = true;
    int32_t flags = 0;
    for (int32_t i = 0; i < n; i++) {
        TSG_CHECK(check_operands(&A[i], &B[i]));
        if (A[i].dtype != dtype) return fail(TSG_ERR_ARG, "problem %d: value types differ within the batch", i);
        staged = staged && B[i].csr_tile_nnz && B[i].csr_mask && B[i].csr_col_mask && B[i].csr_col_Ptr &&
                 B[i].csr_col_Val && B[i].row_nnz_ptr;
        brow = brow && B[i].tile_rowidx;
        scal = scal && A[i].row_ptr && A[i].row_col && A[i].row_val && B[i].row_ptr && B[i].row_col && B[i].row_val;
        flags |= ((A[i].flags | B[i].flags) & 1) | (B[i].flags & 2);  // repeats; long B tile-rows
        R[i + 1] = R[i] + A[i].tilem;
        K[i + 1] = K[i] + A[i].tilen;
        Cc[i + 1] = Cc[i] + B[i].tilen;
        at[i + 1] = at[i] + A[i].numtile;
        an[i + 1] = an[i] + A[i].nnz;
        bt[i + 1] = bt[i] + B[i].numtile;
        bn[i + 1] = bn[i] + B[i].nnz;
    }
    if (R[n] > MAX_TILE_ROWS || K[n] > MAX_TILE_ROWS || Cc[n] > MAX_TILE_ROWS)
        return fail(TSG_ERR_OVERFLOW, "batch dimensions exceed %lld tile-rows", (long long)MAX_TILE_ROWS);
    // the block-diagonal operands are addressed with int32 tile ids like any other
    if (at[n] > INT32_MAX || bt[n] > INT32_MAX)
        return fail(TSG_ERR_OVERFLOW, "the batch's operand tiles (%lld / %lld) exceed int32 tile ids", (long long)at[n],
                    (long long)bt[n]);
    const size_t vs = vsize(dtype);
    tsg_tile_matrix BA{}, BB{}, BC{};
    BA.m = (int32_t)(R[n] * 16); BA.n = (int32_t)(K[n] * 16); BA.tilem = (int32_t)R[n]; BA.tilen = (int32_t)K[n];
    BA.numtile = at[n]; BA.nnz = an[n]; BA.dtype = dtype; BA.location = TSG_DEVICE; BA.flags = flags & 1;
    BB.m = BA.n; BB.n = (int32_t)(Cc[n] * 16); BB.tilem = BA.tilen; BB.tilen = (int32_t)Cc[n];
    BB.numtile = bt[n]; BB.nnz = bn[n]; BB.dtype = dtype; BB.location = TSG_DEVICE; BB.is_csc = 1; BB.flags = flags;
    auto free_m = [&](tsg_tile_matrix &M) {
        void *ps[] = {M.tile_ptr, M.tile_columnidx, M.tile_rowidx, M.tile_nnz, M.tile_csr_Val, M.tile_csr_Col,
                      M.tile_csr_Ptr, M.mask, M.csc_tile_ptr, M.csc_tile_rowidx, M.csr_tile_nnz, M.csr_mask,
                      M.csr_col_mask, M.csr_col_Ptr, M.csr_col_Val, M.row_nnz_ptr, M.row_ptr, M.row_col,
                      M.row_val};
        for (void *p : ps) pool_free(ctx, p);
        std::memset(&M, 0, sizeof M);
    };
    const int64_t ntA = BA.numtile, ntB = BB.numtile;
    int rc = palloc(ctx, &BA.tile_ptr, (size_t)BA.tilem + 1);
    if (!rc) rc = palloc(ctx, &BA.tile_columnidx, (size_t)ntA);
    if (!rc) rc = palloc(ctx, &BA.tile_nnz, (size_t)ntA + 1);
    if (!rc) rc = palloc(ctx, &BA.tile_csr_Ptr, (size_t)ntA * 16);
    if (!rc) rc = palloc(ctx, &BA.mask, (size_t)ntA * 16);
    if (!rc) rc = palloc(ctx, &BA.tile_csr_Col, (size_t)BA.nnz);
    if (!rc) rc = pool_alloc(ctx, &BA.tile_csr_Val, (size_t)BA.nnz * vs);
    if (!rc) rc = palloc(ctx, &BB.tile_ptr, (size_t)BB.tilem + 1);
    if (!rc) rc = palloc(ctx, &BB.tile_columnidx, (size_t)ntB);
    if (!rc && brow) rc = palloc(ctx, &BB.tile_rowidx, (size_t)ntB);
    if (!rc) rc = palloc(ctx, &BB.tile_nnz, (size_t)ntB + 1);
    if (!rc) rc = palloc(ctx, &BB.tile_csr_Ptr, (size_t)ntB * 16);
    if (!rc) rc = palloc(ctx, &BB.mask, (size_t)ntB * 16);
    if (!rc) rc = palloc(ctx, &BB.tile_csr_Col, (size_t)BB.nnz);
    if (!rc) rc = pool_alloc(ctx, &BB.tile_csr_Val, (size_t)BB.nnz * vs);
    if (!rc) rc = palloc(ctx, &BB.csc_tile_ptr, (size_t)BB.tilen + 1);
    if (!rc) rc = palloc(ctx, &BB.csc_tile_rowidx, (size_t)ntB);
    if (!rc && staged) {
        rc = palloc(ctx, &BB.csr_tile_nnz, (size_t)ntB + 1);
        if (!rc) rc = palloc(ctx, &BB.csr_mask, (size_t)ntB * 16);
        if (!rc) rc = palloc(ctx, &BB.csr_col_mask, (size_t)ntB * 16);
        if (!rc) rc = palloc(ctx, &BB.csr_col_Ptr, (size_t)ntB * 16);
        if (!rc) rc = pool_alloc(ctx, &BB.csr_col_Val, (size_t)BB.nnz * vs);
        if (!rc) rc = palloc(ctx, &BB.row_nnz_ptr, (size_t)BB.tilem + 1);
    }
    if (!rc && scal) {  // scalar CSR of the padded blocks (rows of each problem padded to its tile-rows)
        rc = palloc(ctx, &BA.row_ptr, (size_t)BA.m + 1);
        if (!rc) rc = palloc(ctx, &BA.row_col, (size_t)BA.nnz);
        if (!rc) rc = pool_alloc(ctx, &BA.row_val, (size_t)BA.nnz * vs);
        if (!rc) rc = palloc(ctx, &BB.row_ptr, (size_t)BB.m + 1);
        if (!rc) rc = palloc(ctx, &BB.row_col, (size_t)BB.nnz);
        if (!rc) rc = pool_alloc(ctx, &BB.row_val, (size_t)BB.nnz * vs);
    }
    if (rc) {
        free_m(BA);
        free_m(BB);
        return rc;
    }
    // concatenation jobs; the final entry of every prefix array is written once
    JobList jl, jv;  // every array; the value arrays alone (re-run of the numeric step)
    const int vk = vs == 8 ? 4 : 5;
    for (int32_t i = 0; i < n; i++) {
        const tsg_tile_matrix &a = A[i], &b = B[i];
        jl.add(a.tile_ptr, BA.tile_ptr + R[i], a.tilem, 3, at[i]);
        jl.add(a.tile_columnidx, BA.tile_columnidx + at[i], a.numtile, 2, K[i]);
        jl.add(a.tile_nnz, BA.tile_nnz + at[i], a.numtile, 3, an[i]);
        jl.add(a.tile_csr_Ptr, BA.tile_csr_Ptr + at[i] * 16, a.numtile, 1);
        jl.add(a.mask, BA.mask + at[i] * 16, a.numtile * 2, 1);
        jl.add(a.tile_csr_Col, BA.tile_csr_Col + an[i], a.nnz, 0);
        jl.add(a.tile_csr_Val, (uint8_t *)BA.tile_csr_Val + an[i] * vs, a.nnz, vk);
        jv.add(a.tile_csr_Val, (uint8_t *)BA.tile_csr_Val + an[i] * vs, a.nnz, vk);
        jl.add(b.tile_ptr, BB.tile_ptr + K[i], b.tilem, 3, bt[i]);
        jl.add(b.tile_columnidx, BB.tile_columnidx + bt[i], b.numtile, 2, Cc[i]);
        if (brow) jl.add(b.tile_rowidx, BB.tile_rowidx + bt[i], b.numtile, 2, K[i]);
        jl.add(b.tile_nnz, BB.tile_nnz + bt[i], b.numtile, 3, bn[i]);
        jl.add(b.tile_csr_Ptr, BB.tile_csr_Ptr + bt[i] * 16, b.numtile, 1);
        jl.add(b.mask, BB.mask + bt[i] * 16, b.numtile * 2, 1);
        jl.add(b.tile_csr_Col, BB.tile_csr_Col + bn[i], b.nnz, 0);
        jl.add(b.tile_csr_Val, (uint8_t *)BB.tile_csr_Val + bn[i] * vs, b.nnz, vk);
        jv.add(b.tile_csr_Val, (uint8_t *)BB.tile_csr_Val + bn[i] * vs, b.nnz, vk);
        jl.add(b.csc_tile_ptr, BB.csc_tile_ptr + Cc[i], b.tilen, 3, bt[i]);
        jl.add(b.csc_tile_rowidx, BB.csc_tile_rowidx + bt[i], b.numtile, 2, K[i]);
        if (scal) {  // rows: the problem's m rows, then its padding rows (empty: the end offset repeated)
            jl.add(a.row_ptr, BA.row_ptr + R[i] * 16, a.m, 3, an[i]);
            jl.fill(BA.row_ptr + R[i] * 16 + a.m, (int64_t)a.tilem * 16 - a.m, an[i + 1]);
            jl.add(a.row_col, BA.row_col + an[i], a.nnz, 2, K[i] * 16);
            jl.add(a.row_val, (uint8_t *)BA.row_val + an[i] * vs, a.nnz, vk);
            jv.add(a.row_val, (uint8_t *)BA.row_val + an[i] * vs, a.nnz, vk);
            jl.add(b.row_ptr, BB.row_ptr + K[i] * 16, b.m, 3, bn[i]);
            jl.fill(BB.row_ptr + K[i] * 16 + b.m, (int64_t)b.tilem * 16 - b.m, bn[i + 1]);
            jl.add(b.row_col, BB.row_col + bn[i], b.nnz, 2, Cc[i] * 16);
            jl.add(b.row_val, (uint8_t *)BB.row_val + bn[i] * vs, b.nnz, vk);
            jv.add(b.row_val, (uint8_t *)BB.row_val + bn[i] * vs, b.nnz, vk);
        }
        if (staged) {
            jl.add(b.csr_tile_nnz, BB.csr_tile_nnz + bt[i], b.numtile, 3, bn[i]);
            jl.add(b.csr_mask, BB.csr_mask + bt[i] * 16, b.numtile * 2, 1);
            jl.add(b.csr_col_mask, BB.csr_col_mask + bt[i] * 16, b.numtile * 2, 1);
            jl.add(b.csr_col_Ptr, BB.csr_col_Ptr + bt[i] * 16, b.numtile, 1);
            jl.add(b.csr_col_Val, (uint8_t *)BB.csr_col_Val + bn[i] * vs, b.nnz, vk);
            jv.add(b.csr_col_Val, (uint8_t *)BB.csr_col_Val + bn[i] * vs, b.nnz, vk);
            jl.add(b.row_nnz_ptr, BB.row_nnz_ptr + K[i], b.tilem, 3, bn[i]);
        }
    }
    // closing entries of the prefix arrays (tile and payload totals)
    int64_t *tails = ctx->pinned + 16;
    tails[0] = at[n]; tails[1] = an[n]; tails[2] = bt[n]; tails[3] = bn[n];
    if (scal) {  // closing row offsets of the scalar CSR
        jl.fill(BA.row_ptr + BA.m, 1, an[n]);
        jl.fill(BB.row_ptr + BB.m, 1, bn[n]);
    }
    hipStream_t s = ctx->stream;
    if (hipMemcpyAsync(BA.tile_ptr + R[n], &tails[0], 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(BA.tile_nnz + at[n], &tails[1], 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(BB.tile_ptr + K[n], &tails[2], 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(BB.tile_nnz + bt[n], &tails[3], 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(BB.csc_tile_ptr + Cc[n], &tails[2], 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        (staged && (hipMemcpyAsync(BB.csr_tile_nnz + bt[n], &tails[3], 8, hipMemcpyHostToDevice, s) != hipSuccess ||
                    hipMemcpyAsync(BB.row_nnz_ptr + K[n], &tails[3], 8, hipMemcpyHostToDevice, s) != hipSuccess)))
        rc = fail(TSG_ERR_HIP, "batch tail upload failed");
    if (!rc) rc = run_jobs(ctx, jl);
    tsg_stats bst{};
    if (!rc) rc = spgemm_dispatch(ctx, &BA, &BB, &BC, opts, &bst, keep ? &keep->nr : nullptr);
    if (!keep || rc) {
        free_m(BA);
        free_m(BB);
        if (keep) free_plan(ctx, keep->nr.p);
    }
    if (rc) return rc;
    // per-problem boundaries of C (one read)
    int64_t *drow = nullptr, *dbounds = nullptr;
    std::vector<int64_t> hb(2 * ((size_t)n + 1));
    rc = palloc(ctx, &drow, (size_t)n + 1);
    if (!rc) rc = palloc(ctx, &dbounds, 2 * ((size_t)n + 1));
    if (!rc && (hipMemcpyAsync(drow, R.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess))
        rc = fail(TSG_ERR_HIP, "upload failed");
    if (!rc) {
        k_bounds<<<(unsigned)((n + 1 + 255) / 256), 256, 0, s>>>(n, drow, BC.tile_ptr, BC.tile_nnz, dbounds);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hb.data(), dbounds, sizeof(int64_t) * hb.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = fail(TSG_ERR_HIP, "reading the batch boundaries failed");
    }
    pool_free(ctx, drow);
    pool_free(ctx, dbounds);
    // split C into the per-problem outputs (all cleared first: a failure part
    // way frees only what was made)
    for (int32_t i = 0; i < n; i++) std::memset(&C[i], 0, sizeof C[i]);
    for (int32_t i = 0; i < n && !rc; i++) {
        tsg_tile_matrix &c = C[i];
        c.m = A[i].m; c.n = B[i].n; c.tilem = A[i].tilem; c.tilen = B[i].tilen;
        c.dtype = dtype; c.location = TSG_DEVICE;
        c.numtile = hb[2 * i + 2] - hb[2 * i];
        c.nnz = hb[2 * i + 3] - hb[2 * i + 1];
        // offsets are rebased in copies; masks, row offsets, values and col
        // bytes stay where the product wrote them (views into C's arrays)
        const int64_t t0 = hb[2 * i], v0 = hb[2 * i + 1];
        rc = palloc(ctx, &c.tile_ptr, (size_t)c.tilem + 1);
        if (!rc) rc = palloc(ctx, &c.tile_rowidx, (size_t)c.numtile);
        if (!rc) rc = palloc(ctx, &c.tile_columnidx, (size_t)c.numtile);
        if (!rc) rc = palloc(ctx, &c.tile_nnz, (size_t)c.numtile + 1);
        if (!rc && c.numtile > 0) {
            rc = pool_view(ctx, BC.mask, (size_t)t0 * 32, reinterpret_cast<void **>(&c.mask));
            if (!rc) rc = pool_view(ctx, BC.tile_csr_Ptr, (size_t)t0 * 16, reinterpret_cast<void **>(&c.tile_csr_Ptr));
        } else if (!rc) {
            rc = palloc(ctx, &c.mask, 16);
            if (!rc) rc = palloc(ctx, &c.tile_csr_Ptr, 16);
        }
        if (!rc && c.nnz > 0) {
            rc = pool_view(ctx, BC.tile_csr_Val, (size_t)v0 * vs, &c.tile_csr_Val);
            if (!rc) rc = pool_view(ctx, BC.tile_csr_Col, (size_t)v0, reinterpret_cast<void **>(&c.tile_csr_Col));
        } else if (!rc) {
            rc = pool_alloc(ctx, &c.tile_csr_Val, vs);
            if (!rc) rc = palloc(ctx, &c.tile_csr_Col, 1);
        }
    }
    if (!rc) {
        JobList sj;
        for (int32_t i = 0; i < n; i++) {
            tsg_tile_matrix &c = C[i];
            const int64_t t0 = hb[2 * i], v0 = hb[2 * i + 1];
            sj.add(BC.tile_ptr + R[i], c.tile_ptr, (int64_t)c.tilem + 1, 3, -t0);
            sj.add(BC.tile_rowidx + t0, c.tile_rowidx, c.numtile, 2, -R[i]);
            sj.add(BC.tile_columnidx + t0, c.tile_columnidx, c.numtile, 2, -Cc[i]);
            sj.add(BC.tile_nnz + t0, c.tile_nnz, c.numtile + 1, 3, -v0);
        }
        rc = run_jobs(ctx, sj);
    }
    if (!rc && keep) {  // keep the operands, the product and the value copies for re-runs
        JobList so;  // empty: each C_i's values are a view of the product's, which re-runs rewrite
        rc = upload_jobs(ctx, jv, &keep->vin);
        if (!rc) rc = upload_jobs(ctx, so, &keep->vout);
        if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(TSG_ERR_HIP, "job table upload failed");
        keep->nvin = (int64_t)jv.jobs.size();
        keep->vin_max = max_count(jv);
        keep->nvout = (int64_t)so.jobs.size();
        keep->vout_max = max_count(so);
        keep->BA = BA;
        keep->BB = BB;
        keep->BC = BC;
        if (rc) free_batch_keep(ctx, *keep);
    } else {
        if (keep) {  // failed before the hand-over
            free_m(BA);
            free_m(BB);
            free_plan(ctx, keep->nr.p);
        }
        free_c(ctx, &BC);
    }
    if (rc) {
        for (int32_t i = 0; i < n; i++) free_c(ctx, &C[i]);
        return rc;
    }
    if (stats)
        for (int32_t i = 0; i < n; i++) {
            stats[i] = bst;
            stats[i].numtileC = C[i].numtile;
            stats[i].nnzC = C[i].nnz;
            stats[i].host_syncs = bst.host_syncs + 3;  // + copy tables, boundaries, split
        }
    return TSG_OK;
}

}  // namespace tsg

extern "C" int tsg_spgemm_batched(tsg_context *c, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                                  tsg_tile_matrix *C, const tsg_options *opts, tsg_stats *stats) {
    tsg::Context *ctx = reinterpret_cast<tsg::Context *>(c);
    if (!ctx || (count > 0 && (!A || !B || !C))) return tsg::fail(TSG_ERR_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    return tsg::spgemm_batched(ctx, count, A, B, C, opts, stats);
}
