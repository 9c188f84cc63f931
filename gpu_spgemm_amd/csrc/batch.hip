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
#include <cmath>
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

// Offsets of every problem in the block-diagonal operands.
struct BatchLayout {
    int32_t n = 0, dtype = 0, flags = 0;
    bool staged = true, brow = true, scal = true;
    std::vector<int64_t> R, K, Cc, at, an, bt, bn;
};

static int batch_layout(int32_t n, const tsg_tile_matrix *A, const tsg_tile_matrix *B, BatchLayout &L) {
    L = BatchLayout{};
    L.n = n;
    L.dtype = A[0].dtype;
    L.R.assign(n + 1, 0); L.K.assign(n + 1, 0); L.Cc.assign(n + 1, 0); L.at.assign(n + 1, 0);
    L.an.assign(n + 1, 0); L.bt.assign(n + 1, 0); L.bn.assign(n + 1, 0);
    for (int32_t i = 0; i < n; i++) {
        TSG_CHECK(check_operands(&A[i], &B[i]));
        if (A[i].dtype != L.dtype) return fail(TSG_ERR_ARG, "problem %d: value types differ within the batch", i);
        L.staged = L.staged && B[i].csr_tile_nnz && B[i].csr_mask && B[i].csr_col_mask && B[i].csr_col_Ptr &&
                   B[i].csr_col_Val && B[i].row_nnz_ptr;
        L.brow = L.brow && B[i].tile_rowidx;
        L.scal = L.scal && A[i].row_ptr && A[i].row_col && A[i].row_val && B[i].row_ptr && B[i].row_col &&
                 B[i].row_val;
        L.flags |= ((A[i].flags | B[i].flags) & 1) | (B[i].flags & 2);  // repeats; long B tile-rows
        L.R[i + 1] = L.R[i] + A[i].tilem;
        L.K[i + 1] = L.K[i] + A[i].tilen;
        L.Cc[i + 1] = L.Cc[i] + B[i].tilen;
        L.at[i + 1] = L.at[i] + A[i].numtile;
        L.an[i + 1] = L.an[i] + A[i].nnz;
        L.bt[i + 1] = L.bt[i] + B[i].numtile;
        L.bn[i + 1] = L.bn[i] + B[i].nnz;
    }
    if (L.R[n] > MAX_TILE_ROWS || L.K[n] > MAX_TILE_ROWS || L.Cc[n] > MAX_TILE_ROWS)
        return fail(TSG_ERR_OVERFLOW, "batch dimensions exceed %lld tile-rows", (long long)MAX_TILE_ROWS);
    // the block-diagonal operands are addressed with int32 tile ids like any other
    if (L.at[n] > INT32_MAX || L.bt[n] > INT32_MAX)
        return fail(TSG_ERR_OVERFLOW, "the batch's operand tiles (%lld / %lld) exceed int32 tile ids",
                    (long long)L.at[n], (long long)L.bt[n]);
    return TSG_OK;
}

static void free_operand(Context *ctx, tsg_tile_matrix &M) {
    void *ps[] = {M.tile_ptr, M.tile_columnidx, M.tile_rowidx, M.tile_nnz, M.tile_csr_Val, M.tile_csr_Col,
                  M.tile_csr_Ptr, M.mask, M.csc_tile_ptr, M.csc_tile_rowidx, M.csr_tile_nnz, M.csr_mask,
                  M.csr_col_mask, M.csr_col_Ptr, M.csr_col_Val, M.row_nnz_ptr, M.row_ptr, M.row_col,
                  M.row_val};
    for (void *p : ps) pool_free(ctx, p);
    std::memset(&M, 0, sizeof M);
}

// The block-diagonal operands' descriptors and arrays.
static int alloc_operands(Context *ctx, const BatchLayout &L, tsg_tile_matrix &BA, tsg_tile_matrix &BB) {
    const int32_t n = L.n;
    const size_t vs = vsize(L.dtype);
    BA = tsg_tile_matrix{};
    BB = tsg_tile_matrix{};
    BA.m = (int32_t)(L.R[n] * 16); BA.n = (int32_t)(L.K[n] * 16); BA.tilem = (int32_t)L.R[n];
    BA.tilen = (int32_t)L.K[n]; BA.numtile = L.at[n]; BA.nnz = L.an[n]; BA.dtype = L.dtype;
    BA.location = TSG_DEVICE; BA.flags = L.flags & 1;
    BB.m = BA.n; BB.n = (int32_t)(L.Cc[n] * 16); BB.tilem = BA.tilen; BB.tilen = (int32_t)L.Cc[n];
    BB.numtile = L.bt[n]; BB.nnz = L.bn[n]; BB.dtype = L.dtype; BB.location = TSG_DEVICE; BB.is_csc = 1;
    BB.flags = L.flags;
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
    if (!rc && L.brow) rc = palloc(ctx, &BB.tile_rowidx, (size_t)ntB);
    if (!rc) rc = palloc(ctx, &BB.tile_nnz, (size_t)ntB + 1);
    if (!rc) rc = palloc(ctx, &BB.tile_csr_Ptr, (size_t)ntB * 16);
    if (!rc) rc = palloc(ctx, &BB.mask, (size_t)ntB * 16);
    if (!rc) rc = palloc(ctx, &BB.tile_csr_Col, (size_t)BB.nnz);
    if (!rc) rc = pool_alloc(ctx, &BB.tile_csr_Val, (size_t)BB.nnz * vs);
    if (!rc) rc = palloc(ctx, &BB.csc_tile_ptr, (size_t)BB.tilen + 1);
    if (!rc) rc = palloc(ctx, &BB.csc_tile_rowidx, (size_t)ntB);
    if (!rc && L.staged) {
        rc = palloc(ctx, &BB.csr_tile_nnz, (size_t)ntB + 1);
        if (!rc) rc = palloc(ctx, &BB.csr_mask, (size_t)ntB * 16);
        if (!rc) rc = palloc(ctx, &BB.csr_col_mask, (size_t)ntB * 16);
        if (!rc) rc = palloc(ctx, &BB.csr_col_Ptr, (size_t)ntB * 16);
        if (!rc) rc = pool_alloc(ctx, &BB.csr_col_Val, (size_t)BB.nnz * vs);
        if (!rc) rc = palloc(ctx, &BB.row_nnz_ptr, (size_t)BB.tilem + 1);
    }
    if (!rc && L.scal) {  // scalar CSR of the padded blocks (rows of each problem padded to its tile-rows)
        rc = palloc(ctx, &BA.row_ptr, (size_t)BA.m + 1);
        if (!rc) rc = palloc(ctx, &BA.row_col, (size_t)BA.nnz);
        if (!rc) rc = pool_alloc(ctx, &BA.row_val, (size_t)BA.nnz * vs);
        if (!rc) rc = palloc(ctx, &BB.row_ptr, (size_t)BB.m + 1);
        if (!rc) rc = palloc(ctx, &BB.row_col, (size_t)BB.nnz);
        if (!rc) rc = pool_alloc(ctx, &BB.row_val, (size_t)BB.nnz * vs);
    }
    if (rc) {
        free_operand(ctx, BA);
        free_operand(ctx, BB);
    }
    return rc;
}

// Concatenation jobs: every array (jl, closing entries of the prefix arrays
// included), and the value arrays alone (jv, for numeric re-runs).
static void concat_jobs(const BatchLayout &L, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                        const tsg_tile_matrix &BA, const tsg_tile_matrix &BB, JobList &jl, JobList &jv) {
    const int32_t n = L.n;
    const size_t vs = vsize(L.dtype);
    const int vk = vs == 8 ? 4 : 5;
    const std::vector<int64_t> &R = L.R, &K = L.K, &Cc = L.Cc, &at = L.at, &an = L.an, &bt = L.bt, &bn = L.bn;
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
        if (L.brow) jl.add(b.tile_rowidx, BB.tile_rowidx + bt[i], b.numtile, 2, K[i]);
        jl.add(b.tile_nnz, BB.tile_nnz + bt[i], b.numtile, 3, bn[i]);
        jl.add(b.tile_csr_Ptr, BB.tile_csr_Ptr + bt[i] * 16, b.numtile, 1);
        jl.add(b.mask, BB.mask + bt[i] * 16, b.numtile * 2, 1);
        jl.add(b.tile_csr_Col, BB.tile_csr_Col + bn[i], b.nnz, 0);
        jl.add(b.tile_csr_Val, (uint8_t *)BB.tile_csr_Val + bn[i] * vs, b.nnz, vk);
        jv.add(b.tile_csr_Val, (uint8_t *)BB.tile_csr_Val + bn[i] * vs, b.nnz, vk);
        jl.add(b.csc_tile_ptr, BB.csc_tile_ptr + Cc[i], b.tilen, 3, bt[i]);
        jl.add(b.csc_tile_rowidx, BB.csc_tile_rowidx + bt[i], b.numtile, 2, K[i]);
        if (L.scal) {  // rows: the problem's m rows, then its padding rows (empty: the end offset repeated)
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
        if (L.staged) {
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
    jl.fill(BA.tile_ptr + R[n], 1, at[n]);
    jl.fill(BA.tile_nnz + at[n], 1, an[n]);
    jl.fill(BB.tile_ptr + K[n], 1, bt[n]);
    jl.fill(BB.tile_nnz + bt[n], 1, bn[n]);
    jl.fill(BB.csc_tile_ptr + Cc[n], 1, bt[n]);
    if (L.staged) {
        jl.fill(BB.csr_tile_nnz + bt[n], 1, bn[n]);
        jl.fill(BB.row_nnz_ptr + K[n], 1, bn[n]);
    }
    if (L.scal) {  // closing row offsets of the scalar CSR
        jl.fill(BA.row_ptr + BA.m, 1, an[n]);
        jl.fill(BB.row_ptr + BB.m, 1, bn[n]);
    }
}

int spgemm_batched(Context *ctx, int32_t n, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                   const tsg_options *opts, tsg_stats *stats, BatchKeep *keep) {
    if (n < 0) return fail(TSG_ERR_ARG, "negative batch size");
    if (n == 0) return TSG_OK;
    BatchLayout L;
    TSG_CHECK(batch_layout(n, A, B, L));
    const int32_t dtype = L.dtype;
    const size_t vs = vsize(dtype);
    const std::vector<int64_t> &R = L.R, &Cc = L.Cc;
    tsg_tile_matrix BA{}, BB{}, BC{};
    TSG_CHECK(alloc_operands(ctx, L, BA, BB));
    JobList jl, jv;  // every array; the value arrays alone (re-run of the numeric step)
    concat_jobs(L, A, B, BA, BB, jl, jv);
    hipStream_t s = ctx->stream;
    auto free_m = [&](tsg_tile_matrix &M) { free_operand(ctx, M); };
    int rc = run_jobs(ctx, jl);
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


// ---------------------------------------------------------------- batch graph
// The whole batched product as one captured hipGraph: the concatenation
// copies, Steps 1-3 of the block-diagonal product with every size kept on the
// device (launches cover capacities taken from an exact first run times a
// slack), the per-problem split of C and one copy of a result row into pinned
// host memory.  A launch is one hipGraphLaunch and makes no host sync; the
// result call waits for it once.  A replay reads the operands' arrays again:
// their values and structure may change between launches, each problem's
// dimensions, tile count and nonzero count may not.  A product past a
// capacity is flagged on the device and computed by the eager batch instead.

// res[0] = capacity flag, res[1] = numtileC, res[2] = nnzC; res[3 + 2i] = C
// tile offset of problem i's first tile-row, res[4 + 2i] = its first value
// (i = 0..n).
__global__ void __launch_bounds__(256) k_graph_result(int32_t n, const int64_t *__restrict__ row_off,
                                                      const int64_t *__restrict__ tile_ptr,
                                                      const int64_t *__restrict__ tile_nnz, int64_t cap_tiles,
                                                      const int64_t *__restrict__ hdr,
                                                      const PlanFlags *__restrict__ flags, int64_t *__restrict__ res) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        res[0] = flags->cap;
        res[1] = hdr[0];
        res[2] = flags->nnz;
    }
    if (i > n) return;
    const int64_t t = min(tile_ptr[row_off[i]], cap_tiles);  // past the capacity: flagged, kept in bounds
    res[3 + 2 * i] = t;
    res[4 + 2 * i] = tile_nnz[t];
}

// Every problem's C offsets with the batch offsets removed, into arrays laid
// out like C's (problem i's tile offsets at ptr_out + R[i] + i, its per-tile
// arrays from its first tile t0_i, its value offsets at nnz_out + t0_i + i).
__global__ void __launch_bounds__(256)
k_graph_split(int32_t n, int32_t tilem, const int32_t *__restrict__ prob_of_row, const int64_t *__restrict__ row_off,
              const int64_t *__restrict__ col_off, const int64_t *__restrict__ res, int64_t cap_tiles,
              const int64_t *__restrict__ tile_ptr, const int32_t *__restrict__ rowidx,
              const int32_t *__restrict__ colidx, const int64_t *__restrict__ tile_nnz, int64_t *__restrict__ ptr_out,
              int32_t *__restrict__ row_out, int32_t *__restrict__ col_out, int64_t *__restrict__ nnz_out) {
    if (res[0]) return;  // past a capacity: nothing of C is defined
    const int64_t ntile = min(res[1], cap_tiles);
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < ntile; q += stride) {
        const int32_t r = rowidx[q];
        if (r < 0 || r >= tilem) continue;
        const int32_t i = prob_of_row[r];
        row_out[q] = r - (int32_t)row_off[i];
        col_out[q] = colidx[q] - (int32_t)col_off[i];
        nnz_out[q + i] = tile_nnz[q] - res[4 + 2 * i];
    }
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < tilem; r += stride) {
        const int32_t i = prob_of_row[r];
        ptr_out[r + i] = tile_ptr[r] - res[3 + 2 * i];
    }
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {  // closing entries
        const int64_t t1 = res[3 + 2 * (i + 1)];
        ptr_out[row_off[i + 1] + i] = t1 - res[3 + 2 * i];
        nnz_out[t1 + i] = tile_nnz[t1] - res[4 + 2 * i];
    }
}

struct BatchGraph {
    Context *ctx = nullptr;
    int32_t n = 0;
    BatchLayout L;
    std::vector<tsg_tile_matrix> A, B;
    tsg_options opts{};
    const tsg_options *po = nullptr;
    tsg_tile_matrix BA{}, BB{}, BC{};
    Plan p;
    DeviceCaps caps;
    CatJob *jobs = nullptr;
    int64_t njobs = 0, maxc = 0;
    int64_t *row_off = nullptr, *col_off = nullptr, *dres = nullptr, *hres = nullptr;
    int32_t *prob_of_row = nullptr;
    int64_t *ptr_out = nullptr, *nnz_out = nullptr;
    int32_t *row_out = nullptr, *col_out = nullptr;
    std::vector<void *> kept;  // arrays the captured graph goes on using
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    std::vector<tsg_tile_matrix> fallback;  // the eager batch's outputs after a short capacity
};

// The graph's work, queued on the context's stream (captured, or run once
// uncaptured to size the pool).
static int graph_body(BatchGraph &g) {
    Context *ctx = g.ctx;
    TSG_CHECK(launch_jobs(ctx, g.jobs, g.njobs, g.maxc));
    TSG_CHECK(spgemm_binned_device(ctx, &g.BA, &g.BB, &g.BC, g.po, g.caps, g.p));
    const int32_t n = g.n;
    k_graph_result<<<(unsigned)((n + 1 + 255) / 256), 256, 0, ctx->stream>>>(
        n, g.row_off, g.BC.tile_ptr, g.BC.tile_nnz, g.caps.tiles, g.p.hdr, g.p.flags, g.dres);
    TSG_HIP(hipGetLastError());
    const int64_t items = std::max<int64_t>(std::max<int64_t>(g.caps.tiles, g.BA.tilem), n);
    const unsigned grid = (unsigned)std::min<int64_t>((items + 255) / 256, 4096);
    k_graph_split<<<grid, 256, 0, ctx->stream>>>(n, g.BA.tilem, g.prob_of_row, g.row_off, g.col_off, g.dres,
                                                 g.caps.tiles, g.BC.tile_ptr, g.BC.tile_rowidx, g.BC.tile_columnidx,
                                                 g.BC.tile_nnz, g.ptr_out, g.row_out, g.col_out, g.nnz_out);
    TSG_HIP(hipGetLastError());
    TSG_HIP(hipMemcpyAsync(g.hres, g.dres, sizeof(int64_t) * (3 + 2 * ((size_t)n + 1)), hipMemcpyDeviceToHost,
                           ctx->stream));
    return TSG_OK;
}

static void free_fallback(BatchGraph &g) {
    for (tsg_tile_matrix &c : g.fallback) free_c(g.ctx, &c);
    g.fallback.clear();
}

static void graph_release(BatchGraph &g) {
    Context *ctx = g.ctx;
    (void)hipStreamSynchronize(ctx->stream);
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.done) (void)hipEventDestroy(g.done);
    free_fallback(g);
    free_c(ctx, &g.BC);
    free_plan(ctx, g.p);
    for (void *q : g.kept) pool_free(ctx, q);
    g.kept.clear();
    free_operand(ctx, g.BA);
    free_operand(ctx, g.BB);
    void *ps[] = {g.jobs, g.row_off, g.col_off, g.dres, g.prob_of_row, g.ptr_out, g.nnz_out, g.row_out, g.col_out};
    for (void *q : ps) pool_free(ctx, q);
    if (g.hres) (void)hipHostFree(g.hres);
}

// Restores the context's build-time switches on every exit.
struct DeviceSizesScope {
    Context *ctx;
    explicit DeviceSizesScope(Context *c) : ctx(c) { ctx->device_sizes = true; }
    ~DeviceSizesScope() {
        ctx->device_sizes = false;
        ctx->deferred = nullptr;
    }
};

static int graph_build(BatchGraph &g, double slack) {
    Context *ctx = g.ctx;
    const int32_t n = g.n;
    const BatchLayout &L = g.L;
    hipStream_t s = ctx->stream;
    DeviceSizesScope scope(ctx);
    TSG_CHECK(alloc_operands(ctx, L, g.BA, g.BB));
    if (g.BA.tilem <= 0) return fail(TSG_ERR_ARG, "a batch graph needs tile-rows");
    JobList jl, jv;
    concat_jobs(L, g.A.data(), g.B.data(), g.BA, g.BB, jl, jv);
    TSG_CHECK(upload_jobs(ctx, jl, &g.jobs));
    g.njobs = (int64_t)jl.jobs.size();
    g.maxc = max_count(jl);
    // exact sizes of the first product (the same bin table as the graph's)
    TSG_CHECK(launch_jobs(ctx, g.jobs, g.njobs, g.maxc));
    {
        Plan p0;
        tsg_tile_matrix C0{};
        TSG_CHECK(spgemm_binned(ctx, &g.BA, &g.BB, &C0, g.po, nullptr, &p0));
        const double sl = slack < 1.0 ? 1.0 : slack;
        g.caps.tiles = (int64_t)std::ceil((double)C0.numtile * sl) + 64;
        g.caps.nnz = (int64_t)std::ceil((double)C0.nnz * sl) + 4096;
        for (int b = 0; b < p0.nbins && b < MAX_BINS; b++)
            g.caps.bins[b] = std::min<int64_t>((int64_t)std::ceil((double)(p0.bin_off[b + 1] - p0.bin_off[b]) * sl) + 64,
                                               g.BA.tilem);
        free_plan(ctx, p0);
        free_c(ctx, &C0);
    }
    if (g.caps.tiles > INT32_MAX) return fail(TSG_ERR_OVERFLOW, "C's tile capacity exceeds int32 tile ids");
    // tables of the split and the result row
    const int32_t tilem = g.BA.tilem;
    std::vector<int32_t> pr((size_t)tilem);
    for (int32_t i = 0; i < n; i++)
        for (int64_t r = L.R[i]; r < L.R[i + 1]; r++) pr[(size_t)r] = i;
    TSG_CHECK(palloc(ctx, &g.row_off, (size_t)n + 1));
    TSG_CHECK(palloc(ctx, &g.col_off, (size_t)n + 1));
    TSG_CHECK(palloc(ctx, &g.prob_of_row, (size_t)tilem));
    TSG_CHECK(palloc(ctx, &g.dres, 3 + 2 * ((size_t)n + 1)));
    TSG_CHECK(palloc(ctx, &g.ptr_out, (size_t)tilem + n + 1));
    TSG_CHECK(palloc(ctx, &g.nnz_out, (size_t)g.caps.tiles + n + 1));
    TSG_CHECK(palloc(ctx, &g.row_out, (size_t)std::max<int64_t>(g.caps.tiles, 1)));
    TSG_CHECK(palloc(ctx, &g.col_out, (size_t)std::max<int64_t>(g.caps.tiles, 1)));
    TSG_HIP(hipHostMalloc(reinterpret_cast<void **>(&g.hres), sizeof(int64_t) * (3 + 2 * ((size_t)n + 1))));
    TSG_HIP(hipMemcpyAsync(g.row_off, L.R.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
    TSG_HIP(hipMemcpyAsync(g.col_off, L.Cc.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
    TSG_HIP(hipMemcpyAsync(g.prob_of_row, pr.data(), sizeof(int32_t) * tilem, hipMemcpyHostToDevice, s));
    TSG_HIP(hipStreamSynchronize(s));
    // one uncaptured run: it checks the capacities and leaves the pool holding
    // blocks of every size the capture asks for (no allocation while capturing)
    std::vector<void *> warm;
    ctx->deferred = &warm;
    int rc = graph_body(g);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(TSG_ERR_HIP, "batch graph warm-up failed");
    if (!rc && g.hres[0]) rc = fail(TSG_ERR_OVERFLOW, "batch graph: the first product exceeds its own capacities");
    ctx->deferred = nullptr;
    free_c(ctx, &g.BC);
    free_plan(ctx, g.p);
    for (void *q : warm) pool_free(ctx, q);
    TSG_CHECK(rc);
    // capture
    ctx->deferred = &g.kept;
    TSG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    rc = graph_body(g);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    ctx->deferred = nullptr;
    g.graph = graph;
    if (!rc && ec != hipSuccess) rc = fail(TSG_ERR_HIP, "capturing the batch graph failed: %s", hipGetErrorString(ec));
    TSG_CHECK(rc);
    TSG_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    TSG_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    return TSG_OK;
}

}  // namespace tsg

struct tsg_batch_graph {
    tsg::BatchGraph g;
};

extern "C" int tsg_batch_graph_create(tsg_context *c, int32_t count, const tsg_tile_matrix *A,
                                      const tsg_tile_matrix *B, const tsg_options *opts, double slack,
                                      tsg_batch_graph **out) {
    using namespace tsg;
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || !out || count <= 0 || !A || !B) return fail(TSG_ERR_ARG, "bad arguments");
    *out = nullptr;
    if (opts && opts->timing) return fail(TSG_ERR_ARG, "a batch graph takes no per-step timing");
    (void)hipSetDevice(ctx->device);
    tsg_batch_graph *h = new tsg_batch_graph;
    BatchGraph &g = h->g;
    g.ctx = ctx;
    g.n = count;
    g.A.assign(A, A + count);
    g.B.assign(B, B + count);
    if (opts) {
        g.opts = *opts;
        g.po = &g.opts;
    }
    int rc = batch_layout(count, A, B, g.L);
    if (!rc) rc = graph_build(g, slack);
    if (rc) {
        graph_release(g);
        delete h;
        return rc;
    }
    *out = h;
    return TSG_OK;
}

extern "C" int tsg_batch_graph_launch(tsg_batch_graph *h) {
    using namespace tsg;
    if (!h) return fail(TSG_ERR_ARG, "null graph");
    BatchGraph &g = h->g;
    (void)hipSetDevice(g.ctx->device);
    TSG_HIP(hipGraphLaunch(g.exec, g.ctx->stream));
    TSG_HIP(hipEventRecord(g.done, g.ctx->stream));
    g.launched = true;
    return TSG_OK;
}

extern "C" int tsg_batch_graph_result(tsg_batch_graph *h, tsg_tile_matrix *C, tsg_stats *stats) {
    using namespace tsg;
    if (!h || !C) return fail(TSG_ERR_ARG, "null argument");
    BatchGraph &g = h->g;
    if (!g.launched) return fail(TSG_ERR_ARG, "the batch graph has not been launched");
    Context *ctx = g.ctx;
    (void)hipSetDevice(ctx->device);
    TSG_HIP(hipEventSynchronize(g.done));
    const int32_t n = g.n;
    const int64_t *r = g.hres;
    free_fallback(g);
    if (r[0]) {  // a capacity was short: the exact eager batch
        g.fallback.assign((size_t)n, tsg_tile_matrix{});
        TSG_CHECK(spgemm_batched(ctx, n, g.A.data(), g.B.data(), g.fallback.data(), g.po, stats));
        for (int32_t i = 0; i < n; i++) {
            C[i] = g.fallback[(size_t)i];
            if (stats) stats[i].fallback_used = 1;
        }
        return TSG_OK;
    }
    const size_t vs = vsize(g.L.dtype);
    for (int32_t i = 0; i < n; i++) {
        const int64_t t0 = r[3 + 2 * i], v0 = r[4 + 2 * i], t1 = r[5 + 2 * i], v1 = r[6 + 2 * i];
        tsg_tile_matrix &c = C[i];
        std::memset(&c, 0, sizeof c);
        c.m = g.A[(size_t)i].m; c.n = g.B[(size_t)i].n; c.tilem = g.A[(size_t)i].tilem;
        c.tilen = g.B[(size_t)i].tilen; c.dtype = g.L.dtype; c.location = TSG_DEVICE;
        c.numtile = t1 - t0;
        c.nnz = v1 - v0;
        c.tile_ptr = g.ptr_out + g.L.R[(size_t)i] + i;
        c.tile_rowidx = g.row_out + t0;
        c.tile_columnidx = g.col_out + t0;
        c.tile_nnz = g.nnz_out + t0 + i;
        c.mask = g.BC.mask + t0 * 16;
        c.tile_csr_Ptr = g.BC.tile_csr_Ptr + t0 * 16;
        c.tile_csr_Val = static_cast<uint8_t *>(g.BC.tile_csr_Val) + v0 * vs;
        c.tile_csr_Col = g.BC.tile_csr_Col + v0;
        if (stats) {
            stats[i] = tsg_stats{};
            stats[i].numtileC = c.numtile;
            stats[i].nnzC = c.nnz;
            stats[i].host_syncs = 1;  // the wait for the launch
        }
    }
    return TSG_OK;
}

extern "C" int tsg_batch_graph_capacity(const tsg_batch_graph *h, int64_t *cap_tiles, int64_t *cap_nnz) {
    if (!h) return tsg::fail(TSG_ERR_ARG, "null graph");
    if (cap_tiles) *cap_tiles = h->g.caps.tiles;
    if (cap_nnz) *cap_nnz = h->g.caps.nnz;
    return TSG_OK;
}

extern "C" void tsg_batch_graph_destroy(tsg_batch_graph *h) {
    if (!h) return;
    (void)hipSetDevice(h->g.ctx->device);
    tsg::graph_release(h->g);
    delete h;
}


extern "C" int tsg_spgemm_batched(tsg_context *c, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                                  tsg_tile_matrix *C, const tsg_options *opts, tsg_stats *stats) {
    tsg::Context *ctx = reinterpret_cast<tsg::Context *>(c);
    if (!ctx || (count > 0 && (!A || !B || !C))) return tsg::fail(TSG_ERR_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    return tsg::spgemm_batched(ctx, count, A, B, C, opts, stats);
}
