/*
 * tsg.h -- C ABI of the MI355X tiled SpGEMM library (libtsg.so).
 *
 * C = A*B in the 16x16 tile format of the reference (src/tile_format.h:15-42),
 * computed by hand-written HIP kernels for gfx950.  Plain pointers and sizes
 * only; every call returns TSG_OK (0) or a negative tsg_status, and
 * tsg_last_error_string() describes the last failure of the calling thread.
 * No exception crosses this boundary.
 *
 * The reference has no library API: its path is the host program
 * src/spgemm_raw_v2.cu:61-372 (main) driving host conversion
 * (src/tile_format.h:93-222, :424-546), upload (src/spgemm_raw_v2.cu:25-48)
 * and three kernel steps.  Each entry point below names the reference code it
 * replaces.
 *
 * Layout differences from the reference, all widening only:
 *   - prefix arrays (tile_ptr, csc_tile_ptr, tile_nnz, rowptr) are int64, so
 *     nnz(C) above 2^31 is representable (R-MAT scale 20 reaches 9.7e9);
 *   - values are double (TSG_F64) or float (TSG_F32), chosen at run time
 *     instead of the compile-time TVAL typedef (src/tile_format.h:13).
 */
#ifndef TSG_H
#define TSG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSG_TILE 16  /* src/tile_format.h:11 TILE_SIZE */

typedef enum tsg_status {
    TSG_OK = 0,
    TSG_ERR_ARG = -1,         /* invalid argument / inconsistent shapes      */
    TSG_ERR_HIP = -2,         /* HIP runtime error                          */
    TSG_ERR_OOM = -3,         /* device or host allocation failed            */
    TSG_ERR_OVERFLOW = -4,    /* size beyond a documented limit              */
    TSG_ERR_UNSUPPORTED = -5, /* unsupported input (e.g. unsorted CSR rows)   */
    TSG_ERR_IO = -6,          /* file open / parse failure                   */
    TSG_ERR_NODEVICE = -7     /* no GPU visible                               */
} tsg_status;

typedef enum tsg_dtype { TSG_F64 = 0, TSG_F32 = 1 } tsg_dtype;
typedef enum tsg_location { TSG_HOST = 0, TSG_DEVICE = 1 } tsg_location;

/* Compressed sparse rows.  Column indices sorted ascending within each row
 * (what load_mtx, the transposes and every generator produce). */
typedef struct tsg_csr {
    int32_t m, n;
    int64_t nnz;
    int32_t dtype;     /* tsg_dtype */
    int32_t location;  /* tsg_location */
    int64_t *rowptr;   /* [m+1] */
    int32_t *col;      /* [nnz] */
    void *val;         /* [nnz] double or float */
} tsg_csr;

/* Tile matrix: the reference TileMatrix (src/tile_format.h:15-42) and the
 * device mirror TileDevPtrs (src/spgemm_raw_v2.cu:15-23) in one struct.
 * For a B operand (is_csc = 1) tile_ptr/tile_columnidx keep the CSR-tile
 * order while tile_nnz, tile_csr_Ptr, mask, tile_csr_Col and tile_csr_Val are
 * in CSC-tile order (csc_tile_ptr / csc_tile_rowidx), and tile_csr_Col holds
 * the local column only (src/tile_format.h:535-542).  For A and C the col
 * byte is (row<<4)|col (src/tile_format.h:208). */
typedef struct tsg_tile_matrix {
    int32_t m, n;
    int32_t tilem, tilen;
    int64_t nnz;
    int64_t numtile;
    int32_t dtype;            /* tsg_dtype */
    int32_t location;         /* tsg_location */
    int32_t is_csc;           /* 1 for a B operand laid out by tsg_build_csc_tiles */
    int32_t flags;            /* bit 0: some row holds a column twice (duplicates kept);
                                 bit 1: some tile-row holds more than 64 tiles (set by
                                 tsg_csr2tile; products then keep bitmaps of such B tile-rows);
                                 bit 2: bit 1 is known (tsg_csr2tile checked every tile-row) */
    int64_t *tile_ptr;        /* [tilem+1] */
    int32_t *tile_columnidx;  /* [numtile] */
    int32_t *tile_rowidx;     /* [numtile] row of each tile (C output), may be NULL for A/B */
    int64_t *tile_nnz;        /* [nt+1] exclusive prefix; nt = numtile */
    void *tile_csr_Val;       /* [nnz] */
    uint8_t *tile_csr_Col;    /* [nnz] */
    uint8_t *tile_csr_Ptr;    /* [nt*16] exclusive row offsets inside each tile */
    uint16_t *mask;           /* [nt*16] bit (15-c) = local column c present */
    int64_t *csc_tile_ptr;    /* [tilen+1] (B only) */
    int32_t *csc_tile_rowidx; /* [numtile] (B only) */
    /* B only, optional: the same tiles in CSR-tile order, kept by
     * tsg_build_csc_tiles so that a B tile-row is one contiguous range for the
     * row-staged kernels, with each tile's values stored column by column.
     * NULL when absent (the C-tile kernels are used). */
    int64_t *csr_tile_nnz;    /* [numtile+1] exclusive prefix (CSR-tile order) */
    uint16_t *csr_mask;       /* [numtile*16] row masks, bit (15-c) = column c */
    uint16_t *csr_col_mask;   /* [numtile*16] column masks, bit (15-r) = row r */
    uint8_t *csr_col_Ptr;     /* [numtile*16] exclusive column offsets inside each tile */
    void *csr_col_Val;        /* [nnz] values, column-major inside each tile */
    int64_t *row_nnz_ptr;     /* [tilem+1] first payload entry of each tile-row (= CSR rowptr[16*i]) */
    /* Optional: the same matrix in scalar CSR (columns ascending in each row),
     * kept by tsg_csr2tile.  Heavy tile-rows whose tiles are sparse run a
     * row-by-row product on it; NULL when absent (C outputs). */
    int64_t *row_ptr;         /* [m+1] */
    int32_t *row_col;         /* [nnz] */
    void *row_val;            /* [nnz] */
    /* Optional (kept by tsg_csr2tile / tsg_build_csc_tiles): where each CSR
     * entry's value lives, so tsg_tile_set_values can refresh every copy. */
    int64_t *val_perm;        /* [nnz] position in tile_csr_Val */
    int64_t *val_perm_col;    /* [nnz] B only: position in csr_col_Val */
    /* Optional (tsg_build_tile_csc): every tile's entries column by column,
     * rows ascending, in the tile order of the data arrays above.  For a B
     * operand these are the reference's tile_csc_Ptr/Row/Val
     * (src/tile_format.h:35-37, built at :547-584); for row tiles (A) its
     * tile_a_csc_Ptr/Row/Val (:39-41, built by build_a_tile_csc :589-627).
     * NULL when absent. */
    uint8_t *tile_csc_Ptr;    /* [nt*16] exclusive column offsets inside each tile */
    uint8_t *tile_csc_Row;    /* [nnz] local row of each entry */
    void *tile_csc_Val;       /* [nnz] values in that order */
    /* Rows and columns per tile: 0 = 16 (TILE_SIZE, src/tile_format.h:11),
     * else 1-15 for tiles made by tsg_csr2tile_ts.  Per-tile row arrays
     * (tile_csr_Ptr, mask) then hold tile_size entries per tile, and only the
     * conversion entry points (tsg_build_csc_tiles, transfers, free) accept
     * such a matrix. */
    int32_t tile_size;
    int32_t reserved_ts;
} tsg_tile_matrix;

typedef struct tsg_context tsg_context;

/* Scheduler choice for tsg_spgemm. */
typedef enum tsg_scheduler {
    TSG_SCHED_PLAIN = 0,    /* the reference's 3-step method: exact counts, host reads of numtileC / nnzC */
    TSG_SCHED_LEARNED = 1   /* exact structure, MLP-predicted value capacity: no read between the symbolic
                               and numeric steps; exact numeric re-run past the capacity */
} tsg_scheduler;

typedef struct tsg_options {
    int32_t scheduler;      /* tsg_scheduler */
    int32_t numeric_variant;/* 0 = auto: tile-rows binned to the row-staged kernels, heavy rows to the
                               row-by-row or C-tile kernels -- or, when both operands' tiles hold 64
                               entries or more on average and no tuned configuration says otherwise,
                               variant 2 with its default gate; 1 = C-tile kernels for every tile (the
                               reference's method); 2 = C-tile kernels with matrix cores (MFMA) for dense
                               tile pairs; 3 = C-tile kernels in outer-product form over A's
                               within-tile CSC (needs tsg_build_tile_csc on A; plain scheduler) */
    int32_t mfma_min_products; /* variant 2: a pair goes to the matrix cores when nnz(A tile) * nnz(B tile)
                                  / 16 (its expected product count) >= this (0 = 256) */
    int32_t timing;         /* 1 = record per-step hipEvent times in tsg_stats */
    double  pred_scale;     /* learned: scale of the predicted capacities (0 or 1.0 = normal; below 1 drops the
                               margin, a test hook that forces the exact fallback) */
    int32_t tuning;         /* 0 = use the context's autotuned configuration for this operand class when
                               tsg_autotune stored one; 1 = built-in defaults */
    int32_t sum_order;      /* 0 = every output adds its products in ascending k, one FMA each (the
                               reference's tiled order, src/spgemm_raw_kernels_v2.cuh:83-94): values are
                               bit-identical across kernels and runs; 1 = the row-by-row numeric kernel
                               (heavy sparse tile-rows) adds products in any order with LDS atomics:
                               faster, values equal to rounding (not bit-reproducible) */
} tsg_options;

typedef struct tsg_stats {
    int64_t numtileC;       /* tiles of C, including structurally empty ones */
    int64_t nnzC;
    uint64_t products;      /* sum_{(i,k) in A} nnz(B[k,:]) (src/spgemm_raw_v2.cu:95-99) */
    double ms_step1, ms_step2, ms_step3, ms_total;
    int32_t fallback_used;  /* learned: 1 when the exact path re-ran after an overflow, 2 when the
                               predicted layout exceeded int32 tile ids or free device memory */
    int32_t host_syncs;     /* blocking device-to-host reads made inside the call */
    int64_t pred_numtileC, pred_nnzC;  /* learned: predicted C tiles; C's value capacity */
    int32_t pred_reused;    /* learned: 1 when the predictions were this operand pair's previous ones */
    int32_t numeric_kernels; /* binned pipeline: numeric kernels launched, bit 0 the row-lane kernel
                                (k_num_rows), 1-3 the staged kernels S, W, M (k_num_staged), 4 the
                                C-tile list kernel (k_numeric_list), 5 the row-by-row kernel
                                (k_rgw_numeric); 0 elsewhere */
} tsg_stats;

/* ---- context ---- */
int tsg_create(int device, tsg_context **ctx);
int tsg_destroy(tsg_context *ctx);
const char *tsg_last_error_string(void);
int tsg_device_count(int *count);
int tsg_synchronize(tsg_context *ctx);
/* Device name (src/spgemm_raw_v2.cu:70-71). */
int tsg_device_name(tsg_context *ctx, char *buf, size_t len);
/* The HIP stream every call of this context runs on (hipStream_t). */
void *tsg_stream(tsg_context *ctx);
/* Runs every later call of this context on the caller's hipStream_t (NULL:
 * back to the context's own stream), so library work orders with the caller's
 * own kernels and copies.  The reference queues everything on the legacy
 * default stream (src/spgemm_raw_v2.cu:120-129, :204-240).  Work already queued
 * on the previous stream is waited for first; the caller keeps the stream
 * alive while the context uses it.  Calls that return host values (sizes,
 * downloads) still synchronize that stream. */
int tsg_set_stream(tsg_context *ctx, void *stream);
/* Bytes held by the context's device pool (in use, cached). */
int tsg_pool_stats(tsg_context *ctx, int64_t *in_use, int64_t *cached);
int tsg_pool_trim(tsg_context *ctx);
/* Raw device arrays from the context's pool (for caller-owned outputs, e.g. of
 * tsg_spgemm_noalloc), and a blocking copy between any host and device
 * addresses on the context's stream (hipMemcpyDefault).  The reference uses
 * cudaMalloc / cudaMemcpy directly (src/spgemm_raw_v2.cu:25-48). */
int tsg_device_malloc(tsg_context *ctx, size_t bytes, void **ptr);
int tsg_device_free(tsg_context *ctx, void *ptr);
int tsg_memcpy(tsg_context *ctx, void *dst, const void *src, size_t bytes);

/* In-place inclusive scan of n int64 values in device memory, on the
 * context's stream (asynchronous): the library's replacement for CUB
 * DeviceScan::ExclusiveSum (src/spgemm_raw_v2.cu:131-138, :202-208).  With 0 in
 * data[0] (or zero_first = 1, which takes data[0] as 0 whatever it holds) and
 * counts in data[1..n-1] it leaves exclusive offsets plus the total in
 * data[n-1]; total (device, may be NULL) also receives data[n-1]. */
int tsg_scan_i64(tsg_context *ctx, int64_t *data, int64_t n, int32_t zero_first, int64_t *total);

/* ---- host matrices (generators, MTX I/O) ---- */
/* MTX reader, replaces load_mtx (src/tile_format.h:51-90): symmetric mirrored,
 * pattern -> 1.0, complex -> real part, sorted by (row,col), duplicates kept. */
int tsg_load_mtx(const char *path, tsg_csr *out_host);
int tsg_write_mtx(const char *path, const tsg_csr *host);
/* B = A^T (src/spgemm_raw_v2.cu:78-89, the --aat path). */
int tsg_csr_transpose_host(const tsg_csr *in, tsg_csr *out_host);
/* flops of A*B (src/spgemm_raw_v2.cu:95-99). */
int tsg_csr_products(const tsg_csr *A_host, const tsg_csr *B_host, uint64_t *products);
int tsg_csr_free_host(tsg_csr *h);
int tsg_csr_convert_dtype_host(tsg_csr *h, int32_t dtype);

/* Synthetic inputs (no reference counterpart: the reference reads .mtx files).
 * All use a counter-based SplitMix64 stream, so results do not depend on threads. */
int tsg_gen_er(int32_t n, int32_t per_row, uint64_t seed, tsg_csr *out);          /* exactly per_row distinct uniform columns, U[-1,1] */
int tsg_gen_er_var(int32_t n, int32_t lo, int32_t hi, uint64_t seed, tsg_csr *out); /* nnz/row uniform in [lo,hi] */
int tsg_gen_laplace2d(int32_t nx, int32_t ny, tsg_csr *out);                      /* 5-point, 4/-1 */
int tsg_gen_laplace3d(int32_t nx, int32_t ny, int32_t nz, tsg_csr *out);          /* 7-point, 6/-1 */
int tsg_gen_stencil27(int32_t nx, int32_t ny, int32_t nz, tsg_csr *out);          /* 27-point, 26/-1 */
int tsg_gen_aggregation(int32_t nx, int32_t ny, int32_t nz, tsg_csr *out);        /* 2x2x2 piecewise-constant P */
int tsg_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                 int32_t dtype, tsg_csr *out);                                     /* duplicates merged, U(0,1] */

/* ---- device memory ---- */
int tsg_csr_upload(tsg_context *ctx, const tsg_csr *host, tsg_csr *dev);
int tsg_csr_download(tsg_context *ctx, const tsg_csr *dev, tsg_csr *host);
int tsg_csr_free(tsg_context *ctx, tsg_csr *dev);
/* Host<->device copies of a tile matrix; replaces upload_tiles
 * (src/spgemm_raw_v2.cu:25-48).  Download allocates host arrays with malloc. */
int tsg_tile_upload(tsg_context *ctx, const tsg_tile_matrix *host, tsg_tile_matrix *dev);
int tsg_tile_download(tsg_context *ctx, const tsg_tile_matrix *dev, tsg_tile_matrix *host);
/* Tile-rows [tr_lo, tr_hi) of a row-tiled matrix (A or C) to the host, as a
 * matrix of tr_hi - tr_lo tile-rows: tile_ptr, tile_nnz and tile_rowidx
 * rebased to the range, everything else as in tsg_tile_download (malloc'ed,
 * free with tsg_tile_free_host).  No reference counterpart: the reference
 * copies all of C back to check it (src/spgemm_raw_v2.cu:253-266), which at
 * R-MAT scale 20 (162 GB of C) does not fit; checks then run on row blocks. */
int tsg_tile_download_rows(tsg_context *ctx, const tsg_tile_matrix *dev, int32_t tr_lo, int32_t tr_hi,
                           tsg_tile_matrix *host);
/* Frees device arrays (free_dev, src/spgemm_raw_v2.cu:50-59). */
int tsg_tile_free(tsg_context *ctx, tsg_tile_matrix *dev);
/* Frees host arrays (free_tile_matrix, src/tile_format.h:647-659). */
int tsg_tile_free_host(tsg_tile_matrix *host);

/* ---- conversion on the GPU ---- */
/* CSR -> row tiles; replaces csr2tile (src/tile_format.h:93-222). */
int tsg_csr2tile(tsg_context *ctx, const tsg_csr *dev_csr, tsg_tile_matrix *out_dev);
/* Re-lays a row-tiled matrix's data arrays in CSC-tile order, making it a B
 * operand; replaces build_csc_tiles (src/tile_format.h:424-546).  Works on any
 * row-tiled matrix, including a C output (so products chain without host work).
 * Asynchronous on the context's stream (no host value is returned).  The row
 * masks and row offsets of 16 x 16 tiles must start on 16-byte boundaries, as
 * every array this library allocates does. */
int tsg_build_csc_tiles(tsg_context *ctx, tsg_tile_matrix *B_dev);
/* New values for a matrix made by tsg_csr2tile (and tsg_build_csc_tiles), same
 * structure: vals = nnz device values in the CSR order of the original input;
 * every value copy (tiles, column-major copy, scalar CSR) is rewritten.  No
 * reference counterpart (the reference rebuilds the tiles on the host). */
int tsg_tile_set_values(tsg_context *ctx, tsg_tile_matrix *dev, const void *vals);
/* CSR -> tiles of ts x ts (1 <= ts <= 16) on the device; replaces csr2tile_ts
 * (src/tile_format.h:225-324): mask bit ts-1-c is local column c, col byte
 * (row<<4)|col, ts row offsets and masks per tile.  ts = 16 is tsg_csr2tile.
 * tsg_build_csc_tiles on the result replaces build_csc_tiles_ts (:327-417).
 * No product runs on ts != 16 (neither does the reference's). */
int tsg_csr2tile_ts(tsg_context *ctx, const tsg_csr *csr_dev, int32_t ts, tsg_tile_matrix *out_dev);
/* Within-tile CSC of every tile (tile_csc_Ptr/Row/Val above), on the
 * device; replaces the within-tile transpose of build_csc_tiles
 * (src/tile_format.h:547-584) for a B operand and build_a_tile_csc
 * (src/tile_format.h:589-627) for row tiles.  Entries of a column keep row
 * order, and repeated columns of a row (flags bit 0) keep storage order, as
 * the reference's counting scatter does.  Rebuilding replaces the arrays;
 * tsg_tile_set_values refreshes tile_csc_Val. */
int tsg_build_tile_csc(tsg_context *ctx, tsg_tile_matrix *dev);
/* Per A tile-row and per B tile column, a bitmap over the k tile index:
 * replaces build_bitmasks (src/tile_format.h:628-645), which the reference's
 * driver does not call.  *bm_len = ceil(A.tilen / 32) words per row; bit
 * 31 - k % 32 of word k / 32 is set when A's tile-row holds a tile in tile
 * column k (bmA, tilemA * bm_len words) or B's tile column holds a tile in
 * tile row k (bmB, tilenB * bm_len words); AND-ing them finds the matched k.
 * A: row tiles, B: a B operand (tsg_build_csc_tiles); bmA / bmB are HOST
 * arrays the caller sizes (NULL: skipped), as the reference's are. */
int tsg_build_bitmasks(tsg_context *ctx, const tsg_tile_matrix *A_dev, const tsg_tile_matrix *B_dev,
                       uint32_t *bmA, uint32_t *bmB, int32_t *bm_len);
/* Row tiles -> CSR (the reference's verification unpacking, src/spgemm_raw_v2.cu:268-284). */
int tsg_tile2csr(tsg_context *ctx, const tsg_tile_matrix *dev, tsg_csr *out_dev);

/* ---- SpGEMM ---- */
/* C = A*B; replaces Steps 1-3 of src/spgemm_raw_v2.cu:111-240 (kernels in
 * src/spgemm_raw_kernels.cuh:23-215 and src/spgemm_raw_kernels_v2.cuh:10-139).
 * A: row tiles; B: CSC-laid tiles (tsg_build_csc_tiles); C allocated by the
 * library in the context pool (free with tsg_tile_free).  opts/stats may be NULL. */
int tsg_spgemm(tsg_context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
               tsg_tile_matrix *C, const tsg_options *opts, tsg_stats *stats);

/* C = A*B with no device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The exact binned
 * pipeline of tsg_spgemm (plain scheduler, numeric_variant 0; timing ignored)
 * with its work arrays in a static device scratch array of 64 MiB per GPU,
 * held by the library until the process ends, and C written into the
 * caller's device arrays, which C carries in: tile_ptr [A->tilem + 1],
 * tile_rowidx and tile_columnidx [cap_tiles], tile_nnz [cap_tiles + 1], mask and
 * tile_csr_Ptr [cap_tiles * 16], tile_csr_Val and tile_csr_Col [cap_nnz] (values
 * of A's dtype).  On success C's sizes are set and its arrays stay the
 * caller's (do not pass C to tsg_tile_free).  TSG_ERR_OVERFLOW when C exceeds a
 * capacity (stats->numtileC, and once the tiles fit stats->nnzC, give the sizes
 * needed) or when A's work arrays exceed the scratch (about 24 bytes per A
 * tile plus 10 per tile-row; e.g. the 2048^2 5-point Laplacian needs 33 MB).
 * Heavy tile-rows take the C-tile kernels here, never the row-by-row ones
 * (whose cursor area does not fit): results are the same bits.  Every call on
 * one GPU shares its scratch: calls that overlap on one GPU race. */
int tsg_spgemm_noalloc(tsg_context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, tsg_tile_matrix *C,
                       int64_t cap_tiles, int64_t cap_nnz, const tsg_options *opts, tsg_stats *stats);

/* ---- Learned scheduler (no reference counterpart: the reference sizes C with
 * exact counts and two blocking reads, numtileC after Step 1 and nnzC after
 * Step 2, src/spgemm_raw_v2.cu:119-141, :191-211) ----
 * An MLP (8 features -> 32 -> 32 -> 2, ReLU, FP32; tsg_learned_num_params()
 * weights: W1[8][32], b1[32], W2[32][32], b2[32], W3[32][2], b3[2], row-major,
 * x @ W) predicts log1p of each C tile-row's tile count and nonzero count from
 * per-tile-row features of A*B that the Step-1 count emits: log1p of (A tiles,
 * B tiles reached, A nonzeros, B nonzeros reached, span of B tile columns,
 * longest B tile-row reached, summed B tile-row spans, nonempty B tile-rows
 * reached).  With scheduler TSG_SCHED_LEARNED, tsg_spgemm keeps the exact
 * Step-1 count (C's tile arrays are exact) and sizes C's value arrays from the
 * calibrated sum of the predicted nonzeros, so no host read separates the
 * symbolic and numeric steps (stats->pred_nnzC: the capacity; pred_numtileC:
 * the predicted tiles).  The numeric kernels write nothing past the capacity;
 * when nnz(C) exceeds it, the numeric step runs again into exact arrays
 * (stats->fallback_used = 1; 2 when the predicted arrays did not fit in
 * memory): the result is always the exact product. */
int tsg_learned_num_params(void);
/* Appends (features, exact targets) of every tile-row of A*B to the context's
 * training set (runs the exact product once). */
int tsg_learned_collect(tsg_context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B);
int tsg_learned_clear(tsg_context *ctx);
/* Adam (beta 0.9/0.999, eps 1e-8) on the mean squared error over the training
 * set, shuffled minibatches of `batch`; weights start Glorot-uniform from `seed`
 * when none are set.  final_loss: mean loss of the last epoch (may be NULL). */
int tsg_learned_train(tsg_context *ctx, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                      float *final_loss);
int tsg_learned_get_weights(tsg_context *ctx, float *w, int32_t n);
/* Log-space offsets (tiles, nonzeros) applied to the summed predictions before
 * C is sized; training sets each to log 1.02 + the largest log(actual /
 * predicted total) over the collected products, clamped to [log 1.02, log 16].
 * get and/or set: 2 floats each (may be NULL). */
int tsg_learned_calibration(tsg_context *ctx, float *get, const float *set);
int tsg_learned_set_weights(tsg_context *ctx, const float *w, int32_t n);
/* Host arrays: feats tilem x 8, pred tilem x 2 (expm1 of the network's outputs:
 * predicted tiles, nonzeros). */
int tsg_learned_features(tsg_context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, float *feats);
int tsg_learned_predict(tsg_context *ctx, const tsg_tile_matrix *A, const tsg_tile_matrix *B, float *pred);
/* The network's GPU forward pass and mean-squared-error gradient on host arrays
 * X (n x 8), T (n x 2 log1p targets): Y (n x 2 raw outputs), grad (num_params). */
int tsg_learned_forward(tsg_context *ctx, const float *X, int64_t n, float *Y);
int tsg_learned_gradient(tsg_context *ctx, const float *X, const float *T, int64_t n, float *grad,
                         float *loss);

/* ---- Autotuner (no reference counterpart) ----
 * A configuration chooses which row-staged kernel variants may take rows
 * (var_mask bit v: variant v, v = 0 small, 1 wide, 2 many-tiles, 3 small
 * tile-rows of short scalar rows whose steps run 16 lanes per tile-row and a
 * lane per scalar row -- with fewer than 32,768 such tile-rows they run
 * variant 0's kernels unless bit 4 is set; the default admits all four) and the A
 * tile density (entries per A tile) at or below which a heavy tile-row runs
 * the row-by-row kernels (rg_density; 0 = never), or, with
 * mfma_min_products > 0, the C-tile kernels with matrix cores for every tile
 * (numeric_variant 2 with that density gate).  tsg_autotune times every
 * configuration on the given product (count = 1) or batch (count > 1, as
 * tsg_spgemm_batched), median of `reps` runs, and stores the fastest for the
 * operands' class (log2 buckets of tile-rows, A tiles per tile-row, A entries
 * per tile, B tiles per tile-row, value type); later products of that class
 * use it unless opts->tuning = 1.  Results are identical under every
 * configuration: only speed changes. */
typedef struct tsg_tune_entry {
    int32_t var_mask;
    int32_t rg_density;
    int32_t mfma_min_products;  /* > 0: matrix-core C-tile kernels with this gate */
    int32_t sampled;         /* report only: 1 = timed on a sample of the batch (every 7th product) and
                                not among the fastest there, so not timed on the whole batch */
    double ms;               /* median wall time of one product (or batch), host syncs included */
} tsg_tune_entry;
int tsg_autotune(tsg_context *ctx, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                 int32_t reps, tsg_tune_entry *report, int32_t cap, int32_t *nreport, tsg_tune_entry *best);
/* The stored configuration for the operands' class: returns 1 and fills *cfg
 * when one exists, 0 when not (count as in tsg_autotune). */
int tsg_autotune_query(tsg_context *ctx, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                       tsg_tune_entry *cfg);
/* Stores cfg (var_mask, rg_density, mfma_min_products; ms ignored) for the operands' class. */
int tsg_autotune_set(tsg_context *ctx, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                     const tsg_tune_entry *cfg);
int tsg_autotune_clear(tsg_context *ctx);

/* ---- Plans: same structure, new values (no reference counterpart) ----
 * tsg_plan_create computes C_i = A_i * B_i (count = 1: as tsg_spgemm; count >
 * 1: as tsg_spgemm_batched, C an array of count outputs) and keeps what the
 * numeric step needs.  tsg_plan_execute recomputes every C_i's values from the
 * operands' current values (change them with tsg_tile_set_values) without
 * touching C's structure: the first call captures the launches -- value copies
 * into the batch's block-diagonal operands, the numeric kernels, copies out --
 * into a hipGraph, later calls replay it (one graph launch, no host sync, no
 * allocation; asynchronous on the context's stream).  Operands and outputs must
 * keep their arrays while the plan lives; C_i stay the caller's (tsg_tile_free).
 * Plans use the exact pipeline (the learned scheduler re-plans every call). */
typedef struct tsg_plan tsg_plan;
int tsg_plan_create(tsg_context *ctx, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                    tsg_tile_matrix *C, const tsg_options *opts, tsg_plan **plan);
int tsg_plan_execute(tsg_plan *plan);
int tsg_plan_destroy(tsg_plan *plan);

/* Many independent products in one pipeline run (no reference counterpart):
 * the batch is one product of block-diagonal operands; C is an array of count
 * outputs, stats (may be NULL) an array of count entries.  Each C_i is freed
 * with tsg_tile_free as usual; its mask, tile_csr_Ptr, tile_csr_Val and
 * tile_csr_Col lie inside allocations shared by the batch's outputs, which
 * return to the context's pool when the last C_i holding them is freed. */
int tsg_spgemm_batched(tsg_context *ctx, int32_t count, const tsg_tile_matrix *A,
                       const tsg_tile_matrix *B, tsg_tile_matrix *C, const tsg_options *opts,
                       tsg_stats *stats);

/* ---- Batch graphs: a whole batched product in one launch (no reference
 * counterpart; the reference computes one product per process with host
 * syncs between its steps, src/spgemm_raw_v2.cu:111-240) ----
 * tsg_batch_graph_create captures C_i = A_i * B_i for the count problems into
 * one hipGraph: the copies into the block-diagonal operands, Steps 1-3 with
 * every size kept on the device, the split into per-problem C_i and one copy
 * of a result row into pinned host memory.  C's arrays and the rows each bin's
 * kernels cover are capacities: an exact first product's sizes times `slack`
 * (>= 1; 1.25 leaves room for replays whose structure grows).
 * tsg_batch_graph_launch queues one hipGraphLaunch on the context's stream and
 * returns (no host sync).  A launch reads the operands' arrays again: their
 * values and structure may change between launches; each problem's
 * dimensions, tile count and nonzero count, and the arrays' addresses, may
 * not.  tsg_batch_graph_result waits for the last launch and fills C[0..count)
 * with views of the graph's outputs, valid until the next launch or destroy
 * (never tsg_tile_free them).  A product past a capacity is flagged on the
 * device; result then computes the batch with tsg_spgemm_batched's exact
 * pipeline (stats[i].fallback_used = 1) and returns those outputs, owned by the
 * graph as well.  Results are bit-identical to tsg_spgemm_batched.  Runs the
 * binned kernels only: opts->numeric_variant 0 and B's column-major copies. */
typedef struct tsg_batch_graph tsg_batch_graph;
int tsg_batch_graph_create(tsg_context *ctx, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                           const tsg_options *opts, double slack, tsg_batch_graph **graph);
int tsg_batch_graph_launch(tsg_batch_graph *graph);
int tsg_batch_graph_result(tsg_batch_graph *graph, tsg_tile_matrix *C, tsg_stats *stats);
/* C's tile and nonzero capacities over the whole batch. */
int tsg_batch_graph_capacity(const tsg_batch_graph *graph, int64_t *cap_tiles, int64_t *cap_nnz);
void tsg_batch_graph_destroy(tsg_batch_graph *graph);

#ifdef __cplusplus
}
#endif
#endif /* TSG_H */
