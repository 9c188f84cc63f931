/*
 * oracle.h -- CPU restatement of the reference's tiled SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (gpu_spgemm_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates
 * (paths relative to the reference repository root).
 *
 * Layout conventions (identical to the reference's TileMatrix,
 * src/tile_format.h:15-42, with prefix arrays widened to int64):
 *   - 16x16 tiles; tile_ptr[tilem+1]; tile_colidx[numtile] ascending per tile-row
 *   - tile_nnz[nt+1] exclusive prefix of per-tile nonzeros
 *   - tile_ptrRow[nt*16] uint8 exclusive per-row offsets inside a tile
 *   - mask[nt*16] uint16, bit (15-c) set when local column c is present
 *   - tile_col[nnz]: (row<<4)|col for A and C, local col only for B (CSC order)
 */
#ifndef TSG_ORACLE_H
#define TSG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_tile {
    int32_t m, n, tilem, tilen;
    int64_t nnz, numtile;
    int64_t *tile_ptr;        /* [tilem+1]                                */
    int32_t *tile_colidx;     /* [numtile]  CSR-tile order                */
    int32_t *tile_rowidx;     /* [numtile]  (C only)                      */
    int64_t *tile_nnz;        /* [nt+1]     nt = numtile (A,C) / numtile_csc (B) */
    uint8_t *tile_ptrRow;     /* [nt*16]                                  */
    uint16_t *mask;           /* [nt*16]                                  */
    uint8_t *tile_col;        /* [nnz]                                    */
    double *val64;            /* [nnz] when value type is f64             */
    float *val32;             /* [nnz] when value type is f32             */
    int64_t *csc_tile_ptr;    /* [tilen+1]  (B only)                      */
    int32_t *csc_tile_rowidx; /* [numtile_csc] (B only)                   */
    int64_t numtile_csc;
    int32_t is_csc;           /* data arrays in CSC-tile order (B)        */
    int32_t is_f32;
} ora_tile;

typedef struct ora_check {
    int64_t total_ref;    /* structural nnz of the Gustavson product        */
    int64_t total_gpu;    /* entries found in the tiled result (inside MxN) */
    int64_t missing;      /* in reference, absent from tiled result         */
    int64_t extra;        /* in tiled result, absent from reference         */
    double rel_frob;      /* ||C_ref - C||_F / ||C_ref||_F                  */
    double max_scaled;    /* max |c - c_ref| / sum_k |a_ik b_kj|            */
    double max_abs;       /* max |c - c_ref|                                */
    int32_t pass;         /* reference PASS rule: 0 missing, 0 extra, rel_frob < 1e-5 */
} ora_check;

/* --- CSR helpers (values are always carried as double on input) --- */
/* flops = sum_{(i,k) in A} nnz(B[k,:])           spgemm_raw_v2.cu:95-99 */
uint64_t ora_flops(int32_t m, const int64_t *rpA, const int32_t *colA, const int64_t *rpB);
/* nnz(C) of the Gustavson product (structural union) and numtileC of Step 1
 * (tile-pattern product, empty tiles included) of A (m x k) * B (k x n), from
 * the CSRs, over row blocks in nthreads threads. */
int ora_count_product(int32_t m, int32_t k, int32_t n, const int64_t *rpA, const int32_t *colA,
                      const int64_t *rpB, const int32_t *colB, int32_t nthreads, int64_t *nnz_out,
                      int64_t *tiles_out);
/* B = A^T by counting sort                        spgemm_raw_v2.cu:78-89 */
void ora_transpose(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val,
                   int64_t *rpT, int32_t *colT, double *valT);

/* --- tile format --- */
/* CSR -> row tiles                                tile_format.h:93-222 */
int ora_csr2tile(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val,
                 int32_t is_f32, ora_tile *out);
/* Re-lay B's data arrays in CSC-tile order        tile_format.h:424-546 */
int ora_build_csc_tiles(ora_tile *B, int32_t m, int32_t n, const int64_t *rp, const int32_t *col,
                        const double *val);
void ora_tile_free(ora_tile *t);

/* --- tiled SpGEMM, steps 1-3 --- */
/* Step 1: C tile structure                        spgemm_raw_kernels.cuh:23-102 */
int ora_step1(const ora_tile *A, const ora_tile *B, ora_tile *C);
/* Step 2: symbolic masks, tile_nnzC, ptrRowC      spgemm_raw_kernels.cuh:105-215 */
int ora_step2(const ora_tile *A, const ora_tile *B, ora_tile *C);
/* Step 3: numeric values                          spgemm_raw_kernels_v2.cuh:10-139 */
int ora_step3(const ora_tile *A, const ora_tile *B, ora_tile *C);
/* Steps 1-3 back to back. */
int ora_spgemm_tiled(const ora_tile *A, const ora_tile *B, ora_tile *C);

/* Gustavson FP64 recompute and compare           spgemm_raw_v2.cu:253-359
 * Rows [row_lo,row_hi) only (pass 0,m for all).  C may be f64 or f32. */
int ora_gustavson_check(int32_t m, int32_t n, const int64_t *rpA, const int32_t *colA,
                        const double *valA, const int64_t *rpB, const int32_t *colB,
                        const double *valB, const ora_tile *C, int32_t row_lo, int32_t row_hi,
                        ora_check *out);

/* Gustavson product into CSR (structural union kept, sorted columns).
 * Two-phase: call with col/val NULL to get nnz into *nnz_out and rp filled. */
int ora_gustavson_csr(int32_t m, int32_t n, const int64_t *rpA, const int32_t *colA,
                      const double *valA, const int64_t *rpB, const int32_t *colB,
                      const double *valB, int64_t *rpC, int32_t *colC, double *valC);

/* CPU baseline: the reference's CPU path (csr2tile + build_csc_tiles + Gustavson
 * product) for rows [row_lo,row_hi) of C = A*B with nthreads workers.
 * Returns elapsed seconds of the Gustavson part in *gus_s and of the two
 * conversions in *conv_s.  Products processed are returned. */
int64_t ora_cpu_baseline(int32_t m, int32_t k, int32_t n, const int64_t *rpA, const int32_t *colA,
                         const double *valA, const int64_t *rpB, const int32_t *colB,
                         const double *valB, int32_t row_lo, int32_t row_hi, int32_t nthreads,
                         int32_t do_convert, double *conv_s, double *gus_s);

/* MTX reader restatement                          tile_format.h:51-90
 * Two-phase: first call with out arrays NULL returns m,n,nnz. */
int ora_load_mtx(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int64_t *rp,
                 int32_t *col, double *val);

#ifdef __cplusplus
}
#endif
#endif
