/*
 * oracle.c -- CPU restatement of the reference's tiled SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Citations are reference-repository paths (src/...).  The numeric step uses
 * one fixed summation order -- tile pairs in ascending k, A entries in storage
 * order, one fused multiply-add per product -- which is one of the orders the
 * reference's kernel may produce (its pair cache order depends on LDS atomics,
 * src/spgemm_raw_kernels.cuh:150,156), chosen so a GPU path that keeps the same
 * order can be compared bit for bit.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define TS 16

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void ora_tile_free(ora_tile *t) {
    if (!t) return;
    free(t->tile_ptr); free(t->tile_colidx); free(t->tile_rowidx); free(t->tile_nnz);
    free(t->tile_ptrRow); free(t->mask); free(t->tile_col); free(t->val64); free(t->val32);
    free(t->csc_tile_ptr); free(t->csc_tile_rowidx);
    memset(t, 0, sizeof *t);
}

/* spgemm_raw_v2.cu:95-99 */
uint64_t ora_flops(int32_t m, const int64_t *rpA, const int32_t *colA, const int64_t *rpB) {
    uint64_t f = 0;
    for (int32_t i = 0; i < m; i++)
        for (int64_t j = rpA[i]; j < rpA[i + 1]; j++) f += (uint64_t)(rpB[colA[j] + 1] - rpB[colA[j]]);
    return f;
}

/* spgemm_raw_v2.cu:78-89: counting-sort transpose; rows of A^T ascending per column. */
void ora_transpose(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val,
                   int64_t *rpT, int32_t *colT, double *valT) {
    memset(rpT, 0, sizeof(int64_t) * (size_t)(n + 1));
    int64_t nnz = rp[m];
    for (int64_t i = 0; i < nnz; i++) rpT[col[i] + 1]++;
    for (int32_t i = 0; i < n; i++) rpT[i + 1] += rpT[i];
    int64_t *off = malloc(sizeof(int64_t) * (size_t)(n + 1));
    memcpy(off, rpT, sizeof(int64_t) * (size_t)(n + 1));
    for (int32_t r = 0; r < m; r++)
        for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
            int64_t p = off[col[j]]++;
            colT[p] = r;
            valT[p] = val[j];
        }
    free(off);
}

/* tile_format.h:93-222 (csr2tile).  Three passes per tile-row: distinct tile
 * columns (:99-115), sorted columns + per-tile and per-row counts (:117-159),
 * then row-major scatter of values with packed (row<<4|col) bytes and
 * mask bit 15-col (:161-215). */
int ora_csr2tile(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val,
                 int32_t is_f32, ora_tile *T) {
    memset(T, 0, sizeof *T);
    T->m = m; T->n = n; T->nnz = rp[m]; T->is_f32 = is_f32;
    T->tilem = (m + TS - 1) / TS;
    T->tilen = (n + TS - 1) / TS;
    int32_t tilem = T->tilem, tilen = T->tilen;
    T->tile_ptr = calloc((size_t)tilem + 1, sizeof(int64_t));
    char *flag = calloc((size_t)tilen + 1, 1);
    int32_t *touched = malloc(sizeof(int32_t) * ((size_t)tilen + 1));
    for (int32_t bi = 0; bi < tilem; bi++) {
        int32_t nt = 0, rs = bi * TS, re = rs + TS < m ? rs + TS : m;
        for (int64_t j = rp[rs]; j < rp[re]; j++) {
            int32_t tc = col[j] / TS;
            if (!flag[tc]) { flag[tc] = 1; touched[nt++] = tc; }
        }
        T->tile_ptr[bi + 1] = nt;
        for (int32_t k = 0; k < nt; k++) flag[touched[k]] = 0;
    }
    for (int32_t i = 0; i < tilem; i++) T->tile_ptr[i + 1] += T->tile_ptr[i];
    int64_t numtile = T->numtile = T->tile_ptr[tilem];

    T->tile_colidx = calloc((size_t)numtile + 1, sizeof(int32_t));
    T->tile_nnz = calloc((size_t)numtile + 1, sizeof(int64_t));
    T->tile_ptrRow = calloc((size_t)numtile * TS + 1, 1);
    int32_t *nnz_per_tc = calloc((size_t)tilen + 1, sizeof(int32_t));
    uint8_t *ptr_flat = calloc((size_t)tilen * TS + 1, 1);
    for (int32_t bi = 0; bi < tilem; bi++) {
        int32_t nt = 0, rs = bi * TS, re = rs + TS < m ? rs + TS : m;
        for (int32_t ri = 0; ri < re - rs; ri++)
            for (int64_t j = rp[rs + ri]; j < rp[rs + ri + 1]; j++) {
                int32_t tc = col[j] / TS;
                if (!flag[tc]) { flag[tc] = 1; touched[nt++] = tc; }
                nnz_per_tc[tc]++;
                ptr_flat[(size_t)tc * TS + ri]++;
            }
        qsort(touched, (size_t)nt, sizeof(int32_t), cmp_i32);
        int64_t base = T->tile_ptr[bi];
        for (int32_t i = 0; i < nt; i++) {
            int32_t tc = touched[i];
            T->tile_colidx[base + i] = tc;
            T->tile_nnz[base + i + 1] = nnz_per_tc[tc];
            memcpy(T->tile_ptrRow + (size_t)(base + i) * TS, ptr_flat + (size_t)tc * TS, TS);
            flag[tc] = 0; nnz_per_tc[tc] = 0;
            memset(ptr_flat + (size_t)tc * TS, 0, TS);
        }
    }
    for (int64_t i = 0; i < numtile; i++) T->tile_nnz[i + 1] += T->tile_nnz[i];
    for (int64_t t = 0; t < numtile; t++) {     /* :155-159 exclusive scan per tile */
        uint8_t *p = T->tile_ptrRow + t * TS, s = 0;
        for (int r = 0; r < TS; r++) { uint8_t v = p[r]; p[r] = s; s = (uint8_t)(s + v); }
    }

    T->tile_col = calloc((size_t)T->nnz + 1, 1);
    if (is_f32) T->val32 = calloc((size_t)T->nnz + 1, sizeof(float));
    else T->val64 = calloc((size_t)T->nnz + 1, sizeof(double));
    T->mask = calloc((size_t)numtile * TS + 1, sizeof(uint16_t));
    int64_t *tc_pos = malloc(sizeof(int64_t) * ((size_t)tilen + 1));   /* :167 direct map */
    for (int32_t bi = 0; bi < tilem; bi++) {
        int32_t rs = bi * TS, re = rs + TS < m ? rs + TS : m;
        int64_t tb = T->tile_ptr[bi], te = T->tile_ptr[bi + 1];
        for (int64_t t = tb; t < te; t++) tc_pos[T->tile_colidx[t]] = t;
        /* the reference gathers each tile's entries in CSR traversal order
         * (row-major, :185-192) and copies them row by row (:195-211); writing
         * them at a running per-tile cursor gives the same positions. */
        int64_t *cur = malloc(sizeof(int64_t) * (size_t)(te - tb + 1));
        for (int64_t t = tb; t < te; t++) cur[t - tb] = T->tile_nnz[t];
        for (int32_t ri = 0; ri < re - rs; ri++)
            for (int64_t j = rp[rs + ri]; j < rp[rs + ri + 1]; j++) {
                int32_t tc = col[j] / TS, lc = col[j] - tc * TS;
                int64_t t = tc_pos[tc];
                int64_t o = cur[t - tb]++;
                if (is_f32) T->val32[o] = (float)val[j]; else T->val64[o] = val[j];
                T->tile_col[o] = (uint8_t)((ri << 4) | lc);
                T->mask[t * TS + ri] |= (uint16_t)(1u << (TS - 1 - lc));
            }
        free(cur);
    }
    free(tc_pos); free(nnz_per_tc); free(ptr_flat); free(flag); free(touched);
    return 0;
}

/* tile_format.h:424-546 (build_csc_tiles).  B's tiles re-enumerated by tile
 * column then ascending tile row (:439-475); inside each tile row-major with
 * ascending columns and the LOCAL column only in tile_col (:535-542).
 * tile_ptr/tile_colidx keep the CSR-tile order from csr2tile. */
int ora_build_csc_tiles(ora_tile *B, int32_t m, int32_t n, const int64_t *rp, const int32_t *col,
                        const double *val) {
    int64_t nnz = rp[m];
    int32_t tilem = B->tilem, tilen = B->tilen;
    int64_t *cp = calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *ri_ = malloc(sizeof(int32_t) * ((size_t)nnz + 1));
    double *cv = malloc(sizeof(double) * ((size_t)nnz + 1));
    ora_transpose(m, n, rp, col, val, cp, ri_, cv);        /* :428-437 */

    B->csc_tile_ptr = calloc((size_t)tilen + 1, sizeof(int64_t));
    char *flag = calloc((size_t)tilem + 1, 1);
    int32_t *touched = malloc(sizeof(int32_t) * ((size_t)tilem + 1));
    for (int32_t bt = 0; bt < tilen; bt++) {                 /* :439-453 */
        int32_t nt = 0, cs = bt * TS, ce = cs + TS < n ? cs + TS : n;
        for (int32_t c = cs; c < ce; c++)
            for (int64_t j = cp[c]; j < cp[c + 1]; j++) {
                int32_t tr = ri_[j] / TS;
                if (!flag[tr]) { flag[tr] = 1; touched[nt++] = tr; }
            }
        B->csc_tile_ptr[bt + 1] = nt;
        for (int32_t k = 0; k < nt; k++) flag[touched[k]] = 0;
    }
    for (int32_t i = 0; i < tilen; i++) B->csc_tile_ptr[i + 1] += B->csc_tile_ptr[i];
    int64_t nt_csc = B->numtile_csc = B->csc_tile_ptr[tilen];
    B->csc_tile_rowidx = calloc((size_t)nt_csc + 1, sizeof(int32_t));
    for (int32_t bt = 0; bt < tilen; bt++) {                 /* :457-475 */
        int32_t nt = 0, cs = bt * TS, ce = cs + TS < n ? cs + TS : n;
        int64_t pos = B->csc_tile_ptr[bt];
        for (int32_t c = cs; c < ce; c++)
            for (int64_t j = cp[c]; j < cp[c + 1]; j++) {
                int32_t tr = ri_[j] / TS;
                if (!flag[tr]) { flag[tr] = 1; touched[nt++] = tr; B->csc_tile_rowidx[pos++] = tr; }
            }
        qsort(B->csc_tile_rowidx + B->csc_tile_ptr[bt], (size_t)nt, sizeof(int32_t), cmp_i32);
        for (int32_t k = 0; k < nt; k++) flag[touched[k]] = 0;
    }
    /* replace the data arrays with CSC-tile-ordered ones (:480-546) */
    free(B->tile_nnz); free(B->tile_ptrRow); free(B->tile_col); free(B->mask);
    free(B->val64); free(B->val32);
    B->val64 = NULL; B->val32 = NULL;
    B->tile_nnz = calloc((size_t)nt_csc + 1, sizeof(int64_t));
    int64_t *tr_pos = malloc(sizeof(int64_t) * ((size_t)tilem + 1));
    int32_t *rowcnt = calloc((size_t)nt_csc * TS + 1, sizeof(int32_t));
    for (int32_t bt = 0; bt < tilen; bt++) {                 /* :482-517 counts */
        for (int64_t t = B->csc_tile_ptr[bt]; t < B->csc_tile_ptr[bt + 1]; t++)
            tr_pos[B->csc_tile_rowidx[t]] = t;
        int32_t cs = bt * TS, ce = cs + TS < n ? cs + TS : n;
        for (int32_t c = cs; c < ce; c++)
            for (int64_t j = cp[c]; j < cp[c + 1]; j++) {
                int64_t t = tr_pos[ri_[j] / TS];
                B->tile_nnz[t + 1]++;
                rowcnt[t * TS + ri_[j] % TS]++;
            }
    }
    for (int64_t i = 0; i < nt_csc; i++) B->tile_nnz[i + 1] += B->tile_nnz[i];
    B->tile_ptrRow = calloc((size_t)nt_csc * TS + 1, 1);
    for (int64_t t = 0; t < nt_csc; t++) {                   /* :519-525 */
        uint8_t s = 0;
        for (int r = 0; r < TS; r++) { B->tile_ptrRow[t * TS + r] = s; s = (uint8_t)(s + rowcnt[t * TS + r]); }
    }
    int64_t tot = B->tile_nnz[nt_csc];
    B->tile_col = calloc((size_t)tot + 1, 1);
    if (B->is_f32) B->val32 = calloc((size_t)tot + 1, sizeof(float));
    else B->val64 = calloc((size_t)tot + 1, sizeof(double));
    B->mask = calloc((size_t)nt_csc * TS + 1, sizeof(uint16_t));
    memset(rowcnt, 0, sizeof(int32_t) * (size_t)nt_csc * TS);   /* reused as tileRowOff */
    for (int32_t bt = 0; bt < tilen; bt++) {                 /* :527-546 fill */
        for (int64_t t = B->csc_tile_ptr[bt]; t < B->csc_tile_ptr[bt + 1]; t++)
            tr_pos[B->csc_tile_rowidx[t]] = t;
        int32_t cs = bt * TS, ce = cs + TS < n ? cs + TS : n;
        for (int32_t c = cs; c < ce; c++)
            for (int64_t j = cp[c]; j < cp[c + 1]; j++) {
                int32_t brow = ri_[j], lr = brow % TS, lc = c % TS;
                int64_t t = tr_pos[brow / TS];
                int64_t off = B->tile_nnz[t] + B->tile_ptrRow[t * TS + lr] + rowcnt[t * TS + lr]++;
                if (B->is_f32) B->val32[off] = (float)cv[j]; else B->val64[off] = cv[j];
                B->tile_col[off] = (uint8_t)lc;
                B->mask[t * TS + lr] |= (uint16_t)(1u << (TS - 1 - lc));
            }
    }
    B->is_csc = 1;
    free(rowcnt); free(tr_pos); free(flag); free(touched); free(cp); free(ri_); free(cv);
    return 0;
}

/* spgemm_raw_kernels.cuh:23-102 (step1_count_kernel + step1_fill_kernel):
 * C tile-row bi holds the union over k in A's tile-row bi of B's CSR-tile
 * columns of tile-row k (the driver passes B's CSR tile_ptr/tile_colidx,
 * spgemm_raw_v2.cu:127-128), ascending (bitmap walk, :91-100). */
int ora_step1(const ora_tile *A, const ora_tile *B, ora_tile *C) {
    memset(C, 0, sizeof *C);
    C->m = A->m; C->n = B->n; C->tilem = A->tilem; C->tilen = B->tilen;
    C->is_f32 = A->is_f32;
    int32_t tilem = A->tilem, tilen = B->tilen;
    C->tile_ptr = calloc((size_t)tilem + 1, sizeof(int64_t));
    char *flag = calloc((size_t)tilen + 1, 1);
    int32_t *touched = malloc(sizeof(int32_t) * ((size_t)tilen + 1));
    for (int pass = 0; pass < 2; pass++) {
        for (int32_t bi = 0; bi < tilem; bi++) {
            int32_t nt = 0;
            for (int64_t i = A->tile_ptr[bi]; i < A->tile_ptr[bi + 1]; i++) {
                int32_t k = A->tile_colidx[i];
                for (int64_t j = B->tile_ptr[k]; j < B->tile_ptr[k + 1]; j++) {
                    int32_t c = B->tile_colidx[j];
                    if (!flag[c]) { flag[c] = 1; touched[nt++] = c; }
                }
            }
            for (int32_t q = 0; q < nt; q++) flag[touched[q]] = 0;
            if (pass == 0) { C->tile_ptr[bi + 1] = nt; continue; }
            qsort(touched, (size_t)nt, sizeof(int32_t), cmp_i32);
            int64_t base = C->tile_ptr[bi];
            for (int32_t q = 0; q < nt; q++) {
                C->tile_rowidx[base + q] = bi;
                C->tile_colidx[base + q] = touched[q];
            }
        }
        if (pass == 0) {      /* spgemm_raw_v2.cu:131-141 exclusive scan, numtileC */
            for (int32_t i = 0; i < tilem; i++) C->tile_ptr[i + 1] += C->tile_ptr[i];
            C->numtile = C->tile_ptr[tilem];
            C->tile_rowidx = calloc((size_t)C->numtile + 1, sizeof(int32_t));
            C->tile_colidx = calloc((size_t)C->numtile + 1, sizeof(int32_t));
        }
    }
    free(flag); free(touched);
    return 0;
}

/* Matched tile pairs of C tile (bi,bj): k in A's tile-row bi and in B's CSC
 * tile-column bj, ascending k.  Returns the count; positions into pa/pb. */
static int64_t tile_pairs(const ora_tile *A, const ora_tile *B, int32_t bi, int32_t bj,
                          int64_t *pa, int64_t *pb) {
    int64_t ai = A->tile_ptr[bi], ae = A->tile_ptr[bi + 1];
    int64_t bk = B->csc_tile_ptr[bj], be = B->csc_tile_ptr[bj + 1];
    int64_t cnt = 0;
    while (ai < ae && bk < be) {       /* merge path, spgemm_raw_kernels.cuh:188-195 */
        int32_t ka = A->tile_colidx[ai], kb = B->csc_tile_rowidx[bk];
        if (ka == kb) { pa[cnt] = ai; pb[cnt] = bk; cnt++; ai++; bk++; }
        else if (ka < kb) ai++;
        else bk++;
    }
    return cnt;
}

static inline int64_t row_end(const ora_tile *T, int64_t t, int r) {
    return r == TS - 1 ? T->tile_nnz[t + 1] - T->tile_nnz[t] : (int64_t)T->tile_ptrRow[t * TS + r + 1];
}

/* spgemm_raw_kernels.cuh:105-202 (symbolic) and :205-215 (ptrRowC):
 * maskC[r] = OR over pairs, OR over A entries (r,kk) of maskB[kk]; tile_nnzC
 * = sum of popcounts, then exclusive scan (spgemm_raw_v2.cu:202-211);
 * ptrRowC = per-tile exclusive scan of row popcounts (uint8). */
int ora_step2(const ora_tile *A, const ora_tile *B, ora_tile *C) {
    if (!B->is_csc) return -1;
    int64_t ntc = C->numtile;
    C->mask = calloc((size_t)ntc * TS + 1, sizeof(uint16_t));
    C->tile_ptrRow = calloc((size_t)ntc * TS + 1, 1);
    C->tile_nnz = calloc((size_t)ntc + 1, sizeof(int64_t));
    int64_t cap = 1024;
    int64_t *pa = malloc(sizeof(int64_t) * cap), *pb = malloc(sizeof(int64_t) * cap);
    for (int64_t ti = 0; ti < ntc; ti++) {
        int32_t bi = C->tile_rowidx[ti], bj = C->tile_colidx[ti];
        int64_t need = (A->tile_ptr[bi + 1] - A->tile_ptr[bi]);
        if (need > cap) { cap = need; pa = realloc(pa, sizeof(int64_t) * cap); pb = realloc(pb, sizeof(int64_t) * cap); }
        int64_t np = tile_pairs(A, B, bi, bj, pa, pb);
        uint16_t rm[TS] = {0};
        for (int64_t p = 0; p < np; p++) {
            int64_t a = pa[p], b = pb[p];
            const uint8_t *cA = A->tile_col + A->tile_nnz[a];
            for (int r = 0; r < TS; r++)
                for (int64_t e = A->tile_ptrRow[a * TS + r]; e < row_end(A, a, r); e++)
                    rm[r] |= B->mask[b * TS + (cA[e] & 0xf)];
        }
        int cnt = 0;
        uint8_t s = 0;
        for (int r = 0; r < TS; r++) {
            C->mask[ti * TS + r] = rm[r];
            C->tile_ptrRow[ti * TS + r] = s;
            int pc = __builtin_popcount(rm[r]);
            s = (uint8_t)(s + pc);
            cnt += pc;
        }
        C->tile_nnz[ti + 1] = cnt;
    }
    for (int64_t i = 0; i < ntc; i++) C->tile_nnz[i + 1] += C->tile_nnz[i];
    C->nnz = C->tile_nnz[ntc];
    free(pa); free(pb);
    return 0;
}

/* spgemm_raw_kernels_v2.cuh:10-139 (numeric): for every C tile, every matched
 * pair (ascending k), every A entry (r,kk,va) and every B entry (kk,c,vb):
 * acc[r][c] += va*vb (one fused multiply-add); then the values of the set bits
 * of maskC[r], ascending c, are written at tile_nnzC + ptrRowC[r] with col byte
 * (r<<4)|c (:120-137).  Empty tiles write nothing (:33). */
int ora_step3(const ora_tile *A, const ora_tile *B, ora_tile *C) {
    int64_t ntc = C->numtile, nnzc = C->nnz;
    int f32 = A->is_f32;
    C->tile_col = calloc((size_t)nnzc + 1, 1);
    if (f32) C->val32 = calloc((size_t)nnzc + 1, sizeof(float));
    else C->val64 = calloc((size_t)nnzc + 1, sizeof(double));
    int64_t cap = 1024;
    int64_t *pa = malloc(sizeof(int64_t) * cap), *pb = malloc(sizeof(int64_t) * cap);
    for (int64_t ti = 0; ti < ntc; ti++) {
        int64_t base = C->tile_nnz[ti];
        if (C->tile_nnz[ti + 1] == base) continue;
        int32_t bi = C->tile_rowidx[ti], bj = C->tile_colidx[ti];
        int64_t need = (A->tile_ptr[bi + 1] - A->tile_ptr[bi]);
        if (need > cap) { cap = need; pa = realloc(pa, sizeof(int64_t) * cap); pb = realloc(pb, sizeof(int64_t) * cap); }
        int64_t np = tile_pairs(A, B, bi, bj, pa, pb);
        double acc[TS][TS];
        float accf[TS][TS];
        memset(acc, 0, sizeof acc);
        memset(accf, 0, sizeof accf);
        for (int64_t p = 0; p < np; p++) {
            int64_t a = pa[p], b = pb[p];
            int64_t as = A->tile_nnz[a], bs = B->tile_nnz[b];
            for (int r = 0; r < TS; r++)
                for (int64_t e = A->tile_ptrRow[a * TS + r]; e < row_end(A, a, r); e++) {
                    int kk = A->tile_col[as + e] & 0xf;
                    for (int64_t q = B->tile_ptrRow[b * TS + kk]; q < row_end(B, b, kk); q++) {
                        int c = B->tile_col[bs + q] & 0xf;
                        if (f32) accf[r][c] = fmaf(A->val32[as + e], B->val32[bs + q], accf[r][c]);
                        else acc[r][c] = fma(A->val64[as + e], B->val64[bs + q], acc[r][c]);
                    }
                }
        }
        for (int r = 0; r < TS; r++) {
            uint16_t mk = C->mask[ti * TS + r];
            int64_t o = base + C->tile_ptrRow[ti * TS + r];
            for (int c = 0; c < TS; c++)
                if ((mk >> (TS - 1 - c)) & 1) {
                    if (f32) C->val32[o] = accf[r][c]; else C->val64[o] = acc[r][c];
                    C->tile_col[o] = (uint8_t)((r << 4) | c);
                    o++;
                }
        }
    }
    free(pa); free(pb);
    return 0;
}

int ora_spgemm_tiled(const ora_tile *A, const ora_tile *B, ora_tile *C) {
    int rc = ora_step1(A, B, C);
    if (!rc) rc = ora_step2(A, B, C);
    if (!rc) rc = ora_step3(A, B, C);
    return rc;
}

typedef struct { int32_t c; int64_t src; } colsrc;
static int cmp_colsrc(const void *a, const void *b) {
    const colsrc *x = a, *y = b;
    if (x->c != y->c) return (x->c > y->c) - (x->c < y->c);
    return (x->src > y->src) - (x->src < y->src);
}

/* spgemm_raw_v2.cu:253-359: C tiles -> per-row (col,val) lists restricted to
 * MxN (:268-284), FP64 Gustavson per row with a dense accumulator and the
 * structural union kept (:297-308), merge compare counting missing/extra and
 * the Frobenius sums (:311-336), PASS rule (:355).  max_scaled is added: the
 * error relative to sum_k |a_ik b_kj|, robust to cancellation. */
int ora_gustavson_check(int32_t m, int32_t n, const int64_t *rpA, const int32_t *colA,
                        const double *valA, const int64_t *rpB, const int32_t *colB,
                        const double *valB, const ora_tile *C, int32_t row_lo, int32_t row_hi,
                        ora_check *out) {
    memset(out, 0, sizeof *out);
    if (row_lo < 0) row_lo = 0;
    if (row_hi > m) row_hi = m;
    int32_t nr = row_hi > row_lo ? row_hi - row_lo : 0;
    int64_t *rc = calloc((size_t)nr + 1, sizeof(int64_t));
    /* gather the tiled result's entries of rows [row_lo,row_hi) */
    int32_t tlo = row_lo / TS, thi = (row_hi + TS - 1) / TS;
    for (int pass = 0; pass < 2; pass++) {
        int64_t *cur = NULL;
        colsrc *ent = NULL;
        if (pass == 1) {
            for (int32_t i = 0; i < nr; i++) rc[i + 1] += rc[i];
            cur = malloc(sizeof(int64_t) * ((size_t)nr + 1));
            memcpy(cur, rc, sizeof(int64_t) * ((size_t)nr + 1));
            ent = malloc(sizeof(colsrc) * ((size_t)rc[nr] + 1));
        }
        for (int32_t tr = tlo; tr < thi && tr < C->tilem; tr++)
            for (int64_t ti = C->tile_ptr[tr]; ti < C->tile_ptr[tr + 1]; ti++) {
                int32_t tc = C->tile_colidx[ti];
                for (int64_t e = C->tile_nnz[ti]; e < C->tile_nnz[ti + 1]; e++) {
                    int32_t gr = tr * TS + ((C->tile_col[e] >> 4) & 0xf);
                    int32_t gc = tc * TS + (C->tile_col[e] & 0xf);
                    if (gr < row_lo || gr >= row_hi || gc >= n) continue;
                    if (pass == 0) rc[gr - row_lo + 1]++;
                    else { colsrc cs = {gc, e}; ent[cur[gr - row_lo]++] = cs; }
                }
            }
        if (pass == 1) {
            free(cur);
            double *acc = calloc((size_t)n + 1, sizeof(double));
            double *aabs = calloc((size_t)n + 1, sizeof(double));
            char *used = calloc((size_t)n + 1, 1);
            int32_t *cols = malloc(sizeof(int32_t) * ((size_t)n + 1));
            double fd2 = 0, fr2 = 0;
            for (int32_t i = row_lo; i < row_hi; i++) {
                colsrc *g = ent + rc[i - row_lo];
                int64_t gn = rc[i - row_lo + 1] - rc[i - row_lo];
                qsort(g, (size_t)gn, sizeof(colsrc), cmp_colsrc);
                out->total_gpu += gn;
                int32_t nc = 0;
                for (int64_t ja = rpA[i]; ja < rpA[i + 1]; ja++) {
                    int32_t k = colA[ja];
                    double va = valA[ja];
                    for (int64_t jb = rpB[k]; jb < rpB[k + 1]; jb++) {
                        int32_t c = colB[jb];
                        if (!used[c]) { used[c] = 1; cols[nc++] = c; }
                        acc[c] += va * valB[jb];
                        aabs[c] += fabs(va * valB[jb]);
                    }
                }
                qsort(cols, (size_t)nc, sizeof(int32_t), cmp_i32);
                out->total_ref += nc;
                int64_t gi = 0;
                for (int32_t ci = 0; ci < nc; ci++) {
                    int32_t c = cols[ci];
                    double rv = acc[c];
                    fr2 += rv * rv;
                    while (gi < gn && g[gi].c < c) {
                        double gv = C->is_f32 ? C->val32[g[gi].src] : C->val64[g[gi].src];
                        out->extra++; fd2 += gv * gv; gi++;
                    }
                    if (gi < gn && g[gi].c == c) {
                        double gv = C->is_f32 ? C->val32[g[gi].src] : C->val64[g[gi].src];
                        double d = rv - gv;
                        fd2 += d * d;
                        double ad = fabs(d);
                        if (ad > out->max_abs) out->max_abs = ad;
                        double sc = aabs[c] > 0 ? ad / aabs[c] : (ad > 0 ? INFINITY : 0.0);
                        if (sc > out->max_scaled) out->max_scaled = sc;
                        gi++;
                    } else {
                        out->missing++; fd2 += rv * rv;
                    }
                }
                while (gi < gn) {
                    double gv = C->is_f32 ? C->val32[g[gi].src] : C->val64[g[gi].src];
                    out->extra++; fd2 += gv * gv; gi++;
                }
                for (int32_t ci = 0; ci < nc; ci++) { acc[cols[ci]] = 0; aabs[cols[ci]] = 0; used[cols[ci]] = 0; }
            }
            out->rel_frob = fr2 > 0 ? sqrt(fd2 / fr2) : (fd2 > 0 ? INFINITY : 0.0);
            out->pass = out->missing == 0 && out->extra == 0 && out->rel_frob < 1e-5;
            free(acc); free(aabs); free(used); free(cols); free(ent);
        }
    }
    free(rc);
    return 0;
}

/* Gustavson CSR product with the structural union (spgemm_raw_v2.cu:297-308). */
int ora_gustavson_csr(int32_t m, int32_t n, const int64_t *rpA, const int32_t *colA,
                      const double *valA, const int64_t *rpB, const int32_t *colB,
                      const double *valB, int64_t *rpC, int32_t *colC, double *valC) {
    double *acc = calloc((size_t)n + 1, sizeof(double));
    char *used = calloc((size_t)n + 1, 1);
    int32_t *cols = malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!colC) rpC[0] = 0;
    for (int32_t i = 0; i < m; i++) {
        int32_t nc = 0;
        for (int64_t ja = rpA[i]; ja < rpA[i + 1]; ja++) {
            int32_t k = colA[ja];
            for (int64_t jb = rpB[k]; jb < rpB[k + 1]; jb++) {
                int32_t c = colB[jb];
                if (!used[c]) { used[c] = 1; cols[nc++] = c; }
                acc[c] += valA[ja] * valB[jb];
            }
        }
        if (!colC) {
            rpC[i + 1] = rpC[i] + nc;
        } else {
            qsort(cols, (size_t)nc, sizeof(int32_t), cmp_i32);
            for (int32_t q = 0; q < nc; q++) { colC[rpC[i] + q] = cols[q]; valC[rpC[i] + q] = acc[cols[q]]; }
        }
        for (int32_t q = 0; q < nc; q++) { acc[cols[q]] = 0; used[cols[q]] = 0; }
    }
    free(acc); free(used); free(cols);
    return 0;
}

/* ---- CPU baseline: Gustavson over a row range, row-block parallel ---- */
typedef struct {
    int32_t n, lo, hi;
    const int64_t *rpA, *rpB;
    const int32_t *colA, *colB;
    const double *valA, *valB;
    int64_t products;
    double checksum;
} gus_job;

static void *gus_worker(void *p) {
    gus_job *J = p;
    double *acc = calloc((size_t)J->n + 1, sizeof(double));
    char *used = calloc((size_t)J->n + 1, 1);
    int32_t *cols = malloc(sizeof(int32_t) * ((size_t)J->n + 1));
    size_t ocap = 1 << 16, olen = 0;
    int32_t *ocol = malloc(sizeof(int32_t) * ocap);
    double *oval = malloc(sizeof(double) * ocap);
    int64_t prod = 0;
    double cs = 0;
    for (int32_t i = J->lo; i < J->hi; i++) {
        int32_t nc = 0;
        for (int64_t ja = J->rpA[i]; ja < J->rpA[i + 1]; ja++) {
            int32_t k = J->colA[ja];
            double va = J->valA[ja];
            for (int64_t jb = J->rpB[k]; jb < J->rpB[k + 1]; jb++) {
                int32_t c = J->colB[jb];
                if (!used[c]) { used[c] = 1; cols[nc++] = c; }
                acc[c] += va * J->valB[jb];
            }
            prod += J->rpB[k + 1] - J->rpB[k];
        }
        qsort(cols, (size_t)nc, sizeof(int32_t), cmp_i32);
        if (olen + (size_t)nc > ocap) olen = 0;      /* output ring: written, not kept */
        if ((size_t)nc > ocap) { ocap = (size_t)nc; ocol = realloc(ocol, sizeof(int32_t) * ocap); oval = realloc(oval, sizeof(double) * ocap); }
        for (int32_t q = 0; q < nc; q++) {
            ocol[olen + q] = cols[q]; oval[olen + q] = acc[cols[q]];
            acc[cols[q]] = 0; used[cols[q]] = 0;
        }
        if (nc) cs += oval[olen];
        olen += (size_t)nc;
    }
    J->products = prod;
    J->checksum = cs;
    free(acc); free(used); free(cols); free(ocol); free(oval);
    return NULL;
}

int64_t ora_cpu_baseline(int32_t m, int32_t k, int32_t n, const int64_t *rpA, const int32_t *colA,
                         const double *valA, const int64_t *rpB, const int32_t *colB,
                         const double *valB, int32_t row_lo, int32_t row_hi, int32_t nthreads,
                         int32_t do_convert, double *conv_s, double *gus_s) {
    *conv_s = 0;
    if (do_convert) {  /* A is m x k, B is k x n */
        double t0 = now_s();
        ora_tile tA, tB;
        ora_csr2tile(m, k, rpA, colA, valA, 0, &tA);
        ora_csr2tile(k, n, rpB, colB, valB, 0, &tB);
        ora_build_csc_tiles(&tB, k, n, rpB, colB, valB);
        *conv_s = now_s() - t0;
        ora_tile_free(&tA); ora_tile_free(&tB);
    }
    if (nthreads < 1) nthreads = 1;
    gus_job *jobs = calloc((size_t)nthreads, sizeof(gus_job));
    pthread_t *th = calloc((size_t)nthreads, sizeof(pthread_t));
    /* split by products so threads get even work */
    int64_t totp = 0;
    for (int32_t i = row_lo; i < row_hi; i++)
        for (int64_t j = rpA[i]; j < rpA[i + 1]; j++) totp += rpB[colA[j] + 1] - rpB[colA[j]];
    int32_t r = row_lo;
    int64_t acc = 0;
    for (int t = 0; t < nthreads; t++) {
        gus_job *J = &jobs[t];
        J->n = n; J->rpA = rpA; J->rpB = rpB; J->colA = colA; J->colB = colB; J->valA = valA; J->valB = valB;
        J->lo = r;
        int64_t target = totp * (t + 1) / nthreads;
        while (r < row_hi && (acc < target || t == nthreads - 1)) {
            for (int64_t j = rpA[r]; j < rpA[r + 1]; j++) acc += rpB[colA[j] + 1] - rpB[colA[j]];
            r++;
        }
        J->hi = r;
    }
    double t0 = now_s();
    if (nthreads == 1) gus_worker(&jobs[0]);
    else {
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, gus_worker, &jobs[t]);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    *gus_s = now_s() - t0;
    int64_t prod = 0;
    for (int t = 0; t < nthreads; t++) prod += jobs[t].products;
    free(jobs); free(th);
    return prod;
}

/* ---- whole-product counts: nnz(C) of the Gustavson product with its
 * structural union (spgemm_raw_v2.cu:297-308), and numtileC of Step 1, the
 * boolean product of the tile patterns (spgemm_raw_kernels.cuh:23-57: for
 * tile-row I the distinct tile columns of B's tile-rows k reached by A's
 * tiles (I, k)), both straight from the CSRs, over row blocks in threads. */
typedef struct {
    int32_t n, k, tlo, thi, m;
    const int64_t *rpA, *rpB, *btp;   /* btp/btc: B's tile-row patterns */
    const int32_t *colA, *colB, *btc;
    int64_t nnz, tiles;
} cnt_job;

static void *cnt_worker(void *p) {
    cnt_job *J = p;
    const int32_t tn = (J->n + 15) / 16, tk = (J->k + 15) / 16;
    char *used = calloc((size_t)J->n + 1, 1);
    int32_t *cols = malloc(sizeof(int32_t) * ((size_t)J->n + 1));
    char *tused = calloc((size_t)tn + 1, 1), *kused = calloc((size_t)tk + 1, 1);
    int32_t *tc = malloc(sizeof(int32_t) * ((size_t)tn + 1)), *kc = malloc(sizeof(int32_t) * ((size_t)tk + 1));
    int64_t nnz = 0, tiles = 0;
    for (int32_t ti = J->tlo; ti < J->thi; ti++) {
        int32_t nk = 0, nt = 0;
        const int32_t r1 = (ti + 1) * 16 < J->m ? (ti + 1) * 16 : J->m;
        for (int32_t i = ti * 16; i < r1; i++) {
            int32_t nc = 0;
            for (int64_t ja = J->rpA[i]; ja < J->rpA[i + 1]; ja++) {
                const int32_t k = J->colA[ja];
                if (!kused[k >> 4]) { kused[k >> 4] = 1; kc[nk++] = k >> 4; }
                for (int64_t jb = J->rpB[k]; jb < J->rpB[k + 1]; jb++) {
                    const int32_t c = J->colB[jb];
                    if (!used[c]) { used[c] = 1; cols[nc++] = c; }
                }
            }
            nnz += nc;
            for (int32_t q = 0; q < nc; q++) used[cols[q]] = 0;
        }
        for (int32_t q = 0; q < nk; q++) {      /* union of the reached B tile-rows' patterns */
            const int32_t kt = kc[q];
            kused[kt] = 0;
            for (int64_t e = J->btp[kt]; e < J->btp[kt + 1]; e++) {
                const int32_t c = J->btc[e];
                if (!tused[c]) { tused[c] = 1; tc[nt++] = c; }
            }
        }
        tiles += nt;
        for (int32_t q = 0; q < nt; q++) tused[tc[q]] = 0;
    }
    J->nnz = nnz;
    J->tiles = tiles;
    free(used); free(cols); free(tused); free(kused); free(tc); free(kc);
    return NULL;
}

int ora_count_product(int32_t m, int32_t k, int32_t n, const int64_t *rpA, const int32_t *colA,
                      const int64_t *rpB, const int32_t *colB, int32_t nthreads, int64_t *nnz_out,
                      int64_t *tiles_out) {
    const int32_t tmA = (m + 15) / 16, tmB = (k + 15) / 16, tn = (n + 15) / 16;
    /* B's tile pattern per tile-row: distinct tile columns, any order */
    int64_t *btp = calloc((size_t)tmB + 1, sizeof(int64_t));
    char *f = calloc((size_t)tn + 1, 1);
    int32_t *tmp = malloc(sizeof(int32_t) * ((size_t)tn + 1));
    int32_t *btc = NULL;
    for (int pass = 0; pass < 2; pass++) {
        for (int32_t t = 0; t < tmB; t++) {
            int32_t c = 0;
            const int32_t r1 = (t + 1) * 16 < k ? (t + 1) * 16 : k;
            for (int32_t i = t * 16; i < r1; i++)
                for (int64_t j = rpB[i]; j < rpB[i + 1]; j++)
                    if (!f[colB[j] >> 4]) { f[colB[j] >> 4] = 1; tmp[c++] = colB[j] >> 4; }
            if (pass == 0) btp[t + 1] = c;
            else memcpy(btc + btp[t], tmp, sizeof(int32_t) * (size_t)c);
            for (int32_t q = 0; q < c; q++) f[tmp[q]] = 0;
        }
        if (pass == 0) {
            for (int32_t t = 0; t < tmB; t++) btp[t + 1] += btp[t];
            btc = malloc(sizeof(int32_t) * ((size_t)btp[tmB] + 1));
        }
    }
    free(f); free(tmp);
    if (nthreads < 1) nthreads = 1;
    cnt_job *jobs = calloc((size_t)nthreads, sizeof(cnt_job));
    pthread_t *th = calloc((size_t)nthreads, sizeof(pthread_t));
    /* interleaved blocks of 64 tile-rows would balance better; contiguous
     * ranges split by A's nonzeros are enough here */
    int32_t t = 0;
    for (int q = 0; q < nthreads; q++) {
        cnt_job *J = &jobs[q];
        J->n = n; J->k = k; J->m = m; J->rpA = rpA; J->colA = colA; J->rpB = rpB; J->colB = colB; J->btp = btp; J->btc = btc;
        J->tlo = t;
        const int64_t target = rpA[m] * (q + 1) / nthreads;
        while (t < tmA && (rpA[(t + 1) * 16 < m ? (t + 1) * 16 : m] <= target || q == nthreads - 1)) t++;
        J->thi = t;
    }
    for (int q = 0; q < nthreads; q++) pthread_create(&th[q], NULL, cnt_worker, &jobs[q]);
    int64_t nz = 0, tl = 0;
    for (int q = 0; q < nthreads; q++) {
        pthread_join(th[q], NULL);
        nz += jobs[q].nnz;
        tl += jobs[q].tiles;
    }
    *nnz_out = nz;
    *tiles_out = tl;
    free(jobs); free(th); free(btp); free(btc);
    return 0;
}

/* tile_format.h:51-90 (load_mtx): banner flags from the first line, '%'
 * comments skipped, 1-based -> 0-based, symmetric off-diagonals mirrored,
 * pattern -> 1.0, complex -> real part, entries sorted by (row,col) and
 * duplicates kept (ties keep file order here). */
typedef struct { int32_t r, c; int64_t k; double v; } mtx_ent;
static int cmp_mtx(const void *a, const void *b) {
    const mtx_ent *x = a, *y = b;
    if (x->r != y->r) return (x->r > y->r) - (x->r < y->r);
    if (x->c != y->c) return (x->c > y->c) - (x->c < y->c);
    return (x->k > y->k) - (x->k < y->k);
}

int ora_load_mtx(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int64_t *rp,
                 int32_t *col, double *val) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[1024];
    int sym = 0, pat = 0, cpx = 0;
    if (fgets(line, sizeof line, f)) {
        sym = strstr(line, "symmetric") != NULL;
        pat = strstr(line, "pattern") != NULL;
        cpx = strstr(line, "complex") != NULL;
    }
    while (fgets(line, sizeof line, f)) if (line[0] != '%') break;
    int mm = 0, nn = 0, nf = 0;
    if (sscanf(line, "%d %d %d", &mm, &nn, &nf) != 3) { fclose(f); return -2; }
    mtx_ent *e = malloc(sizeof(mtx_ent) * (2 * (size_t)nf + 1));
    int64_t ne = 0;
    for (int i = 0; i < nf; i++) {
        int r, c;
        double v = 1.0, vi = 0.0;
        int got;
        if (pat) got = fscanf(f, "%d %d", &r, &c) == 2;
        else if (cpx) got = fscanf(f, "%d %d %lf %lf", &r, &c, &v, &vi) == 4;
        else got = fscanf(f, "%d %d %lf", &r, &c, &v) == 3;
        if (!got) { free(e); fclose(f); return -3; }
        r--; c--;
        mtx_ent x = {r, c, ne, v};
        e[ne++] = x;
        if (sym && r != c) { mtx_ent y = {c, r, ne, v}; e[ne++] = y; }
    }
    fclose(f);
    *m = mm; *n = nn; *nnz = ne;
    if (rp) {
        qsort(e, (size_t)ne, sizeof(mtx_ent), cmp_mtx);
        memset(rp, 0, sizeof(int64_t) * ((size_t)mm + 1));
        for (int64_t i = 0; i < ne; i++) { col[i] = e[i].c; val[i] = e[i].v; rp[e[i].r + 1]++; }
        for (int32_t i = 0; i < mm; i++) rp[i + 1] += rp[i];
    }
    free(e);
    return 0;
}
