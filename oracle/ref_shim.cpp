// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Exposes the reference's own host conversion functions, compiled from
// src/tile_format.h where it lies in the reference checkout (include path set
// by oracle/Makefile), through a C ABI that fills the oracle's ora_tile struct.
// Used only to validate the oracle's restatement and to write golden fixtures.
// Nothing here is shipped or called by the product.
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "oracle.h"
#include "tile_format.h"   // reference: src/tile_format.h

template <typename T>
static T *dup(const T *src, size_t n) {
    T *d = (T *)malloc(sizeof(T) * (n ? n : 1));
    if (n) memcpy(d, src, sizeof(T) * n);
    return d;
}

static int64_t *widen(const int *src, size_t n) {
    int64_t *d = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) d[i] = src[i];
    return d;
}

static HostCSR make_csr(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val) {
    HostCSR h;
    h.m = m; h.n = n; h.nnz = (int)rp[m];
    h.row_ptr.resize(m + 1);
    for (int i = 0; i <= m; i++) h.row_ptr[i] = (int)rp[i];
    h.col_idx.assign(col, col + h.nnz);
    h.val.assign(val, val + h.nnz);
    return h;
}

// Copy a reference TileMatrix into an ora_tile (values as f64); rows = the
// tile size (per-tile entries of tile_csr_Ptr and mask).
static void to_ora(const TileMatrix &T, bool is_B, ora_tile *o, int rows = TILE_SIZE) {
    memset(o, 0, sizeof *o);
    o->m = T.m; o->n = T.n; o->tilem = T.tilem; o->tilen = T.tilen;
    o->numtile = T.numtile;
    int nt = is_B ? T.numtileB_csc : T.numtile;
    o->nnz = T.tile_nnz[nt];
    o->tile_ptr = widen(T.tile_ptr, T.tilem + 1);
    o->tile_colidx = dup(T.tile_columnidx, T.numtile);
    o->tile_nnz = widen(T.tile_nnz, nt + 1);
    o->tile_ptrRow = dup(T.tile_csr_Ptr, (size_t)nt * rows);
    o->mask = dup(T.mask, (size_t)nt * rows);
    o->tile_col = dup(T.tile_csr_Col, (size_t)o->nnz);
    o->val64 = dup(T.tile_csr_Val, (size_t)o->nnz);
    if (is_B) {
        o->csc_tile_ptr = widen(T.csc_tile_ptr, T.tilen + 1);
        o->csc_tile_rowidx = dup(T.csc_tile_rowidx, T.numtileB_csc);
        o->numtile_csc = T.numtileB_csc;
        o->is_csc = 1;
    }
}

extern "C" {

int ref_csr2tile(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val,
                 ora_tile *out) {
    HostCSR h = make_csr(m, n, rp, col, val);
    TileMatrix T = csr2tile(h);
    to_ora(T, false, out);
    free_tile_matrix(T);
    return 0;
}

int ref_csr2tile_csc(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val,
                     ora_tile *out) {
    HostCSR h = make_csr(m, n, rp, col, val);
    TileMatrix T = csr2tile(h);
    build_csc_tiles(T, h);
    to_ora(T, true, out);
    free_tile_matrix(T);
    return 0;
}

// The reference's csr2tile_ts (tile_format.h:225-324), and for as_b its
// build_csc_tiles_ts (:327-417), at tile size ts.
int ref_csr2tile_ts(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val, int32_t ts,
                    int32_t as_b, ora_tile *out) {
    HostCSR h = make_csr(m, n, rp, col, val);
    TileMatrix T = csr2tile_ts(h, ts);
    if (as_b) build_csc_tiles_ts(T, h, ts);
    to_ora(T, as_b != 0, out, ts);
    free_tile_matrix(T);
    return 0;
}

// Reference bitmaps (tile_format.h:628-645) for A (CSR tiles) and B (CSC tiles).
int ref_build_bitmasks(int32_t mA, int32_t nA, const int64_t *rpA, const int32_t *colA,
                       const double *valA, int32_t mB, int32_t nB, const int64_t *rpB,
                       const int32_t *colB, const double *valB, uint32_t *bmA, uint32_t *bmB,
                       int32_t *bm_len) {
    HostCSR a = make_csr(mA, nA, rpA, colA, valA), b = make_csr(mB, nB, rpB, colB, valB);
    TileMatrix TA = csr2tile(a), TB = csr2tile(b);
    build_csc_tiles(TB, b);
    unsigned int *pa, *pb;
    int len;
    build_bitmasks(TA, TB, pa, pb, len);
    *bm_len = len;
    if (bmA) memcpy(bmA, pa, sizeof(unsigned) * (size_t)TA.tilem * len);
    if (bmB) memcpy(bmB, pb, sizeof(unsigned) * (size_t)TB.tilen * len);
    free(pa); free(pb);
    free_tile_matrix(TA); free_tile_matrix(TB);
    return 0;
}

// The reference's within-tile CSC: of a B operand (build_csc_tiles,
// tile_format.h:547-584) or of row tiles (build_a_tile_csc, :589-627).  The
// caller sizes the outputs from ref_csr2tile(_csc): ptr [nt*16], row/val [nnz].
int ref_tile_csc(int32_t m, int32_t n, const int64_t *rp, const int32_t *col, const double *val, int32_t as_b,
                 uint8_t *ptr, uint8_t *row, double *vout) {
    HostCSR h = make_csr(m, n, rp, col, val);
    TileMatrix T = csr2tile(h);
    int nt;
    const unsigned char *p, *r;
    const TVAL *v;
    if (as_b) {
        build_csc_tiles(T, h);
        nt = T.numtileB_csc; p = T.tile_csc_Ptr; r = T.tile_csc_Row; v = T.tile_csc_Val;
    } else {
        build_a_tile_csc(T);
        nt = T.numtile; p = T.tile_a_csc_Ptr; r = T.tile_a_csc_Row; v = T.tile_a_csc_Val;
    }
    const int nnz = T.tile_nnz[nt];
    memcpy(ptr, p, (size_t)nt * TILE_SIZE);
    memcpy(row, r, (size_t)nnz);
    memcpy(vout, v, sizeof(double) * (size_t)nnz);
    free_tile_matrix(T);
    return 0;
}

// Reference MTX reader (tile_format.h:51-90).  Two-phase: rp == NULL returns sizes.
int ref_load_mtx(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int64_t *rp, int32_t *col,
                 double *val) {
    HostCSR h = load_mtx(path);
    *m = h.m; *n = h.n; *nnz = h.nnz;
    if (rp) {
        for (int i = 0; i <= h.m; i++) rp[i] = h.row_ptr[i];
        memcpy(col, h.col_idx.data(), sizeof(int) * h.nnz);
        memcpy(val, h.val.data(), sizeof(double) * h.nnz);
    }
    return 0;
}

void ref_tile_free(ora_tile *t) { ora_tile_free(t); }

}  // extern "C"
