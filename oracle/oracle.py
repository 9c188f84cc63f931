"""ctypes wrapper of the oracle (liboracle.so) and of the reference-compiled
checker (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  Results come back as dicts
with the reference's field names (tile_ptr, tile_columnidx, tile_nnz,
tile_csr_Ptr, mask, tile_csr_Col, tile_csr_Val, csc_tile_ptr, csc_tile_rowidx),
the same layout gpu_spgemm_amd.DeviceTiles.download() returns.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")


class OraTile(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("tilem", C.c_int32), ("tilen", C.c_int32),
                ("nnz", C.c_int64), ("numtile", C.c_int64),
                ("tile_ptr", C.POINTER(C.c_int64)), ("tile_colidx", C.POINTER(C.c_int32)),
                ("tile_rowidx", C.POINTER(C.c_int32)), ("tile_nnz", C.POINTER(C.c_int64)),
                ("tile_ptrRow", C.POINTER(C.c_uint8)), ("mask", C.POINTER(C.c_uint16)),
                ("tile_col", C.POINTER(C.c_uint8)), ("val64", C.POINTER(C.c_double)),
                ("val32", C.POINTER(C.c_float)), ("csc_tile_ptr", C.POINTER(C.c_int64)),
                ("csc_tile_rowidx", C.POINTER(C.c_int32)), ("numtile_csc", C.c_int64),
                ("is_csc", C.c_int32), ("is_f32", C.c_int32)]


class OraCheck(C.Structure):
    _fields_ = [("total_ref", C.c_int64), ("total_gpu", C.c_int64), ("missing", C.c_int64),
                ("extra", C.c_int64), ("rel_frob", C.c_double), ("max_scaled", C.c_double),
                ("max_abs", C.c_double), ("pass_", C.c_int32)]


_lib = None
_ref = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.ora_flops.restype = C.c_uint64
        _lib.ora_cpu_baseline.restype = C.c_int64
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
    return _ref


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _csr_args(csr):
    rp = np.ascontiguousarray(csr.rowptr, dtype=np.int64)
    col = np.ascontiguousarray(csr.col, dtype=np.int32)
    val = np.ascontiguousarray(csr.val, dtype=np.float64)
    return rp, col, val


def _arr(ptr, n, dt):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dt)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,)).copy()


def tile_to_dict(t, free_fn, rows=16):
    nt = t.numtile_csc if t.is_csc else t.numtile
    d = dict(m=t.m, n=t.n, tilem=t.tilem, tilen=t.tilen, nnz=t.nnz, numtile=t.numtile,
             is_csc=t.is_csc, dtype=1 if t.is_f32 else 0,
             tile_ptr=_arr(t.tile_ptr, t.tilem + 1, np.int64),
             tile_columnidx=_arr(t.tile_colidx, t.numtile, np.int32),
             tile_rowidx=_arr(t.tile_rowidx, t.numtile, np.int32) if t.tile_rowidx else None,
             tile_nnz=_arr(t.tile_nnz, nt + 1, np.int64),
             tile_csr_Ptr=_arr(t.tile_ptrRow, nt * rows, np.uint8),
             mask=_arr(t.mask, nt * rows, np.uint16),
             tile_csr_Col=_arr(t.tile_col, t.nnz, np.uint8),
             tile_csr_Val=(_arr(t.val32, t.nnz, np.float32) if t.is_f32 else _arr(t.val64, t.nnz, np.float64)))
    if t.is_csc:
        d["csc_tile_ptr"] = _arr(t.csc_tile_ptr, t.tilen + 1, np.int64)
        d["csc_tile_rowidx"] = _arr(t.csc_tile_rowidx, t.numtile_csc, np.int32)
    free_fn(C.byref(t))
    return d


def _dict_to_tile(d, keep):
    """Builds an OraTile view over numpy arrays of a dict (no ownership)."""
    t = OraTile()
    t.m, t.n, t.tilem, t.tilen = d["m"], d["n"], d["tilem"], d["tilen"]
    t.nnz, t.numtile, t.is_csc = d["nnz"], d["numtile"], d.get("is_csc", 0)
    t.is_f32 = 1 if d["tile_csr_Val"].dtype == np.float32 else 0

    def put(name, key, ct, dt):
        a = d.get(key)
        if a is None:
            return
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        setattr(t, name, _p(a, ct))

    put("tile_ptr", "tile_ptr", C.c_int64, np.int64)
    put("tile_colidx", "tile_columnidx", C.c_int32, np.int32)
    put("tile_rowidx", "tile_rowidx", C.c_int32, np.int32)
    put("tile_nnz", "tile_nnz", C.c_int64, np.int64)
    put("tile_ptrRow", "tile_csr_Ptr", C.c_uint8, np.uint8)
    put("mask", "mask", C.c_uint16, np.uint16)
    put("tile_col", "tile_csr_Col", C.c_uint8, np.uint8)
    if t.is_f32:
        put("val32", "tile_csr_Val", C.c_float, np.float32)
    else:
        put("val64", "tile_csr_Val", C.c_double, np.float64)
    if t.is_csc:
        put("csc_tile_ptr", "csc_tile_ptr", C.c_int64, np.int64)
        put("csc_tile_rowidx", "csc_tile_rowidx", C.c_int32, np.int32)
        t.numtile_csc = d["numtile"]
    return t


# ---------------------------------------------------------------- oracle calls
def csr2tile(csr, f32=False):
    """tile_format.h:93-222 restated."""
    rp, col, val = _csr_args(csr)
    t = OraTile()
    lib().ora_csr2tile(csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double),
                       1 if f32 else 0, C.byref(t))
    return tile_to_dict(t, lib().ora_tile_free)


def csr2tile_b(csr, f32=False):
    """csr2tile + build_csc_tiles (tile_format.h:424-546) restated."""
    rp, col, val = _csr_args(csr)
    t = OraTile()
    lib().ora_csr2tile(csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double),
                       1 if f32 else 0, C.byref(t))
    lib().ora_build_csc_tiles(C.byref(t), csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32),
                              _p(val, C.c_double))
    return tile_to_dict(t, lib().ora_tile_free)


def spgemm_tiled(A, B, f32=False, steps=3):
    """Steps 1..steps of the tiled method on the CPU; returns the C dict."""
    keep = []
    tA = csr2tile_raw(A, f32, False, keep)
    tB = csr2tile_raw(B, f32, True, keep)
    tC = OraTile()
    lib().ora_step1(C.byref(tA), C.byref(tB), C.byref(tC))
    if steps >= 2:
        lib().ora_step2(C.byref(tA), C.byref(tB), C.byref(tC))
    if steps >= 3:
        lib().ora_step3(C.byref(tA), C.byref(tB), C.byref(tC))
    else:
        tC.nnz = tC.nnz if steps >= 2 else 0
    d = tile_to_dict(tC, lib().ora_tile_free)
    lib().ora_tile_free(C.byref(tA))
    lib().ora_tile_free(C.byref(tB))
    return d


class _Rows:
    """Rows [r0, r1) of a CSR as a CSR of r1 - r0 rows (same columns)."""

    def __init__(self, A, r0, r1):
        a, b = int(A.rowptr[r0]), int(A.rowptr[r1])
        self.m, self.n = r1 - r0, A.n
        self.rowptr = np.asarray(A.rowptr[r0:r1 + 1], dtype=np.int64) - a
        self.col = A.col[a:b]
        self.val = A.val[a:b]


def row_block(A, r0, r1):
    """Rows [r0, r1) of CSR A as a CSR with r1 - r0 rows."""
    return _Rows(A, r0, r1)


def spgemm_tiled_rows(A, B, blocks, f32=False):
    """Steps 1-3 of the tiled method for A's tile-row ranges [lo, hi) in
    `blocks` times all of B (B converted once): one C dict per block, a matrix
    of hi - lo tile-rows (C's tile-rows lo..hi-1 of the whole product, tile_ptr
    and tile_nnz from 0, tile_rowidx relative to lo)."""
    keep = []
    tB = csr2tile_raw(B, f32, True, keep)
    out = []
    try:
        for lo, hi in blocks:
            Ab = _Rows(A, lo * 16, min(hi * 16, A.m))
            tA = csr2tile_raw(Ab, f32, False, keep)
            tC = OraTile()
            lib().ora_step1(C.byref(tA), C.byref(tB), C.byref(tC))
            lib().ora_step2(C.byref(tA), C.byref(tB), C.byref(tC))
            lib().ora_step3(C.byref(tA), C.byref(tB), C.byref(tC))
            out.append(tile_to_dict(tC, lib().ora_tile_free))
            lib().ora_tile_free(C.byref(tA))
    finally:
        lib().ora_tile_free(C.byref(tB))
    return out


def embed_rows(Cd, lo, tilem):
    """A C dict of tile-rows [lo, lo + Cd['tilem']) (download_rows layout) as a
    dict of `tilem` tile-rows whose other tile-rows are empty, for
    gustavson_check over rows of that range."""
    d = dict(Cd)
    tp = np.zeros(tilem + 1, np.int64)
    n = Cd["tilem"]
    tp[lo + 1:lo + n + 1] = Cd["tile_ptr"][1:]
    tp[lo + n + 1:] = Cd["tile_ptr"][-1]
    d["tile_ptr"] = tp
    d["tilem"] = tilem
    d["m"] = tilem * 16
    if Cd.get("tile_rowidx") is not None:
        d["tile_rowidx"] = Cd["tile_rowidx"] + lo
    return d


def csr2tile_raw(csr, f32, as_b, keep):
    rp, col, val = _csr_args(csr)
    keep += [rp, col, val]
    t = OraTile()
    lib().ora_csr2tile(csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double),
                       1 if f32 else 0, C.byref(t))
    if as_b:
        lib().ora_build_csc_tiles(C.byref(t), csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32),
                                  _p(val, C.c_double))
    return t


def gustavson_check(A, B, Cd, row_lo=0, row_hi=None):
    """spgemm_raw_v2.cu:253-359 restated: compare a C tile dict with FP64 Gustavson."""
    if row_hi is None:
        row_hi = A.m
    ra, ca, va = _csr_args(A)
    rb, cb, vb = _csr_args(B)
    keep = []
    tC = _dict_to_tile(Cd, keep)
    out = OraCheck()
    lib().ora_gustavson_check(A.m, B.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(va, C.c_double),
                              _p(rb, C.c_int64), _p(cb, C.c_int32), _p(vb, C.c_double), C.byref(tC),
                              C.c_int32(row_lo), C.c_int32(row_hi), C.byref(out))
    return {k: getattr(out, k) for k, _ in out._fields_}


def gustavson_csr(A, B):
    """Gustavson product with structural union; returns (rowptr, col, val)."""
    ra, ca, va = _csr_args(A)
    rb, cb, vb = _csr_args(B)
    rpc = np.zeros(A.m + 1, dtype=np.int64)
    lib().ora_gustavson_csr(A.m, B.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(va, C.c_double),
                            _p(rb, C.c_int64), _p(cb, C.c_int32), _p(vb, C.c_double),
                            _p(rpc, C.c_int64), None, None)
    nnz = int(rpc[-1])
    colc = np.zeros(max(nnz, 1), dtype=np.int32)
    valc = np.zeros(max(nnz, 1), dtype=np.float64)
    lib().ora_gustavson_csr(A.m, B.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(va, C.c_double),
                            _p(rb, C.c_int64), _p(cb, C.c_int32), _p(vb, C.c_double),
                            _p(rpc, C.c_int64), _p(colc, C.c_int32), _p(valc, C.c_double))
    return rpc, colc[:nnz], valc[:nnz]


def count_product(A, B, nthreads=8):
    """(nnz(C) with the structural union, numtileC of Step 1 with empty tiles)
    of A*B straight from the CSRs, threaded: a whole-product check at sizes
    where C cannot come back to the host."""
    ra, ca, _ = _csr_args(A)
    rb, cb, _ = _csr_args(B)
    nz, nt = C.c_int64(), C.c_int64()
    lib().ora_count_product(A.m, A.n, B.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(rb, C.c_int64),
                            _p(cb, C.c_int32), C.c_int32(nthreads), C.byref(nz), C.byref(nt))
    return nz.value, nt.value


def flops(A, B):
    ra, ca, _ = _csr_args(A)
    rb, _, _ = _csr_args(B)
    return int(lib().ora_flops(A.m, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(rb, C.c_int64)))


def cpu_baseline(A, B, row_lo, row_hi, nthreads=1, convert=False):
    """Times the reference's CPU path on rows [row_lo,row_hi): returns
    (products, gustavson_seconds, conversion_seconds)."""
    ra, ca, va = _csr_args(A)
    rb, cb, vb = _csr_args(B)
    conv, gus = C.c_double(), C.c_double()
    prod = lib().ora_cpu_baseline(A.m, A.n, B.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(va, C.c_double),
                                  _p(rb, C.c_int64), _p(cb, C.c_int32), _p(vb, C.c_double),
                                  C.c_int32(row_lo), C.c_int32(row_hi), C.c_int32(nthreads),
                                  C.c_int32(1 if convert else 0), C.byref(conv), C.byref(gus))
    return int(prod), gus.value, conv.value


def load_mtx(path):
    m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
    lib().ora_load_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), None, None, None)
    rp = np.zeros(m.value + 1, dtype=np.int64)
    col = np.zeros(max(nnz.value, 1), dtype=np.int32)
    val = np.zeros(max(nnz.value, 1), dtype=np.float64)
    lib().ora_load_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), _p(rp, C.c_int64),
                       _p(col, C.c_int32), _p(val, C.c_double))
    return m.value, n.value, rp, col[:nnz.value], val[:nnz.value]


def csr2tile_ts(csr, ts, as_b=False, f32=False):
    """csr2tile_ts (src/tile_format.h:225-324) and, for as_b, build_csc_tiles_ts
    (:327-417): tiles of ts x ts (1 <= ts <= 16), restated with numpy.

    Tiles are the distinct (row // ts, col // ts) pairs, ascending tile column
    within a tile-row; a tile's entries go row by row, each row in CSR storage
    order (duplicates kept), col byte (row<<4)|col, mask bit ts-1-col, row
    offsets the exclusive prefix of the row counts (unsigned char).  A B
    operand keeps tile_ptr / tile_columnidx and re-orders the per-tile data by
    tile column then tile row (csc_tile_ptr / csc_tile_rowidx) with the col
    byte reduced to the local column."""
    m, n = int(csr.m), int(csr.n)
    rp = np.asarray(csr.rowptr, np.int64)
    col = np.asarray(csr.col, np.int64)
    val = np.asarray(csr.val, np.float32 if f32 else np.float64)
    nnz = len(col)
    tilem, tilen = (m + ts - 1) // ts, (n + ts - 1) // ts
    row = np.repeat(np.arange(m, dtype=np.int64), np.diff(rp))
    bi, lr, tc, lc = row // ts, row % ts, col // ts, col % ts
    key = bi * max(tilen, 1) + tc
    tiles = np.unique(key)
    nt = len(tiles)
    tid = np.searchsorted(tiles, key)
    trow, tcol = tiles // max(tilen, 1), tiles % max(tilen, 1)
    tile_ptr = np.zeros(tilem + 1, np.int64)
    np.add.at(tile_ptr, trow + 1, 1)
    tile_ptr = np.cumsum(tile_ptr)
    cnt = np.bincount(tid * ts + lr, minlength=nt * ts).reshape(nt, ts)
    ptr = ((np.cumsum(cnt, axis=1) - cnt) & 0xFF).astype(np.uint8)
    tnz = np.zeros(nt + 1, np.int64)
    tnz[1:] = np.cumsum(cnt.sum(axis=1))
    order = np.lexsort((np.arange(nnz), lr, tid))
    mask = np.zeros(nt * ts, np.int64)
    np.bitwise_or.at(mask, tid * ts + lr, np.left_shift(1, ts - 1 - lc))
    d = dict(m=m, n=n, tilem=tilem, tilen=tilen, nnz=nnz, numtile=nt, is_csc=0, dtype=1 if f32 else 0,
             tile_ptr=tile_ptr, tile_columnidx=tcol.astype(np.int32), tile_rowidx=None, tile_nnz=tnz,
             tile_csr_Ptr=ptr.reshape(-1), mask=mask.astype(np.uint16),
             tile_csr_Col=((lr << 4) | lc)[order].astype(np.uint8), tile_csr_Val=val[order])
    if not as_b:
        return d
    perm = np.lexsort((trow, tcol))               # CSC tile order: tile column, then tile row
    cp = np.zeros(tilen + 1, np.int64)
    np.add.at(cp, tcol + 1, 1)
    starts, sizes = tnz[:-1][perm], np.diff(tnz)[perm]
    gather = np.repeat(starts - np.r_[0, np.cumsum(sizes)[:-1]], sizes) + np.arange(nnz)
    nz_new = np.zeros(nt + 1, np.int64)
    nz_new[1:] = np.cumsum(sizes)
    d.update(is_csc=1, csc_tile_ptr=np.cumsum(cp), csc_tile_rowidx=trow[perm].astype(np.int32), tile_nnz=nz_new,
             tile_csr_Ptr=ptr[perm].reshape(-1), mask=d["mask"].reshape(nt, ts)[perm].reshape(-1),
             tile_csr_Col=(d["tile_csr_Col"][gather] & 15).astype(np.uint8), tile_csr_Val=d["tile_csr_Val"][gather])
    return d


def build_bitmasks(dA, dB):
    """build_bitmasks (src/tile_format.h:628-645) restated with numpy from tile
    dicts: A's tile-rows and B's tile columns (CSC-tile order) as bitmaps over
    the k tile index, bm_len = ceil(A.tilen / 32), bit 31 - k % 32 of word k / 32."""
    bl = (int(dA["tilen"]) + 31) // 32

    def bits(n, ptr, idx):
        bm = np.zeros((n, bl), np.uint32)
        ptr = np.asarray(ptr, np.int64)
        lst = np.repeat(np.arange(n), np.diff(ptr))
        k = np.asarray(idx, np.int64)
        np.bitwise_or.at(bm, (lst, k // 32), (np.uint32(1) << (31 - k % 32).astype(np.uint32)))
        return bm

    return (bits(int(dA["tilem"]), dA["tile_ptr"], dA["tile_columnidx"]),
            bits(int(dB["tilen"]), dB["csc_tile_ptr"], dB["csc_tile_rowidx"]))


def tile_csc(d):
    """Within-tile CSC of every tile of a tile dict (the download() layout, data
    arrays in their own tile order): the reference's tile_csc_Ptr/Row/Val of a
    B operand (src/tile_format.h:547-584) or tile_a_csc_Ptr/Row/Val of row
    tiles (build_a_tile_csc, src/tile_format.h:589-627).

    Restated as a stable sort of the entries, in storage order, by (tile, local
    column): the reference's counting scatter walks each tile's rows ascending
    and entries in storage order (:567-582, :609-625), which is that order.
    Ptr is the exclusive prefix of the column counts, kept as unsigned char."""
    tnz = np.asarray(d["tile_nnz"], np.int64)
    nt = len(tnz) - 1
    nnz = int(tnz[-1])
    tile_of = np.repeat(np.arange(nt, dtype=np.int64), np.diff(tnz))
    local = np.arange(nnz, dtype=np.int64) - tnz[tile_of]
    ptr = np.asarray(d["tile_csr_Ptr"], np.int64).reshape(nt, 16)
    # row r holds entries [ptr[r], ptr[r+1]): the row is the count of later row
    # starts (r' = 1..15) at or below the entry
    row = (ptr[tile_of, 1:] <= local[:, None]).sum(axis=1).astype(np.uint8) if nnz else np.zeros(0, np.uint8)
    col = np.asarray(d["tile_csr_Col"], np.uint8).astype(np.int64) & 15
    order = np.lexsort((np.arange(nnz), col, tile_of))
    cnt = np.bincount(tile_of * 16 + col, minlength=nt * 16).reshape(nt, 16)
    cptr = (np.cumsum(cnt, axis=1) - cnt).astype(np.int64) & 0xFF
    return dict(tile_csc_Ptr=cptr.astype(np.uint8).reshape(-1), tile_csc_Row=row[order],
                tile_csc_Val=np.asarray(d["tile_csr_Val"])[order])


# ---------------------------------------------------------------- reference-compiled checker
def ref_csr2tile(csr, as_b=False):
    """The reference's own csr2tile (+ build_csc_tiles) compiled from its sources."""
    rp, col, val = _csr_args(csr)
    t = OraTile()
    fn = ref().ref_csr2tile_csc if as_b else ref().ref_csr2tile
    fn(csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double), C.byref(t))
    return tile_to_dict(t, ref().ref_tile_free)


def ref_csr2tile_ts(csr, ts, as_b=False):
    """The reference's own csr2tile_ts (+ build_csc_tiles_ts) compiled from its sources."""
    rp, col, val = _csr_args(csr)
    t = OraTile()
    ref().ref_csr2tile_ts(csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double), int(ts),
                          1 if as_b else 0, C.byref(t))
    return tile_to_dict(t, ref().ref_tile_free, rows=ts)


def ref_tile_csc(csr, as_b=False):
    """The reference's own within-tile CSC (build_csc_tiles :547-584 for a B
    operand, build_a_tile_csc :589-627 for row tiles), compiled from its sources."""
    d = ref_csr2tile(csr, as_b)
    nt, nnz = len(d["tile_nnz"]) - 1, int(d["tile_nnz"][-1])
    rp, col, val = _csr_args(csr)
    ptr = np.zeros(max(nt * 16, 1), np.uint8)
    row = np.zeros(max(nnz, 1), np.uint8)
    v = np.zeros(max(nnz, 1), np.float64)
    ref().ref_tile_csc(csr.m, csr.n, _p(rp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double), 1 if as_b else 0,
                       _p(ptr, C.c_uint8), _p(row, C.c_uint8), _p(v, C.c_double))
    return dict(tile_csc_Ptr=ptr[:nt * 16], tile_csc_Row=row[:nnz], tile_csc_Val=v[:nnz])


def ref_build_bitmasks(A, B):
    """The reference's own build_bitmasks (tile_format.h:628-645) on csr2tile(A)
    and build_csc_tiles(B), compiled from its sources."""
    ra, ca, va = _csr_args(A)
    rb, cb, vb = _csr_args(B)
    bl = C.c_int32()
    ref().ref_build_bitmasks(A.m, A.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(va, C.c_double), B.m, B.n,
                             _p(rb, C.c_int64), _p(cb, C.c_int32), _p(vb, C.c_double), None, None, C.byref(bl))
    tma, tnb = (A.m + 15) // 16, (B.n + 15) // 16
    bmA = np.zeros(max(tma * bl.value, 1), np.uint32)
    bmB = np.zeros(max(tnb * bl.value, 1), np.uint32)
    ref().ref_build_bitmasks(A.m, A.n, _p(ra, C.c_int64), _p(ca, C.c_int32), _p(va, C.c_double), B.m, B.n,
                             _p(rb, C.c_int64), _p(cb, C.c_int32), _p(vb, C.c_double), _p(bmA, C.c_uint32),
                             _p(bmB, C.c_uint32), C.byref(bl))
    return bmA[:tma * bl.value].reshape(tma, bl.value), bmB[:tnb * bl.value].reshape(tnb, bl.value)


def ref_load_mtx(path):
    m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
    ref().ref_load_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), None, None, None)
    rp = np.zeros(m.value + 1, dtype=np.int64)
    col = np.zeros(max(nnz.value, 1), dtype=np.int32)
    val = np.zeros(max(nnz.value, 1), dtype=np.float64)
    ref().ref_load_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), _p(rp, C.c_int64),
                       _p(col, C.c_int32), _p(val, C.c_double))
    return m.value, n.value, rp, col[:nnz.value], val[:nnz.value]
