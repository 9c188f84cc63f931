"""gpu_spgemm_amd -- MI355X tiled SpGEMM (C = A*B in 16x16 tiles).

Python mirror of the reference's host path (src/spgemm_raw_v2.cu:61-372) over
the C ABI of libtsg.so (include/tsg.h).  Everything numeric runs in the HIP
kernels of libtsg; this module only moves arrays and calls entry points.  The
library must be built (gpu_spgemm_amd.build) -- there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsg.so")

F64, F32 = 0, 1
HOST, DEVICE = 0, 1
SCHED_PLAIN, SCHED_LEARNED = 0, 1

STATUS = {0: "OK", -1: "ERR_ARG", -2: "ERR_HIP", -3: "ERR_OOM", -4: "ERR_OVERFLOW",
          -5: "ERR_UNSUPPORTED", -6: "ERR_IO", -7: "ERR_NODEVICE"}


class TsgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class CSRStruct(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64), ("dtype", C.c_int32),
                ("location", C.c_int32), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int32)), ("val", C.c_void_p)]


class TileStruct(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("tilem", C.c_int32), ("tilen", C.c_int32),
                ("nnz", C.c_int64), ("numtile", C.c_int64), ("dtype", C.c_int32),
                ("location", C.c_int32), ("is_csc", C.c_int32), ("flags", C.c_int32),
                ("tile_ptr", C.POINTER(C.c_int64)), ("tile_columnidx", C.POINTER(C.c_int32)),
                ("tile_rowidx", C.POINTER(C.c_int32)), ("tile_nnz", C.POINTER(C.c_int64)),
                ("tile_csr_Val", C.c_void_p), ("tile_csr_Col", C.POINTER(C.c_uint8)),
                ("tile_csr_Ptr", C.POINTER(C.c_uint8)), ("mask", C.POINTER(C.c_uint16)),
                ("csc_tile_ptr", C.POINTER(C.c_int64)), ("csc_tile_rowidx", C.POINTER(C.c_int32)),
                ("csr_tile_nnz", C.POINTER(C.c_int64)), ("csr_mask", C.POINTER(C.c_uint16)),
                ("csr_col_mask", C.POINTER(C.c_uint16)), ("csr_col_Ptr", C.POINTER(C.c_uint8)),
                ("csr_col_Val", C.c_void_p), ("row_nnz_ptr", C.POINTER(C.c_int64)),
                ("row_ptr", C.POINTER(C.c_int64)), ("row_col", C.POINTER(C.c_int32)), ("row_val", C.c_void_p),
                ("val_perm", C.POINTER(C.c_int64)), ("val_perm_col", C.POINTER(C.c_int64)),
                ("tile_csc_Ptr", C.POINTER(C.c_uint8)), ("tile_csc_Row", C.POINTER(C.c_uint8)),
                ("tile_csc_Val", C.c_void_p), ("tile_size", C.c_int32), ("reserved_ts", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("scheduler", C.c_int32), ("numeric_variant", C.c_int32),
                ("mfma_min_products", C.c_int32), ("timing", C.c_int32), ("pred_scale", C.c_double),
                ("tuning", C.c_int32), ("sum_order", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("var_mask", C.c_int32), ("rg_density", C.c_int32), ("mfma_min_products", C.c_int32),
                ("sampled", C.c_int32), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("numtileC", C.c_int64), ("nnzC", C.c_int64), ("products", C.c_uint64),
                ("ms_step1", C.c_double), ("ms_step2", C.c_double), ("ms_step3", C.c_double),
                ("ms_total", C.c_double), ("fallback_used", C.c_int32), ("host_syncs", C.c_int32),
                ("pred_numtileC", C.c_int64), ("pred_nnzC", C.c_int64), ("pred_reused", C.c_int32),
                ("numeric_kernels", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libtsg.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `python -m gpu_spgemm_amd.build` first")
        _lib = C.CDLL(LIB_PATH)
        _lib.tsg_last_error_string.restype = C.c_char_p
        _lib.tsg_stream.restype = C.c_void_p
        _lib.tsg_tile_set_values.argtypes = [C.c_void_p, C.POINTER(TileStruct), C.c_void_p]
    return _lib


def _check(rc):
    if rc != 0:
        raise TsgError(rc, lib().tsg_last_error_string().decode(errors="replace"))


def device_count():
    n = C.c_int(0)
    _check(lib().tsg_device_count(C.byref(n)))
    return n.value


def _np(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                 shape=(count,)).copy()


class HostCSR:
    """CSR on the host as numpy arrays (rowptr int64, col int32, val f64/f32)."""

    def __init__(self, m, n, rowptr, col, val):
        self.m, self.n = int(m), int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val)
        if self.val.dtype not in (np.float64, np.float32):
            self.val = self.val.astype(np.float64)

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    @property
    def dtype(self):
        return F32 if self.val.dtype == np.float32 else F64

    def struct(self):
        s = CSRStruct()
        s.m, s.n, s.nnz, s.dtype, s.location = self.m, self.n, self.nnz, self.dtype, HOST
        s.rowptr = self.rowptr.ctypes.data_as(C.POINTER(C.c_int64))
        s.col = self.col.ctypes.data_as(C.POINTER(C.c_int32))
        s.val = self.val.ctypes.data_as(C.c_void_p)
        return s

    @classmethod
    def _take(cls, s):
        """Copies a library-allocated host CSR into numpy and frees it."""
        vdt = np.float32 if s.dtype == F32 else np.float64
        out = cls(s.m, s.n, _np(s.rowptr, s.m + 1, np.int64), _np(s.col, s.nnz, np.int32),
                  _np(s.val, s.nnz, vdt))
        _check(lib().tsg_csr_free_host(C.byref(s)))
        return out

    def astype(self, dtype):
        vdt = np.float32 if dtype == F32 else np.float64
        return HostCSR(self.m, self.n, self.rowptr, self.col, self.val.astype(vdt))

    def transpose(self):
        """B = A^T, the reference's --aat path (src/spgemm_raw_v2.cu:78-89)."""
        s, o = self.struct(), CSRStruct()
        _check(lib().tsg_csr_transpose_host(C.byref(s), C.byref(o)))
        return HostCSR._take(o)

    def products(self, B):
        """flops = sum_{(i,k) in A} nnz(B[k,:]) (src/spgemm_raw_v2.cu:95-99)."""
        a, b, f = self.struct(), B.struct(), C.c_uint64(0)
        _check(lib().tsg_csr_products(C.byref(a), C.byref(b), C.byref(f)))
        return f.value

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.m, self.n))


def _gen(fn, *args):
    o = CSRStruct()
    _check(fn(*args, C.byref(o)))
    return HostCSR._take(o)


def gen_er(n, per_row, seed=0):
    return _gen(lib().tsg_gen_er, C.c_int32(n), C.c_int32(per_row), C.c_uint64(seed))


def gen_er_var(n, lo, hi, seed=0):
    return _gen(lib().tsg_gen_er_var, C.c_int32(n), C.c_int32(lo), C.c_int32(hi), C.c_uint64(seed))


def gen_laplace2d(nx, ny):
    return _gen(lib().tsg_gen_laplace2d, C.c_int32(nx), C.c_int32(ny))


def gen_laplace3d(nx, ny, nz):
    return _gen(lib().tsg_gen_laplace3d, C.c_int32(nx), C.c_int32(ny), C.c_int32(nz))


def gen_stencil27(nx, ny, nz):
    return _gen(lib().tsg_gen_stencil27, C.c_int32(nx), C.c_int32(ny), C.c_int32(nz))


def gen_aggregation(nx, ny, nz):
    return _gen(lib().tsg_gen_aggregation, C.c_int32(nx), C.c_int32(ny), C.c_int32(nz))


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, dtype=F32):
    return _gen(lib().tsg_gen_rmat, C.c_int32(scale), C.c_int32(edge_factor), C.c_double(a),
                C.c_double(b), C.c_double(c), C.c_uint64(seed), C.c_int32(dtype))


def gen_band(n, half, seed=0):
    """Dense band |i - j| <= half (values U[-1,1] from numpy's default_rng(seed)):
    tile-rows of full 16x16 tiles, the case for matrix cores."""
    lo = np.maximum(np.arange(n) - half, 0)
    hi = np.minimum(np.arange(n) + half + 1, n)
    cnt = hi - lo
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(cnt)
    col = (np.arange(rp[-1]) - np.repeat(rp[:-1], cnt) + np.repeat(lo, cnt)).astype(np.int32)
    val = np.random.default_rng(seed).uniform(-1, 1, int(rp[-1]))
    return HostCSR(n, n, rp, col, val)


def gen_batch_uneven(count=512, seed=4, lo=256, hi=65536):
    """The uneven batch: count problems, n_i = round(exp(U(ln lo, ln hi)));
    even i: 7-point 3D Laplacian on a near-cubic nx*ny*nz grid (values 6/-1),
    odd i: Erdos-Renyi with nnz/row drawn uniformly from {2..32} per matrix
    (values U[-1,1]); sizes and draws from numpy's default_rng(seed).
    Returns [(name, HostCSR)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        if i % 2 == 0:
            nz = max(1, int(round(n ** (1.0 / 3.0))))
            ny = max(1, int(round((n / nz) ** 0.5)))
            nx = max(1, int(round(n / (nz * ny))))
            out.append((f"lap3d_{nx}x{ny}x{nz}", gen_laplace3d(nx, ny, nz)))
        else:
            per = int(rng.integers(2, 33))
            out.append((f"er_{n}_{per}", gen_er(n, per, seed=1000 + i)))
    return out


def load_mtx(path):
    """MTX reader with load_mtx semantics (src/tile_format.h:51-90)."""
    return _gen(lib().tsg_load_mtx, os.fsencode(path))


def write_mtx(path, csr):
    s = csr.struct()
    _check(lib().tsg_write_mtx(os.fsencode(path), C.byref(s)))


class Context:
    """One GPU, one HIP stream, one device memory pool (tsg_create)."""

    def __init__(self, device=0):
        self._p = C.c_void_p()
        _check(lib().tsg_create(C.c_int(device), C.byref(self._p)))

    def close(self):
        if self._p:
            lib().tsg_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._p

    def synchronize(self):
        _check(lib().tsg_synchronize(self._p))

    def device_name(self):
        buf = C.create_string_buffer(256)
        _check(lib().tsg_device_name(self._p, buf, C.c_size_t(256)))
        return buf.value.decode()

    def stream(self):
        """The hipStream_t every call of this context is queued on (an int)."""
        return lib().tsg_stream(self._p) or 0

    def set_stream(self, stream):
        """Queues every later call on the caller's hipStream_t (an int handle,
        e.g. torch.cuda.current_stream().cuda_stream); 0 or None: back to the
        context's own stream."""
        _check(lib().tsg_set_stream(self._p, C.c_void_p(stream or None)))

    def scan_i64(self, dev_ptr, n, zero_first=False, total_ptr=0):
        """In-place inclusive scan of n int64 at a device address (asynchronous,
        on the context's stream); total_ptr (device, optional) gets data[n-1]."""
        _check(lib().tsg_scan_i64(self._p, C.c_void_p(dev_ptr), C.c_int64(n), C.c_int32(1 if zero_first else 0),
                                  C.c_void_p(total_ptr or None)))

    def malloc(self, nbytes):
        """A raw device array of nbytes from the context's pool (an int address)."""
        p = C.c_void_p()
        _check(lib().tsg_device_malloc(self._p, C.c_size_t(nbytes), C.byref(p)))
        return p.value

    def free(self, ptr):
        _check(lib().tsg_device_free(self._p, C.c_void_p(ptr)))

    def to_device(self, arr):
        """Copies a numpy array into a new device array; returns its address."""
        a = np.ascontiguousarray(arr)
        p = self.malloc(a.nbytes)
        _check(lib().tsg_memcpy(self._p, C.c_void_p(p), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes)))
        return p

    def to_host(self, ptr, count, dtype):
        """Copies count elements of dtype from a device address into numpy."""
        out = np.empty(count, dtype=dtype)
        _check(lib().tsg_memcpy(self._p, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes)))
        return out

    def copy(self, dst, src, nbytes):
        """Copies nbytes between two addresses (device or host; blocking)."""
        _check(lib().tsg_memcpy(self._p, C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes)))

    def pool_trim(self):
        """Returns the pool's cached (freed) blocks to the device."""
        _check(lib().tsg_pool_trim(self._p))

    def pool_stats(self):
        a, b = C.c_int64(), C.c_int64()
        _check(lib().tsg_pool_stats(self._p, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- device objects
    def upload_csr(self, h):
        return DeviceCSR(self, h)

    def csr2tile(self, dcsr):
        """csr2tile (src/tile_format.h:93-222) on the GPU."""
        t = DeviceTiles(self)
        _check(lib().tsg_csr2tile(self._p, C.byref(dcsr.s), C.byref(t.s)))
        return t

    def tiles_from_csr(self, h, as_b=False, ts=16):
        """Host CSR -> device tiles of ts x ts (csr2tile, or csr2tile_ts for
        ts != 16; as_b: then build_csc_tiles / build_csc_tiles_ts)."""
        d = self.upload_csr(h)
        try:
            if ts == 16:
                t = self.csr2tile(d)
            else:
                t = DeviceTiles(self)
                _check(lib().tsg_csr2tile_ts(self._p, C.byref(d.s), int(ts), C.byref(t.s)))
        finally:
            d.free()
        if as_b:
            t.build_csc_tiles()
        return t

    def spgemm(self, A, B, scheduler=SCHED_PLAIN, timing=False, variant=0, pred_scale=1.0,
               mfma_min_products=0, tuning=0, sum_order=0):
        """C = A*B (Steps 1-3 of src/spgemm_raw_v2.cu:111-240); returns (C, stats dict).
        sum_order=1 lets the heavy-row numeric kernel add products in any order."""
        o = Options(scheduler, variant, mfma_min_products, 1 if timing else 0, pred_scale, tuning, sum_order)
        st = Stats()
        Cm = DeviceTiles(self)
        _check(lib().tsg_spgemm(self._p, C.byref(A.s), C.byref(B.s), C.byref(Cm.s), C.byref(o),
                                C.byref(st)))
        return Cm, st.as_dict()

    def spgemm_noalloc(self, A, B, out, cap_tiles, cap_nnz):
        """C = A*B with no device allocation (tsg_spgemm_noalloc): C is written
        into the caller's device arrays, `out` a dict of device addresses (ints)
        named as TileStruct's fields (tile_ptr, tile_rowidx, tile_columnidx,
        tile_nnz, mask, tile_csr_Ptr, tile_csr_Val, tile_csr_Col).  Returns
        (TileStruct describing C in those arrays, stats dict); raises TsgError
        (ERR_OVERFLOW, stats in .stats) when a capacity is too small."""
        o = Options(SCHED_PLAIN, 0, 0, 0, 1.0, 0)
        st = Stats()
        c = TileStruct()
        ftype = dict(TileStruct._fields_)
        for k, v in out.items():
            t = ftype[k]
            setattr(c, k, v if t is C.c_void_p else C.cast(C.c_void_p(v), t))
        rc = lib().tsg_spgemm_noalloc(self._p, C.byref(A.s), C.byref(B.s), C.byref(c), C.c_int64(cap_tiles),
                                      C.c_int64(cap_nnz), C.byref(o), C.byref(st))
        if rc != 0:
            err = TsgError(rc, lib().tsg_last_error_string().decode(errors="replace"))
            err.stats = st.as_dict()
            raise err
        return c, st.as_dict()

    def spgemm_batched(self, As, Bs, scheduler=SCHED_PLAIN, timing=False, variant=0, pred_scale=1.0,
                       mfma_min_products=0, tuning=0, sum_order=0):
        """C_i = A_i*B_i for every pair in one pipeline run; returns ([C_i], [stats_i])."""
        n = len(As)
        assert len(Bs) == n
        o = Options(scheduler, variant, mfma_min_products, 1 if timing else 0, pred_scale, tuning, sum_order)
        a = (TileStruct * max(n, 1))(*[x.s for x in As])
        b = (TileStruct * max(n, 1))(*[x.s for x in Bs])
        c = (TileStruct * max(n, 1))()
        st = (Stats * max(n, 1))()
        _check(lib().tsg_spgemm_batched(self._p, C.c_int32(n), a, b, c, C.byref(o), st))
        out = []
        for i in range(n):
            t = DeviceTiles(self)
            t.s = TileStruct.from_buffer_copy(c[i])
            out.append(t)
        return out, [st[i].as_dict() for i in range(n)]

    # ---- new values, same structure
    def set_values(self, T, host_csr):
        """Rewrites every value copy of T (made by csr2tile / tiles_from_csr)
        from a host CSR with T's structure (its values in CSR order)."""
        d = self.upload_csr(host_csr)
        try:
            # a device pointer: pass it as void*, never as a default (32-bit) int
            _check(lib().tsg_tile_set_values(self._p, C.byref(T.s), C.c_void_p(d.s.val)))
        finally:
            self.synchronize()
            d.free()

    def build_bitmasks(self, A, B):
        """build_bitmasks (src/tile_format.h:628-645) on the GPU: returns
        (bmA [tilemA, bm_len], bmB [tilenB, bm_len]) as uint32 numpy arrays."""
        bl = C.c_int32()
        _check(lib().tsg_build_bitmasks(self._p, C.byref(A.s), C.byref(B.s), None, None, C.byref(bl)))
        bmA = np.zeros((A.s.tilem, bl.value), np.uint32)
        bmB = np.zeros((B.s.tilen, bl.value), np.uint32)
        _check(lib().tsg_build_bitmasks(self._p, C.byref(A.s), C.byref(B.s),
                                        bmA.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        bmB.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(bl)))
        return bmA, bmB

    def batch_graph(self, As, Bs, slack=1.25, tuning=0, sum_order=0):
        """One captured hipGraph for the whole batch product (tsg_batch_graph_*)."""
        return BatchGraph(self, As, Bs, slack, Options(SCHED_PLAIN, 0, 0, 0, 1.0, tuning, sum_order))

    def plan(self, As, Bs, variant=0, mfma_min_products=0, tuning=0, sum_order=0):
        """Product (or batch, for lists) plus a plan that recomputes the values
        from the operands' current values as one captured hipGraph."""
        return Plan(self, As, Bs, Options(SCHED_PLAIN, variant, mfma_min_products, 0, 1.0, tuning, sum_order))

    # ---- autotuner
    def autotune(self, As, Bs, reps=3):
        """Times every kernel-variant configuration on the product (or batch, for
        lists) and stores the fastest for the operands' class; returns (best, report)."""
        As = As if isinstance(As, (list, tuple)) else [As]
        Bs = Bs if isinstance(Bs, (list, tuple)) else [Bs]
        n = len(As)
        a = (TileStruct * n)(*[x.s for x in As])
        b = (TileStruct * n)(*[x.s for x in Bs])
        rep = (TuneEntry * 64)()
        nrep = C.c_int32(0)
        best = TuneEntry()
        _check(lib().tsg_autotune(self._p, C.c_int32(n), a, b, C.c_int32(reps), rep, C.c_int32(64),
                                  C.byref(nrep), C.byref(best)))
        return best.as_dict(), [rep[i].as_dict() for i in range(nrep.value)]

    def autotune_query(self, As, Bs):
        As = As if isinstance(As, (list, tuple)) else [As]
        Bs = Bs if isinstance(Bs, (list, tuple)) else [Bs]
        n = len(As)
        a = (TileStruct * n)(*[x.s for x in As])
        b = (TileStruct * n)(*[x.s for x in Bs])
        cfg = TuneEntry()
        found = lib().tsg_autotune_query(self._p, C.c_int32(n), a, b, C.byref(cfg))
        return cfg.as_dict() if found == 1 else None

    def autotune_set(self, As, Bs, var_mask, rg_density, mfma_min_products=0):
        As = As if isinstance(As, (list, tuple)) else [As]
        Bs = Bs if isinstance(Bs, (list, tuple)) else [Bs]
        n = len(As)
        a = (TileStruct * n)(*[x.s for x in As])
        b = (TileStruct * n)(*[x.s for x in Bs])
        cfg = TuneEntry(var_mask, rg_density, mfma_min_products, 0, 0.0)
        _check(lib().tsg_autotune_set(self._p, C.c_int32(n), a, b, C.byref(cfg)))

    def autotune_clear(self):
        _check(lib().tsg_autotune_clear(self._p))

    # ---- learned scheduler (MLP 8 -> 32 -> 32 -> 2 on the GPU)
    def learned_collect(self, A, B):
        """Adds A*B's per-tile-row (features, exact targets) to the training set."""
        _check(lib().tsg_learned_collect(self._p, C.byref(A.s), C.byref(B.s)))

    def learned_clear(self):
        _check(lib().tsg_learned_clear(self._p))

    def learned_train(self, epochs=200, batch=4096, lr=3e-3, seed=0):
        """Adam on the collected set; returns the last epoch's mean loss."""
        loss = C.c_float(0)
        _check(lib().tsg_learned_train(self._p, C.c_int32(epochs), C.c_int32(batch), C.c_float(lr),
                                       C.c_uint64(seed), C.byref(loss)))
        return loss.value

    @staticmethod
    def learned_num_params():
        return int(lib().tsg_learned_num_params())

    def learned_weights(self):
        n = self.learned_num_params()
        w = np.zeros(n, np.float32)
        _check(lib().tsg_learned_get_weights(self._p, w.ctypes.data_as(C.c_void_p), C.c_int32(n)))
        return w

    def learned_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(lib().tsg_learned_set_weights(self._p, w.ctypes.data_as(C.c_void_p), C.c_int32(w.size)))

    def learned_calibration(self, set_to=None):
        """Log-space offsets (tiles, nonzeros) added to the predictions; optionally sets them."""
        get = (C.c_float * 2)()
        st = None
        if set_to is not None:
            st = (C.c_float * 2)(*[float(x) for x in set_to])
        _check(lib().tsg_learned_calibration(self._p, get, st))
        return np.array(list(get), np.float32)

    def learned_features(self, A, B):
        X = np.zeros((A.s.tilem, 8), np.float32)
        _check(lib().tsg_learned_features(self._p, C.byref(A.s), C.byref(B.s), X.ctypes.data_as(C.c_void_p)))
        return X

    def learned_predict(self, A, B):
        """Per tile-row (predicted tiles, predicted nonzeros)."""
        Y = np.zeros((A.s.tilem, 2), np.float32)
        _check(lib().tsg_learned_predict(self._p, C.byref(A.s), C.byref(B.s), Y.ctypes.data_as(C.c_void_p)))
        return Y

    def learned_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 2), np.float32)
        _check(lib().tsg_learned_forward(self._p, X.ctypes.data_as(C.c_void_p), C.c_int64(X.shape[0]),
                                         Y.ctypes.data_as(C.c_void_p)))
        return Y

    def learned_gradient(self, X, T):
        """(gradient of the mean squared error over the batch, loss)."""
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.learned_num_params(), np.float32)
        loss = C.c_float(0)
        _check(lib().tsg_learned_gradient(self._p, X.ctypes.data_as(C.c_void_p), T.ctypes.data_as(C.c_void_p),
                                          C.c_int64(X.shape[0]), g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return g, loss.value


class Plan:
    """tsg_plan: outputs[i] = As[i]*Bs[i]; execute() recomputes their values."""

    def __init__(self, ctx, As, Bs, opts):
        self.ctx = ctx
        single = not isinstance(As, (list, tuple))
        As = [As] if single else list(As)
        Bs = [Bs] if single else list(Bs)
        n = len(As)
        a = (TileStruct * n)(*[x.s for x in As])
        b = (TileStruct * n)(*[x.s for x in Bs])
        c = (TileStruct * n)()
        self._p = C.c_void_p()
        _check(lib().tsg_plan_create(ctx.handle, C.c_int32(n), a, b, c, C.byref(opts), C.byref(self._p)))
        self.outputs = []
        for i in range(n):
            t = DeviceTiles(ctx)
            t.s = TileStruct.from_buffer_copy(c[i])
            self.outputs.append(t)
        self._keep = (As, Bs)

    def execute(self):
        _check(lib().tsg_plan_execute(self._p))

    def destroy(self):
        if self._p:
            _check(lib().tsg_plan_destroy(self._p))
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BatchGraph:
    """tsg_batch_graph: launch() queues one graph launch of the whole batch
    product; result() waits for it and returns ([C_i views], [stats_i]).  The
    views are valid until the next launch or destroy()."""

    def __init__(self, ctx, As, Bs, slack, opts):
        self.ctx = ctx
        n = len(As)
        assert len(Bs) == n
        self.n = n
        a = (TileStruct * n)(*[x.s for x in As])
        b = (TileStruct * n)(*[x.s for x in Bs])
        self._p = C.c_void_p()
        _check(lib().tsg_batch_graph_create(ctx.handle, C.c_int32(n), a, b, C.byref(opts), C.c_double(slack),
                                            C.byref(self._p)))
        self._keep = (As, Bs)

    def launch(self):
        _check(lib().tsg_batch_graph_launch(self._p))

    def result(self):
        c = (TileStruct * self.n)()
        st = (Stats * self.n)()
        _check(lib().tsg_batch_graph_result(self._p, c, st))
        out = []
        for i in range(self.n):
            t = TileView(self.ctx)
            t.s = TileStruct.from_buffer_copy(c[i])
            out.append(t)
        return out, [st[i].as_dict() for i in range(self.n)]

    def capacity(self):
        t, z = C.c_int64(), C.c_int64()
        _check(lib().tsg_batch_graph_capacity(self._p, C.byref(t), C.byref(z)))
        return t.value, z.value

    def destroy(self):
        if self._p:
            lib().tsg_batch_graph_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceCSR:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self._host = h
        self.s = CSRStruct()
        hs = h.struct()
        _check(lib().tsg_csr_upload(ctx.handle, C.byref(hs), C.byref(self.s)))

    def free(self):
        if self.s.rowptr:
            _check(lib().tsg_csr_free(self.ctx.handle, C.byref(self.s)))

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass


def _host_tiles_to_dict(h):
    """A library-allocated host TileStruct -> dict of numpy arrays (frees h)."""
    nt, nnz = h.numtile, h.nnz
    rows = h.tile_size or 16     # per-tile row entries of tile_csr_Ptr and mask
    vdt = np.float32 if h.dtype == F32 else np.float64
    d = dict(m=h.m, n=h.n, tilem=h.tilem, tilen=h.tilen, nnz=nnz, numtile=nt, is_csc=h.is_csc,
             dtype=h.dtype, tile_size=rows,
             tile_ptr=_np(h.tile_ptr, h.tilem + 1, np.int64),
             tile_columnidx=_np(h.tile_columnidx, nt, np.int32),
             tile_rowidx=_np(h.tile_rowidx, nt, np.int32) if h.tile_rowidx else None,
             tile_nnz=_np(h.tile_nnz, nt + 1, np.int64),
             tile_csr_Val=_np(h.tile_csr_Val, nnz, vdt),
             tile_csr_Col=_np(h.tile_csr_Col, nnz, np.uint8),
             tile_csr_Ptr=_np(h.tile_csr_Ptr, nt * rows, np.uint8),
             mask=_np(h.mask, nt * rows, np.uint16))
    if h.is_csc:
        d["csc_tile_ptr"] = _np(h.csc_tile_ptr, h.tilen + 1, np.int64)
        d["csc_tile_rowidx"] = _np(h.csc_tile_rowidx, nt, np.int32)
        if h.csr_tile_nnz:      # the CSR-order, column-major copy for the staged kernels
            d["csr_tile_nnz"] = _np(h.csr_tile_nnz, nt + 1, np.int64)
            d["csr_mask"] = _np(h.csr_mask, nt * 16, np.uint16)
            d["csr_col_mask"] = _np(h.csr_col_mask, nt * 16, np.uint16)
            d["csr_col_Ptr"] = _np(h.csr_col_Ptr, nt * 16, np.uint8)
            d["csr_col_Val"] = _np(h.csr_col_Val, nnz, vdt)
            d["row_nnz_ptr"] = _np(h.row_nnz_ptr, h.tilem + 1, np.int64)
    if h.tile_csc_Ptr:          # within-tile CSC (tsg_build_tile_csc)
        d["tile_csc_Ptr"] = _np(h.tile_csc_Ptr, nt * 16, np.uint8)
        d["tile_csc_Row"] = _np(h.tile_csc_Row, nnz, np.uint8)
        d["tile_csc_Val"] = _np(h.tile_csc_Val, nnz, vdt)
    _check(lib().tsg_tile_free_host(C.byref(h)))
    return d


class DeviceTiles:
    """A tile matrix in device memory (owned by a Context's pool)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.s = TileStruct()

    def __getattr__(self, k):
        if k in ("m", "n", "tilem", "tilen", "nnz", "numtile", "dtype", "is_csc"):
            return getattr(self.s, k)
        raise AttributeError(k)

    def build_csc_tiles(self):
        """build_csc_tiles (src/tile_format.h:424-546) on the GPU, in place."""
        _check(lib().tsg_build_csc_tiles(self.ctx.handle, C.byref(self.s)))
        return self

    def build_tile_csc(self):
        """Within-tile CSC of every tile on the GPU, in place: the reference's
        tile_csc_* of a B operand (src/tile_format.h:547-584) or tile_a_csc_* of
        row tiles (build_a_tile_csc, src/tile_format.h:589-627)."""
        _check(lib().tsg_build_tile_csc(self.ctx.handle, C.byref(self.s)))
        return self

    def download(self):
        """Returns a dict of numpy arrays in the reference's field names."""
        h = TileStruct()
        _check(lib().tsg_tile_download(self.ctx.handle, C.byref(self.s), C.byref(h)))
        return _host_tiles_to_dict(h)

    def download_rows(self, tr_lo, tr_hi):
        """Tile-rows [tr_lo, tr_hi) as a dict (download() layout) of a matrix of
        tr_hi - tr_lo tile-rows: tile_ptr, tile_nnz, tile_rowidx rebased."""
        h = TileStruct()
        _check(lib().tsg_tile_download_rows(self.ctx.handle, C.byref(self.s), C.c_int32(tr_lo), C.c_int32(tr_hi),
                                            C.byref(h)))
        return _host_tiles_to_dict(h)

    def to_csr(self):
        """Row tiles -> host CSR through the GPU tile2csr kernels."""
        o, h = CSRStruct(), CSRStruct()
        _check(lib().tsg_tile2csr(self.ctx.handle, C.byref(self.s), C.byref(o)))
        try:
            _check(lib().tsg_csr_download(self.ctx.handle, C.byref(o), C.byref(h)))
        finally:
            lib().tsg_csr_free(self.ctx.handle, C.byref(o))
        return HostCSR._take(h)

    def free(self):
        if self.s.tile_ptr:
            _check(lib().tsg_tile_free(self.ctx.handle, C.byref(self.s)))

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass


class TileView(DeviceTiles):
    """A tile matrix whose arrays another object owns (a batch graph's outputs)."""

    def free(self):
        pass


def upload_tiles(ctx, d):
    """Uploads a host tile dict (download() layout) -- upload_tiles, src/spgemm_raw_v2.cu:25-48."""
    keep = {}

    def ptr(name, ctype):
        a = d.get(name)
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        keep[name] = a
        return a.ctypes.data_as(C.POINTER(ctype))

    h = TileStruct()
    h.m, h.n, h.tilem, h.tilen = d["m"], d["n"], d["tilem"], d["tilen"]
    h.nnz, h.numtile, h.dtype, h.location, h.is_csc = d["nnz"], d["numtile"], d["dtype"], HOST, d["is_csc"]
    h.tile_size = 0 if d.get("tile_size", 16) == 16 else int(d["tile_size"])
    h.tile_ptr = ptr("tile_ptr", C.c_int64)
    h.tile_columnidx = ptr("tile_columnidx", C.c_int32)
    h.tile_rowidx = ptr("tile_rowidx", C.c_int32)
    h.tile_nnz = ptr("tile_nnz", C.c_int64)
    v = np.ascontiguousarray(d["tile_csr_Val"])
    keep["v"] = v
    h.tile_csr_Val = v.ctypes.data_as(C.c_void_p)
    h.tile_csr_Col = ptr("tile_csr_Col", C.c_uint8)
    h.tile_csr_Ptr = ptr("tile_csr_Ptr", C.c_uint8)
    h.mask = ptr("mask", C.c_uint16)
    h.csc_tile_ptr = ptr("csc_tile_ptr", C.c_int64)
    h.csc_tile_rowidx = ptr("csc_tile_rowidx", C.c_int32)
    if d.get("tile_csc_Ptr") is not None:
        h.tile_csc_Ptr = ptr("tile_csc_Ptr", C.c_uint8)
        h.tile_csc_Row = ptr("tile_csc_Row", C.c_uint8)
        cv = np.ascontiguousarray(d["tile_csc_Val"])
        keep["cv"] = cv
        h.tile_csc_Val = cv.ctypes.data_as(C.c_void_p)
    t = DeviceTiles(ctx)
    _check(lib().tsg_tile_upload(ctx.handle, C.byref(h), C.byref(t.s)))
    return t
