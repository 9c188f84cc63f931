"""Batch graphs (tsg_batch_graph_*): the whole batched product as one captured
hipGraph.  Every C_i bit-identical to tsg_spgemm_batched (itself checked
against single calls and the oracle), on the first launch and on replays;
replays follow new operand contents (values and structure); a replay past a
capacity is flagged on the device and answered by the exact eager batch."""
import ctypes as C

import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O
from tests.conftest import assert_tiles_equal
from tests.test_gpu_batch import CKEYS, problems

pytestmark = pytest.mark.gpu


def _addr(p):
    return C.cast(p, C.c_void_p).value


def _array_sizes(s):
    """Bytes of every array of an A operand made by tiles_from_csr."""
    vs = 4 if s.dtype == g.F32 else 8
    return {"tile_ptr": (s.tilem + 1) * 8, "tile_columnidx": s.numtile * 4, "tile_rowidx": s.numtile * 4,
            "tile_nnz": (s.numtile + 1) * 8, "tile_csr_Ptr": s.numtile * 16, "mask": s.numtile * 32,
            "tile_csr_Col": s.nnz, "tile_csr_Val": s.nnz * vs, "row_ptr": (s.m + 1) * 8, "row_col": s.nnz * 4,
            "row_val": s.nnz * vs}


def overwrite(ctx, dst, src):
    """Copies src's arrays into dst's (same dimensions, tile count, nonzeros)."""
    for k in ("m", "n", "tilem", "tilen", "nnz", "numtile", "dtype"):
        assert getattr(dst.s, k) == getattr(src.s, k), k
    for k, nb in _array_sizes(src.s).items():
        d, s = _addr(getattr(dst.s, k)), _addr(getattr(src.s, k))
        assert (d is None) == (s is None), k
        if d and nb:
            ctx.copy(d, s, nb)


def tile_permuted(A, perm_rows=None, perm_cols=None):
    """A with whole tile-rows and/or tile columns permuted (same tile count and
    nonzeros as A)."""
    S = A.to_scipy().tocsr()
    m, n = S.shape
    r = np.arange(m)
    c = np.arange(n)
    if perm_rows is not None:
        r = perm_rows[r >> 4] * 16 + (r & 15)
    if perm_cols is not None:
        c = perm_cols[c >> 4] * 16 + (c & 15)
    coo = S.tocoo()
    import scipy.sparse as sp
    T = sp.csr_matrix((coo.data, (r[coo.row], c[coo.col])), shape=(m, n))
    T.sort_indices()
    return g.HostCSR(m, n, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data)


@pytest.mark.parametrize("f32", [False, True])
def test_graph_equals_batched(ctx, f32):
    ps = problems(f32)
    print(f"[case] batch graph {len(ps)} problems f32={f32}", flush=True)
    tA = [ctx.tiles_from_csr(a) for a, _ in ps]
    tB = [ctx.tiles_from_csr(b, as_b=True) for _, b in ps]
    ref, rst = ctx.spgemm_batched(tA, tB)
    exp = [c.download() for c in ref]
    G = ctx.batch_graph(tA, tB, slack=1.25)
    try:
        cap_t, cap_z = G.capacity()
        assert cap_t >= sum(s["numtileC"] for s in rst) and cap_z >= sum(s["nnzC"] for s in rst)
        for rep in range(3):   # first launch and replays
            G.launch()
            Cs, sts = G.result()
            for i in range(len(ps)):
                assert sts[i]["fallback_used"] == 0 and sts[i]["host_syncs"] == 1
                assert sts[i]["numtileC"] == rst[i]["numtileC"] and sts[i]["nnzC"] == rst[i]["nnzC"]
                assert_tiles_equal(Cs[i].download(), exp[i], keys=CKEYS, label=f"rep {rep} problem {i} ")
        a, b = ps[0]
        want = O.spgemm_tiled(a.astype(g.F64) if f32 else a, b.astype(g.F64) if f32 else b, f32=f32)
        assert_tiles_equal(Cs[0].download(), want, keys=CKEYS, label="oracle 0 ")
    finally:
        G.destroy()


@pytest.mark.parametrize("f32", [False, True])
def test_graph_row_lane_bin(ctx, f32):
    """2-D Laplacians fill bin R, whose symbolic and numeric kernels run on
    device-side row ranges inside the graph: bit-identical to the eager batch
    and the oracle, on a launch and a replay."""
    ps = [g.gen_laplace2d(96, 80), g.gen_laplace2d(200, 33), g.gen_laplace2d(17, 300)]
    if f32:
        ps = [a.astype(g.F32) for a in ps]
    print(f"[case] batch graph row-lane bin f32={f32}", flush=True)
    tA = [ctx.tiles_from_csr(a) for a in ps]
    tB = [ctx.tiles_from_csr(a, as_b=True) for a in ps]
    ref, _ = ctx.spgemm_batched(tA, tB)
    exp = [c.download() for c in ref]
    G = ctx.batch_graph(tA, tB, slack=1.25)
    try:
        for rep in range(2):
            G.launch()
            Cs, sts = G.result()
            for i, a in enumerate(ps):
                assert sts[i]["fallback_used"] == 0
                got = Cs[i].download()
                assert_tiles_equal(got, exp[i], keys=CKEYS, label=f"rep {rep} problem {i} ")
        want = O.spgemm_tiled(ps[0].astype(g.F64) if f32 else ps[0], ps[0].astype(g.F64) if f32 else ps[0], f32=f32)
        assert_tiles_equal(Cs[0].download(), want, keys=CKEYS, label="oracle 0 ")
    finally:
        G.destroy()


def test_graph_replay_follows_new_structure(ctx):
    """Operands of equal sizes swapped between launches: the replay recomputes
    Steps 1-3 and the outputs swap with them."""
    X = g.gen_er(704, 9, seed=11)
    perm = np.random.default_rng(3).permutation((X.m + 15) // 16)
    Y = tile_permuted(X, perm_rows=perm)
    B = g.gen_er(704, 5, seed=12)
    tX, tY = ctx.tiles_from_csr(X), ctx.tiles_from_csr(Y)
    tA = [ctx.tiles_from_csr(X), ctx.tiles_from_csr(Y)]
    tB = [ctx.tiles_from_csr(B, as_b=True) for _ in range(2)]
    expXB = O.spgemm_tiled(X, B)
    expYB = O.spgemm_tiled(Y, B)
    G = ctx.batch_graph(tA, tB, slack=1.25)
    try:
        G.launch()
        Cs, _ = G.result()
        assert_tiles_equal(Cs[0].download(), expXB, keys=CKEYS, label="XB ")
        assert_tiles_equal(Cs[1].download(), expYB, keys=CKEYS, label="YB ")
        overwrite(ctx, tA[0], tY)
        overwrite(ctx, tA[1], tX)
        G.launch()
        Cs, sts = G.result()
        assert sts[0]["fallback_used"] == 0
        assert_tiles_equal(Cs[0].download(), expYB, keys=CKEYS, label="swapped YB ")
        assert_tiles_equal(Cs[1].download(), expXB, keys=CKEYS, label="swapped XB ")
    finally:
        G.destroy()


def test_graph_capacity_short_falls_back_exactly(ctx):
    """A replay whose C outgrows the capacities (slack 1): flagged on the
    device, computed by the eager batch, equal to the oracle."""
    n = 1024
    rp = np.arange(0, 3 * n + 1, 3, dtype=np.int64)[: n + 1]
    rows = np.repeat(np.arange(n), 3)
    cols = np.clip(rows + np.tile([-1, 0, 1], n), 0, n - 1)
    cols[0], cols[-1] = 2, n - 3  # keep three distinct sorted columns at the ends
    X = g.HostCSR(n, n, rp, cols.astype(np.int32), np.linspace(1.0, 2.0, 3 * n))
    X = tile_permuted(X)  # sorted
    perm = np.random.default_rng(5).permutation(n // 16)
    Y = tile_permuted(X, perm_cols=perm)  # same tiles and nonzeros per tile-row, scattered columns
    tA = [ctx.tiles_from_csr(X)]
    tB = [ctx.tiles_from_csr(X, as_b=True)]
    tY = ctx.tiles_from_csr(Y)
    assert tY.s.numtile == tA[0].s.numtile and tY.s.nnz == tA[0].s.nnz
    G = ctx.batch_graph(tA, tB, slack=1.0)
    try:
        G.launch()
        Cs, sts = G.result()
        assert sts[0]["fallback_used"] == 0
        assert_tiles_equal(Cs[0].download(), O.spgemm_tiled(X, X), keys=CKEYS, label="first ")
        overwrite(ctx, tA[0], tY)
        G.launch()
        Cs, sts = G.result()
        exp = O.spgemm_tiled(Y, X)
        cap_t, cap_z = G.capacity()
        assert exp["numtile"] > cap_t or exp["nnz"] > cap_z, "the test must outgrow a capacity"
        assert sts[0]["fallback_used"] == 1
        assert_tiles_equal(Cs[0].download(), exp, keys=CKEYS, label="fallback ")
    finally:
        G.destroy()


def test_graph_rejects_other_variants(ctx):
    A = g.gen_er(100, 4, seed=1)
    tA, tB = [ctx.tiles_from_csr(A)], [ctx.tiles_from_csr(A, as_b=True)]
    o = g.Options(g.SCHED_PLAIN, 1, 0, 0, 1.0, 0)
    with pytest.raises(g.TsgError):
        g.BatchGraph(ctx, tA, tB, 1.25, o)
