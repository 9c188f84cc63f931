"""GPU tile-format conversion (tsg_csr2tile, tsg_build_csc_tiles, tsg_tile2csr,
tsg_build_tile_csc)
against the reference's own conversion output (golden fixtures written by
src/tile_format.h) and against the oracle, bit for bit."""
import os

import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O
from tests.conftest import assert_csc_equal, assert_tiles_equal

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "tile_format_golden.npz")
CASES = ["er200", "lap20x13", "rect77x300", "dense40", "sparse50"]


def gold(name):
    z = np.load(GOLD, allow_pickle=False)
    A = g.HostCSR(int(z[f"{name}/m"]), int(z[f"{name}/n"]), z[f"{name}/rowptr"], z[f"{name}/col"],
                  z[f"{name}/val"])
    exp = {}
    for tag in ("A", "B"):
        exp[tag] = {k.split("/")[-1]: (z[k] if z[k].ndim else z[k].item())
                    for k in z.files if k.startswith(f"{name}/{tag}/")}
    return A, exp


@pytest.mark.parametrize("name", CASES)
def test_gpu_csr2tile_matches_reference_golden(ctx, name):
    print(f"[case] csr2tile golden {name}", flush=True)
    A, exp = gold(name)
    assert_tiles_equal(ctx.tiles_from_csr(A).download(), exp["A"], label="A ")
    assert_tiles_equal(ctx.tiles_from_csr(A, as_b=True).download(), exp["B"], label="B ")


def _matrices():
    return {
        "er4096": lambda: g.gen_er(4096, 8, seed=0),
        "lap2d_64x48": lambda: g.gen_laplace2d(64, 48),
        "stencil27_9": lambda: g.gen_stencil27(9, 9, 9),
        "aggregation": lambda: g.gen_aggregation(10, 6, 4),
        "rmat10": lambda: g.gen_rmat(10, 16, seed=1, dtype=g.F64),
        "er_var": lambda: g.gen_er_var(1000, 0, 40, seed=2),
    }


@pytest.mark.parametrize("name", list(_matrices()))
@pytest.mark.parametrize("f32", [False, True])
def test_gpu_conversion_matches_oracle(ctx, name, f32):
    print(f"[case] conversion {name} f32={f32}", flush=True)
    A = _matrices()[name]()
    h = A.astype(g.F32) if f32 else A
    assert_tiles_equal(ctx.tiles_from_csr(h).download(), O.csr2tile(A, f32=f32), label="A ")
    assert_tiles_equal(ctx.tiles_from_csr(h, as_b=True).download(), O.csr2tile_b(A, f32=f32), label="B ")


@pytest.mark.parametrize("name", CASES)
def test_gpu_tile_csc_matches_reference_golden(ctx, name):
    """tsg_build_tile_csc against the reference's own within-tile CSC
    (build_csc_tiles :547-584 for B, build_a_tile_csc :589-627 for A)."""
    print(f"[case] tile csc golden {name}", flush=True)
    A, _ = gold(name)
    z = np.load(GOLD, allow_pickle=False)
    for tag, as_b in (("A", False), ("B", True)):
        exp = {k.split("/")[-1]: z[k] for k in z.files if k.startswith(f"{name}/{tag}csc/")}
        assert_csc_equal(ctx.tiles_from_csr(A, as_b=as_b).build_tile_csc().download(), exp, label=f"{tag} ")


@pytest.mark.parametrize("name", list(_matrices()))
@pytest.mark.parametrize("f32", [False, True])
def test_gpu_tile_csc_matches_oracle(ctx, name, f32):
    print(f"[case] tile csc {name} f32={f32}", flush=True)
    A = _matrices()[name]()
    h = A.astype(g.F32) if f32 else A
    assert_csc_equal(ctx.tiles_from_csr(h).build_tile_csc().download(), O.tile_csc(O.csr2tile(A, f32=f32)),
                     label="A ")
    assert_csc_equal(ctx.tiles_from_csr(h, as_b=True).build_tile_csc().download(),
                     O.tile_csc(O.csr2tile_b(A, f32=f32)), label="B ")


def test_gpu_tile_csc_duplicates_refresh_and_relayout(ctx):
    """Repeated columns keep storage order; tsg_tile_set_values refreshes the
    CSC values; re-laying row tiles as a B operand drops the row-order CSC."""
    print("[case] tile csc duplicates / set_values / relayout", flush=True)
    rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9, 12], dtype=np.int64)
    col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6, 0, 0, 0], dtype=np.int32)
    A = g.HostCSR(9, 8, rp, col, np.arange(1.0, 13.0))
    for as_b, conv in ((False, O.csr2tile), (True, O.csr2tile_b)):
        t = ctx.tiles_from_csr(A, as_b=as_b)
        assert t.s.flags & 1
        assert_csc_equal(t.build_tile_csc().download(), O.tile_csc(conv(A)), label=f"dups b={as_b} ")
    B = g.gen_er_var(500, 0, 30, seed=4)
    t = ctx.tiles_from_csr(B).build_tile_csc()
    B2 = g.HostCSR(B.m, B.n, B.rowptr, B.col, np.asarray(B.val) * -3.0 + 0.5)
    ctx.set_values(t, B2)
    assert_csc_equal(t.download(), O.tile_csc(O.csr2tile(B2)), label="refreshed ")
    t.build_csc_tiles()
    assert "tile_csc_Ptr" not in t.download()
    assert_csc_equal(t.build_tile_csc().download(), O.tile_csc(O.csr2tile_b(B2)), label="relaid ")
    Z = g.HostCSR(20, 20, np.zeros(21, np.int64), np.zeros(0, np.int32), np.zeros(0))
    z = ctx.tiles_from_csr(Z).build_tile_csc().download()
    assert len(z["tile_csc_Ptr"]) == 0 and len(z["tile_csc_Row"]) == 0


TS_SIZES = (1, 5, 8, 13)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("ts", TS_SIZES)
def test_gpu_csr2tile_ts_matches_reference_golden(ctx, name, ts):
    """tsg_csr2tile_ts (+ tsg_build_csc_tiles on it) against the reference's
    csr2tile_ts / build_csc_tiles_ts output."""
    print(f"[case] csr2tile_ts golden {name} ts={ts}", flush=True)
    A, _ = gold(name)
    z = np.load(GOLD, allow_pickle=False)
    for tag, as_b in (("A", False), ("B", True)):
        exp = {k.split("/")[-1]: (z[k] if z[k].ndim else z[k].item())
               for k in z.files if k.startswith(f"{name}/{tag}ts{ts}/")}
        got = ctx.tiles_from_csr(A, as_b=as_b, ts=ts).download()
        assert got["tile_size"] == ts
        assert_tiles_equal(got, exp, label=f"{tag} ts{ts} ")


@pytest.mark.parametrize("name", ["er4096", "lap2d_64x48", "rmat10", "er_var"])
@pytest.mark.parametrize("ts", [2, 3, 7, 12, 15, 16])
def test_gpu_csr2tile_ts_matches_oracle(ctx, name, ts):
    print(f"[case] csr2tile_ts {name} ts={ts}", flush=True)
    A = _matrices()[name]()
    f32 = ts % 2 == 1
    h = A.astype(g.F32) if f32 else A
    assert_tiles_equal(ctx.tiles_from_csr(h, ts=ts).download(), O.csr2tile_ts(A, ts, f32=f32), label="A ")
    assert_tiles_equal(ctx.tiles_from_csr(h, as_b=True, ts=ts).download(), O.csr2tile_ts(A, ts, as_b=True, f32=f32),
                       label="B ")


def test_gpu_tile_size_limits(ctx):
    """Duplicates keep storage order at every tile size; products and the
    16-row-only conversions reject other tile sizes; ts outside 1-16 is an error."""
    print("[case] tile size limits", flush=True)
    rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9, 12], dtype=np.int64)
    col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6, 0, 0, 0], dtype=np.int32)
    D = g.HostCSR(9, 8, rp, col, np.arange(1.0, 13.0))
    for ts in (3, 4, 9):
        t = ctx.tiles_from_csr(D, ts=ts)
        assert t.s.flags & 1
        assert_tiles_equal(t.download(), O.csr2tile_ts(D, ts), label=f"dups ts{ts} A ")
        assert_tiles_equal(ctx.tiles_from_csr(D, as_b=True, ts=ts).download(), O.csr2tile_ts(D, ts, as_b=True),
                           label=f"dups ts{ts} B ")
    A = g.gen_er(200, 5, seed=3)
    t8, b8 = ctx.tiles_from_csr(A, ts=8), ctx.tiles_from_csr(A, as_b=True, ts=8)
    with pytest.raises(g.TsgError, match="16 x 16"):
        ctx.spgemm(t8, b8)
    with pytest.raises(g.TsgError, match="16 x 16"):
        t8.build_tile_csc()
    with pytest.raises(g.TsgError, match="16 x 16"):
        t8.to_csr()
    for bad in (0, 17, -1):
        with pytest.raises(g.TsgError, match="tile size"):
            ctx.tiles_from_csr(A, ts=bad)
    # host round trip keeps the tile size
    d = t8.download()
    back = g.upload_tiles(ctx, d).download()
    assert back["tile_size"] == 8
    assert_tiles_equal(back, d, label="round trip ")


@pytest.mark.parametrize("name", ["er4096", "rmat10", "aggregation"])
def test_gpu_build_bitmasks_matches_oracle(ctx, name):
    """tsg_build_bitmasks (build_bitmasks, src/tile_format.h:628-645)."""
    print(f"[case] bitmasks {name}", flush=True)
    A = _matrices()[name]()
    B = A if A.m == A.n else A.transpose()
    bmA, bmB = ctx.build_bitmasks(ctx.tiles_from_csr(A), ctx.tiles_from_csr(B, as_b=True))
    eA, eB = O.build_bitmasks(O.csr2tile(A), O.csr2tile_b(B))
    assert np.array_equal(bmA, eA) and np.array_equal(bmB, eB)


def test_gpu_long_tile_row_flag(ctx):
    """flags bit 1: some tile-row holds more than 64 tiles (products then keep
    column bitmaps of such B tile-rows for the heavy rows' Step 1)."""
    print("[case] long tile-row flag", flush=True)
    E = g.gen_er(4096, 40, seed=2)                 # 640 entries a tile-row over 256 tile columns
    d = O.csr2tile(E)
    assert int(np.diff(d["tile_ptr"]).max()) > 64
    assert ctx.tiles_from_csr(E).s.flags & 2 and ctx.tiles_from_csr(E, as_b=True).s.flags & 2
    L = g.gen_laplace2d(64, 48)                    # about 5 tiles a tile-row
    assert not ctx.tiles_from_csr(L).s.flags & 2


def test_gpu_conversion_wide_rows_take_several_windows(ctx):
    # tile columns span more than one LDS window (32,768 tile columns per pass)
    n = 16 * 70000 + 5
    rows = np.repeat(np.arange(40), 3)
    cols = np.array([0, 16 * 33000 + 3, n - 1] * 40)
    rp = np.arange(0, 121, 3, dtype=np.int64)
    A = g.HostCSR(40, n, rp, cols, np.linspace(-1, 1, 120))
    print("[case] wide rows", flush=True)
    assert_tiles_equal(ctx.tiles_from_csr(A).download(), O.csr2tile(A), label="A ")
    assert_tiles_equal(ctx.tiles_from_csr(A, as_b=True).download(), O.csr2tile_b(A), label="B ")


def test_gpu_conversion_empty(ctx):
    A = g.HostCSR(33, 20, np.zeros(34, np.int64), np.zeros(0, np.int32), np.zeros(0))
    d = ctx.tiles_from_csr(A).download()
    assert d["numtile"] == 0 and d["nnz"] == 0 and np.all(d["tile_ptr"] == 0)
    b = ctx.tiles_from_csr(A, as_b=True).download()
    assert b["numtile"] == 0 and np.all(b["csc_tile_ptr"] == 0)


def test_gpu_conversion_rejects_bad_columns(ctx):
    A = g.HostCSR(2, 10, [0, 2, 3], [3, 12, 7], [1.0, 2.0, 3.0])
    with pytest.raises(g.TsgError) as e:
        ctx.tiles_from_csr(A)
    assert e.value.code == -1
    A = g.HostCSR(2, 10, [0, 2, 3], [-1, 2, 7], [1.0, 2.0, 3.0])
    with pytest.raises(g.TsgError):
        ctx.tiles_from_csr(A)


def test_gpu_conversion_rejects_unsorted_rows(ctx):
    A = g.HostCSR(2, 40, [0, 2, 3], [30, 5, 7], [1.0, 2.0, 3.0])
    with pytest.raises(g.TsgError) as e:
        ctx.tiles_from_csr(A)
    assert e.value.code == -5


@pytest.mark.parametrize("name", ["er4096", "rmat10", "aggregation"])
def test_gpu_tile2csr_roundtrip(ctx, name):
    A = _matrices()[name]()
    t = ctx.tiles_from_csr(A)
    B = t.to_csr()
    assert np.array_equal(B.rowptr, A.rowptr) and np.array_equal(B.col, A.col)
    assert np.array_equal(B.val, A.val)


def colmajor_expected(t):
    """Column-major copy of each row-ordered tile from the oracle's csr2tile
    output: entries re-sorted by (tile, column, row)."""
    nt, nnz = t["numtile"], t["nnz"]
    tn, ptr = np.asarray(t["tile_nnz"], np.int64), np.asarray(t["tile_csr_Ptr"], np.int64).reshape(nt, 16)
    tile = np.repeat(np.arange(nt), np.diff(tn))
    local = np.arange(nnz) - tn[tile]
    # row of each entry: last r with ptr[tile, r] <= local
    row = (ptr[tile] <= local[:, None]).sum(1) - 1
    col = np.asarray(t["tile_csr_Col"], np.int64) & 15
    order = np.lexsort((row, col, tile))
    cv = np.asarray(t["tile_csr_Val"])[order]
    cm = np.zeros(nt * 16, np.uint16)
    np.bitwise_or.at(cm, tile * 16 + col, (0x8000 >> row).astype(np.uint16))
    cnt = np.bincount(tile * 16 + col, minlength=nt * 16).reshape(nt, 16)
    cp = (np.cumsum(cnt, 1) - cnt).astype(np.uint8).ravel()
    return cm, cp, cv


@pytest.mark.parametrize("name", ["er4096", "lap2d_64x48", "rmat10"])
def test_gpu_b_column_major_copy(ctx, name):
    A = _matrices()[name]()
    print(f"[case] column-major B {name}", flush=True)
    b = ctx.tiles_from_csr(A, as_b=True).download()
    t = O.csr2tile(A)
    assert np.array_equal(b["csr_tile_nnz"], t["tile_nnz"]) and np.array_equal(b["csr_mask"], t["mask"])
    cm, cp, cv = colmajor_expected(t)
    assert np.array_equal(b["csr_col_mask"], cm)
    assert np.array_equal(b["csr_col_Ptr"], cp)
    assert np.array_equal(b["csr_col_Val"], cv)
    rp = np.asarray(A.rowptr)
    assert np.array_equal(b["row_nnz_ptr"], rp[np.minimum(np.arange(t["tilem"] + 1) * 16, A.m)])
