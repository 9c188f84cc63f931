"""GPU SpGEMM (tsg_spgemm, plain scheduler) against the oracle.

Structure (tile_ptr, tile_columnidx, tile_rowidx, tile_nnz, tile_csr_Ptr, mask,
tile_csr_Col) must be bit-exact.  Values are compared two ways:
  * bit-exact against the oracle's tiled restatement, which keeps the same
    summation order (pairs ascending k, one FMA per product);
  * against the FP64 Gustavson recompute of the reference's checker
    (src/spgemm_raw_v2.cu:253-359): 0 missing, 0 extra, and
    max |c - c_ref| / sum_k |a_ik b_kj| <= 1e-12 (FP64) or <= 1e-5 (FP32);
    the measured value is printed.
"""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
TOL = {g.F64: 1e-12, g.F32: 1e-5}
# 0: tile-rows binned to the row-staged kernels (heavy rows to the C-tile
# kernels); 1: the reference's C-tile method for every tile; 3: the C-tile
# method in outer-product form over A's within-tile CSC
VARIANTS = [0, 1, 3]
CKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
         "tile_csr_Val"]


def run(ctx, A, B, f32=False, label="", **kw):
    print(f"[case] spgemm {label} A={A.m}x{A.n} nnz={A.nnz} B={B.m}x{B.n} f32={f32} {kw}", flush=True)
    ha, hb = (A.astype(g.F32), B.astype(g.F32)) if f32 else (A, B)
    tA = ctx.tiles_from_csr(ha)
    if kw.get("variant") == 3:
        tA.build_tile_csc()
    tB = ctx.tiles_from_csr(hb, as_b=True)
    Cm, st = ctx.spgemm(tA, tB, **kw)
    got = Cm.download()
    return got, st


def check(got, A, B, f32=False, exact_values=True):
    exp = O.spgemm_tiled(A, B, f32=f32)
    keys = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col"]
    if exact_values:
        keys.append("tile_csr_Val")
    assert_tiles_equal(got, exp, keys=keys, label="C ")
    chk = O.gustavson_check(A, B, got)
    print(f"    gustavson: missing={chk['missing']} extra={chk['extra']} rel_frob={chk['rel_frob']:.3e} "
          f"max_scaled={chk['max_scaled']:.3e}", flush=True)
    assert chk["missing"] == 0 and chk["extra"] == 0
    assert chk["max_scaled"] <= TOL[g.F32 if f32 else g.F64]
    return chk


@pytest.mark.parametrize("variant", VARIANTS)
def test_er4096_fp64_plain(ctx, variant):
    """The CPU-path configuration: ER 4096, exactly 8 nnz/row, U[-1,1], seed 0."""
    A = g.gen_er(4096, 8, seed=0)
    got, st = run(ctx, A, A, label="er4096", variant=variant)
    check(got, A, A)
    assert st["numtileC"] == got["numtile"] and st["nnzC"] == got["nnz"]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("f32", [False, True])
def test_random_rectangular(ctx, f32, variant):
    rng = np.random.default_rng(5)
    import scipy.sparse as sp
    for (m, k, n, d) in [(300, 500, 200, 0.02), (64, 64, 64, 0.3), (1000, 37, 900, 0.05)]:
        SA = sp.random(m, k, density=d, format="csr", random_state=rng)
        SB = sp.random(k, n, density=d, format="csr", random_state=rng)
        SA.data = rng.uniform(-1, 1, SA.nnz)
        SB.data = rng.uniform(-1, 1, SB.nnz)
        SA.sort_indices(); SB.sort_indices()
        A = g.HostCSR(m, k, SA.indptr, SA.indices, SA.data)
        B = g.HostCSR(k, n, SB.indptr, SB.indices, SB.data)
        got, _ = run(ctx, A, B, f32=f32, label=f"rand{m}x{k}x{n}", variant=variant)
        check(got, A, B, f32=f32)


@pytest.mark.parametrize("variant", VARIANTS)
def test_aat(ctx, variant):
    """--aat path: B = A^T (src/spgemm_raw_v2.cu:78-89)."""
    A = g.gen_er_var(700, 1, 30, seed=3)
    B = A.transpose()
    got, _ = run(ctx, A, B, label="aat", variant=variant)
    check(got, A, B)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("n", [48, 200])
def test_dense_tiles(ctx, variant, n):
    import scipy.sparse as sp
    D = sp.csr_matrix(np.random.default_rng(1).uniform(-1, 1, (n, n)))
    A = g.HostCSR(n, n, D.indptr, D.indices, D.data)
    got, _ = run(ctx, A, A, label=f"dense{n}", variant=variant)
    check(got, A, A)


@pytest.mark.parametrize("variant", VARIANTS)
def test_laplace2d_integer_exact(ctx, variant):
    A = g.gen_laplace2d(200, 150)
    got, _ = run(ctx, A, A, label="lap2d", variant=variant)
    chk = check(got, A, A)
    assert chk["max_abs"] == 0.0


@pytest.mark.parametrize("variant", [0, 1])
def test_laplace2d_2048_full(ctx, variant):
    """The bench workload at full size: 5-point Laplacian on a 2048 x 2048 grid."""
    A = g.gen_laplace2d(2048, 2048)
    got, st = run(ctx, A, A, label="lap2d_2048", variant=variant)
    assert got["numtile"] == 3386116 and got["nnz"] == 54484996
    chk = check(got, A, A)
    assert chk["max_abs"] == 0.0 and chk["rel_frob"] == 0.0


@pytest.mark.parametrize("variant", VARIANTS)
def test_rmat_fp32(ctx, variant):
    A = g.gen_rmat(12, 16, seed=1, dtype=g.F32)
    A64 = A.astype(g.F64)
    got, _ = run(ctx, A64, A64, f32=True, label="rmat12", variant=variant)
    check(got, A64, A64, f32=True)


@pytest.mark.parametrize("variant", VARIANTS)
def test_galerkin_chain_on_device(ctx, variant):
    """P^T (A P) with A a 27-point stencil: C1 = A P is re-laid as a B operand on
    the GPU (tsg_build_csc_tiles on a product) and fed to the second product."""
    nx = 12
    A = g.gen_stencil27(nx, nx, nx)
    P = g.gen_aggregation(nx, nx, nx)
    Pt = P.transpose()
    tA = ctx.tiles_from_csr(A)
    tP = ctx.tiles_from_csr(P, as_b=True)
    if variant == 3:
        tA.build_tile_csc()
    C1, _ = ctx.spgemm(tA, tP, variant=variant)
    c1 = C1.download()
    check(c1, A, P)
    C1.build_csc_tiles()
    tPt = ctx.tiles_from_csr(Pt)
    if variant == 3:
        tPt.build_tile_csc()
    C2, _ = ctx.spgemm(tPt, C1, variant=variant)
    c2 = C2.download()
    AP = O.gustavson_csr(A, P)
    APc = g.HostCSR(A.m, P.n, *AP)
    if np.all(np.diff(c1["tile_nnz"]) > 0):
        # no structurally empty tiles in C1: its CSR re-tiles to the same operand
        check(c2, Pt, APc)
    chk = O.gustavson_check(Pt, APc, c2)
    assert chk["missing"] == 0 and chk["extra"] == 0 and chk["max_scaled"] <= 1e-12
    # integer stencil: coarse operator exact
    import scipy.sparse as sp
    ref = (Pt.to_scipy() @ (A.to_scipy() @ P.to_scipy())).toarray()
    dense = np.zeros_like(ref)
    for t in range(c2["numtile"]):
        tr, tc = c2["tile_rowidx"][t], c2["tile_columnidx"][t]
        for e in range(c2["tile_nnz"][t], c2["tile_nnz"][t + 1]):
            b = int(c2["tile_csr_Col"][e])
            dense[tr * 16 + (b >> 4), tc * 16 + (b & 15)] = c2["tile_csr_Val"][e]
    assert np.array_equal(dense, ref)
    assert sp.issparse(P.to_scipy())


def test_empty_and_tiny(ctx):
    Z = g.HostCSR(40, 40, np.zeros(41, np.int64), np.zeros(0, np.int32), np.zeros(0))
    got, st = run(ctx, Z, Z, label="empty")
    assert got["numtile"] == 0 and got["nnz"] == 0
    one = g.HostCSR(1, 1, [0, 1], [0], [3.0])
    got, _ = run(ctx, one, one, label="1x1")
    assert got["nnz"] == 1 and got["tile_csr_Val"][0] == 9.0


@pytest.mark.parametrize("variant", VARIANTS)
def test_duplicate_columns(ctx, variant):
    """load_mtx keeps duplicates (src/tile_format.h:78-89): both entries live in
    the tile, one mask bit; products use every entry."""
    rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9], dtype=np.int64)
    col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6], dtype=np.int32)
    A = g.HostCSR(8, 8, rp, col, np.arange(1.0, 10.0))
    tA = ctx.tiles_from_csr(A)
    assert tA.s.flags & 1
    got, _ = run(ctx, A, A, label="dups", variant=variant)
    exp = O.spgemm_tiled(A, A)
    assert_tiles_equal(got, exp, keys=["tile_ptr", "tile_columnidx", "tile_nnz", "mask", "tile_csr_Col",
                                       "tile_csr_Val"])


def test_outer_product_variant_needs_a_tile_csc(ctx):
    A = g.gen_er(300, 6, seed=2)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    with pytest.raises(g.TsgError, match="within-tile CSC"):
        ctx.spgemm(tA, tB, variant=3)
    tA.build_tile_csc()
    with pytest.raises(g.TsgError, match="plain scheduler"):
        ctx.spgemm(tA, tB, variant=3, scheduler=g.SCHED_LEARNED)
    with pytest.raises(g.TsgError, match="numeric_variant"):
        ctx.spgemm(tA, tB, variant=4)
    C, _ = ctx.spgemm(tA, tB, variant=3)
    check(C.download(), A, A)


def test_variants_agree_bitwise(ctx):
    """The binned pipeline and the C-tile method give identical bits on a
    matrix with light (staged) and heavy (C-tile) rows."""
    A = g.gen_rmat(11, 16, seed=4, dtype=g.F64)
    a, _ = run(ctx, A, A, label="rmat11 binned", variant=0)
    b, _ = run(ctx, A, A, label="rmat11 ctile", variant=1)
    assert_tiles_equal(a, b)
    c, _ = run(ctx, A, A, label="rmat11 outer", variant=3)
    assert_tiles_equal(a, c)


def test_deterministic_repeat(ctx):
    A = g.gen_er(3000, 12, seed=9)
    a, _ = run(ctx, A, A, label="rep1")
    b, _ = run(ctx, A, A, label="rep2")
    assert_tiles_equal(a, b)


def test_shape_mismatch_is_an_error(ctx):
    A = g.gen_er(100, 3, seed=1)
    B = g.gen_er(64, 3, seed=1)
    tA = ctx.tiles_from_csr(A)
    tB = ctx.tiles_from_csr(B, as_b=True)
    with pytest.raises(g.TsgError) as e:
        ctx.spgemm(tA, tB)
    assert e.value.code == -1
    tB2 = ctx.tiles_from_csr(A)  # not laid out as B
    with pytest.raises(g.TsgError):
        ctx.spgemm(tA, tB2)


def banded(n, half, seed=0):
    """Dense band |i-j| <= half: heavy tile-rows with full tiles."""
    rows, cols = [], []
    for i in range(n):
        lo, hi = max(0, i - half), min(n, i + half + 1)
        rows.append(np.arange(lo, hi))
    col = np.concatenate(rows).astype(np.int32)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    val = np.random.default_rng(seed).uniform(-1, 1, len(col))
    return g.HostCSR(n, n, rp, col, val)


@pytest.mark.parametrize("name", ["er20000x20", "er70000x32", "rmat12", "er_rect", "band"])
@pytest.mark.parametrize("f32", [False, True])
def test_heavy_rows_paths_agree(ctx, name, f32):
    """Heavy tile-rows: sparse tiles go row-by-row (scalar CSR walk), full
    tiles to the C-tile kernels; both bit-identical to the reference method
    and the oracle.  er70000x32: tile-rows with ~4300 C tiles over 4375 tile
    columns -- several symbolic windows and work items, numeric sub-windows cut
    by column span and values, and Step-1 counts by the workgroup kernel; it
    is checked against the Gustavson recompute (the tiled restatement would
    take minutes)."""
    if name == "er20000x20":
        A = g.gen_er(20000, 20, seed=5); B = A
    elif name == "er70000x32":
        A = g.gen_er(70000, 32, seed=8); B = A
    elif name == "rmat12":
        A = g.gen_rmat(12, 16, seed=3, dtype=g.F64); B = A
    elif name == "er_rect":
        A = g.gen_er(3000, 24, seed=6)
        Bt = g.gen_er(5000, 30, seed=7)             # 5000 x 5000 -> take a 3000 x 5000 B
        S = Bt.to_scipy()[:3000].tocsr()
        B = g.HostCSR(3000, 5000, S.indptr, S.indices, S.data)
    else:
        A = banded(1500, 300); B = A
    got0, s0 = run(ctx, A, B, f32=f32, label=f"{name} binned", variant=0)
    got1, _ = run(ctx, A, B, f32=f32, label=f"{name} ctile", variant=1)
    assert_tiles_equal(got0, got1, keys=["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr",
                                          "mask", "tile_csr_Col", "tile_csr_Val"], label="C ")
    if name in ("er20000x20", "er70000x32"):
        chk = O.gustavson_check(A, B, got0)
        print(f"    gustavson: missing={chk['missing']} extra={chk['extra']} max_scaled={chk['max_scaled']:.3e}",
              flush=True)
        assert chk["missing"] == 0 and chk["extra"] == 0
        assert chk["max_scaled"] <= TOL[g.F32 if f32 else g.F64]
    else:
        check(got0, A, B, f32=f32)


@pytest.mark.parametrize("name", ["er20000x20", "er70000x32", "rmat12", "rmat15", "er_rect"])
@pytest.mark.parametrize("f32", [False, True])
def test_any_order_sums_heavy_rows(ctx, name, f32):
    """sum_order=1: the row-by-row numeric kernel adds products with LDS atomics
    in any order.  C's structure and col bytes stay bit-identical to the
    ascending-k pipeline's; values agree with the Gustavson recompute within
    the FP tolerance (max |c - c_ref| / sum |a b|).  rmat15 has B rows long
    enough that segments go to a wave each."""
    if name == "er20000x20":
        A = g.gen_er(20000, 20, seed=5); B = A
    elif name == "er70000x32":
        A = g.gen_er(70000, 32, seed=8); B = A
    elif name == "rmat12":
        A = g.gen_rmat(12, 16, seed=3, dtype=g.F64); B = A
    elif name == "rmat15":
        A = g.gen_rmat(15, 16, seed=2, dtype=g.F64); B = A
    else:
        A = g.gen_er(3000, 24, seed=6)
        S = g.gen_er(5000, 30, seed=7).to_scipy()[:3000].tocsr()
        B = g.HostCSR(3000, 5000, S.indptr, S.indices, S.data)
    got0, s0 = run(ctx, A, B, f32=f32, label=f"{name} ordered", variant=0)
    got1, s1 = run(ctx, A, B, f32=f32, label=f"{name} any order", variant=0, sum_order=1)
    assert s1["numeric_kernels"] & 32, "the row-by-row numeric kernel did not run"
    assert_tiles_equal(got1, got0, keys=CKEYS[:-1], label="C structure ")
    chk = O.gustavson_check(A, B, got1)
    print(f"    gustavson: missing={chk['missing']} extra={chk['extra']} max_scaled={chk['max_scaled']:.3e}",
          flush=True)
    assert chk["missing"] == 0 and chk["extra"] == 0
    assert chk["max_scaled"] <= TOL[g.F32 if f32 else g.F64]


def _short_rows_wide_b(n=4096, seed=3, per_row=10):
    """A tridiagonal (3 A entries a scalar row) times a B whose rows hold
    per_row columns within +-60: about 30 products a scalar row (the row-lane
    variant's rows) and C tile-rows of about 450 values, so the row-lane
    numeric kernel runs each in several LDS windows."""
    rng = np.random.default_rng(seed)
    i = np.arange(n)
    ca = np.stack([i - 1, i, i + 1], 1)
    keep = (ca >= 0) & (ca < n)
    rpa = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    A = g.HostCSR(n, n, rpa, ca[keep].astype(np.int32), rng.uniform(-1, 1, int(rpa[-1])))
    cols = []
    for k in range(n):
        lo, hi = max(0, k - 60), min(n, k + 61)
        cols.append(np.sort(rng.choice(np.arange(lo, hi), min(per_row, hi - lo), replace=False)))
    rpb = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
    B = g.HostCSR(n, n, rpb, np.concatenate(cols).astype(np.int32), rng.uniform(-1, 1, int(rpb[-1])))
    return A, B


@pytest.mark.parametrize("var_mask", [None, 24, 31])
@pytest.mark.parametrize("f32", [False, True])
def test_row_lane_variant_with_windows(ctx, var_mask, f32):
    """Staged variant R (16 lanes a tile-row for the symbolic step, a lane per
    scalar row with C values staged in LDS windows for the numeric step) on
    tile-rows needing several windows, bit for bit against the tiled
    restatement.  var_mask bit 4 gives R's rows their own kernels at this size
    (24: R the only staged variant; 31: every variant); None: R's few rows run
    VarS's kernels."""
    A, B = _short_rows_wide_b()
    print(f"[case] row-lane windows var_mask={var_mask} f32={f32}", flush=True)
    H, HB = (A.astype(g.F32), B.astype(g.F32)) if f32 else (A, B)
    tA, tB = ctx.tiles_from_csr(H), ctx.tiles_from_csr(HB, as_b=True)
    if var_mask is not None:
        ctx.autotune_set(tA, tB, var_mask, 64)
    try:
        C, st = ctx.spgemm(tA, tB)
        got = C.download()
    finally:
        ctx.autotune_clear()
    assert st["nnzC"] / max(1, tA.s.tilem) > 416       # more than two 208-value windows a tile-row
    if var_mask is not None:
        assert st["numeric_kernels"] & 1, st["numeric_kernels"]   # k_num_rows ran on rows
    else:
        assert not st["numeric_kernels"] & 1, st["numeric_kernels"]
    assert_tiles_equal(got, O.spgemm_tiled(A, B, f32=f32), keys=CKEYS, label="C ")


@pytest.mark.parametrize("f32", [False, True])
def test_row_lane_kernels_on_wide_columns(ctx, f32):
    """A 2-D Laplacian 8192 points wide: every tile-row fits variant R but its
    C tile columns span more than 1024, so the row-lane symbolic kernel ranks
    them by counting instead of its bitmap; bit for bit against the tiled
    restatement."""
    A = g.gen_laplace2d(8192, 8)
    print(f"[case] row-lane wide columns f32={f32}", flush=True)
    H = A.astype(g.F32) if f32 else A
    tA, tB = ctx.tiles_from_csr(H), ctx.tiles_from_csr(H, as_b=True)
    ctx.autotune_set(tA, tB, 24, 64)
    try:
        C, st = ctx.spgemm(tA, tB)
        got = C.download()
    finally:
        ctx.autotune_clear()
    assert st["numeric_kernels"] == 1, st["numeric_kernels"]   # k_num_rows only: every row in R
    assert_tiles_equal(got, O.spgemm_tiled(A, A, f32=f32), keys=CKEYS, label="C ")
