"""Matrix-core numeric variant (numeric_variant=2): dense tile pairs through
v_mfma_f64/f32_16x16x4, the rest one FMA per product.  Structure bit-exact;
values within max |c - c_ref| / sum|a*b| <= 1e-12 (FP64) / 1e-5 (FP32) of the
FP64 recompute (the matrix cores' accumulation rounding is the hardware's);
the measured error and whether the bits match the other kernels are printed."""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
SKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col"]
TOL = {False: 1e-12, True: 1e-5}


def run(ctx, A, f32, **kw):
    h = A.astype(g.F32) if f32 else A
    tA, tB = ctx.tiles_from_csr(h), ctx.tiles_from_csr(h, as_b=True)
    C, st = ctx.spgemm(tA, tB, **kw)
    return C.download()


@pytest.mark.parametrize("name", ["band", "lap2d", "er"])
@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("minp", [1, 64])
def test_mfma_variant_parity(ctx, name, f32, minp):
    A = {"band": lambda: g.gen_band(700, 40, seed=1), "lap2d": lambda: g.gen_laplace2d(90, 70),
         "er": lambda: g.gen_er(2000, 30, seed=2)}[name]()
    print(f"[case] mfma {name} f32={f32} min_products={minp}", flush=True)
    got = run(ctx, A, f32, variant=2, mfma_min_products=minp)
    ref = run(ctx, A, f32, variant=1)
    assert_tiles_equal(got, ref, keys=SKEYS, label="C ")
    chk = O.gustavson_check(A, A, got)
    same = np.array_equal(got["tile_csr_Val"].view(np.uint8), ref["tile_csr_Val"].view(np.uint8))
    print(f"    max_scaled={chk['max_scaled']:.3e} rel_frob={chk['rel_frob']:.3e} bits equal to C-tile kernel: {same}",
          flush=True)
    assert chk["missing"] == 0 and chk["extra"] == 0
    assert chk["max_scaled"] <= TOL[f32]


def test_mfma_gate_off_is_bit_exact(ctx):
    A = g.gen_band(500, 30, seed=3)
    got = run(ctx, A, False, variant=2, mfma_min_products=1 << 20)   # no pair reaches the gate
    ref = run(ctx, A, False, variant=1)
    assert_tiles_equal(got, ref, keys=SKEYS + ["tile_csr_Val"], label="C ")


def test_mfma_lane_layout_on_one_dense_tile(ctx):
    """One 16x16 dense tile times itself: checks the operand and accumulator
    lane layouts against numpy."""
    r = np.random.default_rng(0)
    D = r.uniform(-1, 1, (16, 16))
    rp = np.arange(0, 257, 16, dtype=np.int64)
    col = np.tile(np.arange(16, dtype=np.int32), 16)
    A = g.HostCSR(16, 16, rp, col, D.ravel())
    got = run(ctx, A, False, variant=2, mfma_min_products=1)
    C = np.zeros((16, 16))
    for rr in range(16):
        s0 = got["tile_csr_Ptr"][rr]
        s1 = got["nnz"] if rr == 15 else got["tile_csr_Ptr"][rr + 1]
        for e in range(s0, s1):
            C[rr, got["tile_csr_Col"][e] & 15] = got["tile_csr_Val"][e]
    assert np.allclose(C, D @ D, rtol=1e-13, atol=1e-13)


def test_mfma_rejects_duplicates(ctx):
    rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9], dtype=np.int64)
    col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6], dtype=np.int32)
    A = g.HostCSR(8, 8, rp, col, np.arange(1.0, 10.0))
    with pytest.raises(g.TsgError):
        run(ctx, A, False, variant=2)
