"""Autotuner: every kernel-variant configuration gives the same bits; the
tuner reports every candidate, stores the fastest and later products of the
class use it."""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
CKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
         "tile_csr_Val"]


def mixed():
    # R-MAT: light rows (staged), heavy sparse rows (row-by-row or C-tile)
    return g.gen_rmat(11, 16, seed=4, dtype=g.F64)


@pytest.mark.parametrize("var_mask", [15, 31, 24, 8, 9, 7, 5, 1, 4, 2, 0])
@pytest.mark.parametrize("rg_density", [0, 2, 8, 1000])
def test_every_configuration_gives_identical_bits(ctx, var_mask, rg_density):
    A = mixed()
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    ref, _ = ctx.spgemm(tA, tB, variant=1)          # reference method
    ctx.autotune_set(tA, tB, var_mask, rg_density)
    assert ctx.autotune_query(tA, tB)["var_mask"] == var_mask
    got, _ = ctx.spgemm(tA, tB)
    assert_tiles_equal(got.download(), ref.download(), keys=CKEYS, label=f"cfg {var_mask}/{rg_density} ")
    dflt, _ = ctx.spgemm(tA, tB, tuning=1)          # built-in defaults ignore the stored one
    assert_tiles_equal(dflt.download(), ref.download(), keys=CKEYS, label="defaults ")
    ctx.autotune_clear()
    assert ctx.autotune_query(tA, tB) is None


def test_autotune_picks_the_fastest_and_stores_it(ctx):
    A = mixed()
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    best, rep = ctx.autotune(tA, tB, reps=2)
    print("    report:", [(r["var_mask"], r["rg_density"], round(r["ms"], 3)) for r in rep], flush=True)
    assert len(rep) == 30   # 7 staged-variant masks x 4 row-by-row densities + 2 matrix-core gates
    assert best["ms"] == min(r["ms"] for r in rep)
    q = ctx.autotune_query(tA, tB)
    assert q["var_mask"] == best["var_mask"] and q["rg_density"] == best["rg_density"]
    # a different matrix of the same class uses it too
    A2 = g.gen_rmat(11, 16, seed=9, dtype=g.F64)
    tA2, tB2 = ctx.tiles_from_csr(A2), ctx.tiles_from_csr(A2, as_b=True)
    if ctx.autotune_query(tA2, tB2) is not None:
        r1, _ = ctx.spgemm(tA2, tB2)
        r2, _ = ctx.spgemm(tA2, tB2, variant=1)
        assert_tiles_equal(r1.download(), r2.download(), keys=CKEYS, label="class member ")
    ctx.autotune_clear()


@pytest.mark.parametrize("mfma", [64, 512])
def test_tuned_matrix_core_configuration(ctx, mfma):
    A = g.gen_band(900, 40, seed=5)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    ref, _ = ctx.spgemm(tA, tB, variant=1)
    ctx.autotune_set(tA, tB, 0, 0, mfma)
    assert ctx.autotune_query(tA, tB)["mfma_min_products"] == mfma
    got, _ = ctx.spgemm(tA, tB)
    assert_tiles_equal(got.download(), ref.download(), keys=CKEYS, label=f"mfma {mfma} ")
    ctx.autotune_clear()


def test_autotune_prefers_matrix_cores_on_full_tiles(ctx):
    A = g.gen_band(16384, 48, seed=6)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    best, rep = ctx.autotune(tA, tB, reps=2)
    print("    band report:", [(r["var_mask"], r["rg_density"], r["mfma_min_products"], round(r["ms"], 3))
                               for r in rep], flush=True)
    assert best["mfma_min_products"] > 0
    ctx.autotune_clear()


def test_autotune_batch(ctx):
    ps = [g.gen_er(int(n), int(k), seed=int(n)) for n, k in [(500, 4), (3000, 20), (1200, 9), (64, 2)]]
    tA = [ctx.tiles_from_csr(a) for a in ps]
    tB = [ctx.tiles_from_csr(a, as_b=True) for a in ps]
    best, rep = ctx.autotune(tA, tB, reps=1)
    assert len(rep) == 30 and ctx.autotune_query(tA, tB) is not None
    Cs, _ = ctx.spgemm_batched(tA, tB)
    for i in range(len(ps)):
        C1, _ = ctx.spgemm(tA[i], tB[i], variant=1)
        assert_tiles_equal(Cs[i].download(), C1.download(), keys=CKEYS, label=f"problem {i} ")
    ctx.autotune_clear()


def test_autotune_large_batch_samples_then_times_finalists(ctx):
    """A batch of 16 or more products: every configuration is timed on every
    7th product, the 3 fastest there on the whole batch; the stored winner is
    the fastest of those, and the batch under it equals the C-tile method."""
    ps = [g.gen_er(300 + 37 * i, 3 + (i % 5), seed=200 + i) if i % 2 else g.gen_laplace2d(20 + i, 17)
          for i in range(20)]
    tA = [ctx.tiles_from_csr(a) for a in ps]
    tB = [ctx.tiles_from_csr(a, as_b=True) for a in ps]
    best, rep = ctx.autotune(tA, tB, reps=1)
    full = [r for r in rep if r["sampled"] == 0]
    print("    finalists:", [(r["var_mask"], r["rg_density"], r["mfma_min_products"], round(r["ms"], 3)) for r in full],
          flush=True)
    assert len(rep) == 30 and len(full) == 3
    assert best["sampled"] == 0 and best["ms"] == min(r["ms"] for r in full)
    q = ctx.autotune_query(tA, tB)
    assert q is not None and q["var_mask"] == best["var_mask"] and q["rg_density"] == best["rg_density"]
    Cs, _ = ctx.spgemm_batched(tA, tB)
    for i in range(len(ps)):
        C1, _ = ctx.spgemm(tA[i], tB[i], variant=1)
        assert_tiles_equal(Cs[i].download(), C1.download(), keys=CKEYS, label=f"problem {i} ")
    ctx.autotune_clear()
