"""Plans: new operand values, same structure, values of C recomputed by a
captured hipGraph -- every replay bit-identical to a fresh product with the
new values (the graph is replayed several times with eager work and syncs in
between, each replay checked)."""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
CKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
         "tile_csr_Val"]


def with_values(A, seed):
    v = np.random.default_rng(seed).uniform(-1, 1, A.nnz).astype(A.val.dtype)
    return g.HostCSR(A.m, A.n, A.rowptr, A.col, v)


def fresh(ctx, A, B, **kw):
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(B, as_b=True)
    C, _ = ctx.spgemm(tA, tB, **kw)
    return C.download()


def test_set_values_rewrites_every_copy(ctx):
    A = g.gen_rmat(10, 8, seed=2, dtype=g.F64)
    A2 = with_values(A, 5)
    t = ctx.tiles_from_csr(A, as_b=True)
    ctx.set_values(t, A2)
    ref = ctx.tiles_from_csr(A2, as_b=True).download()
    got = t.download()
    for k in ("tile_csr_Val", "csr_col_Val"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("name", ["lap2d", "rmat", "er_heavy", "band"])
@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("f32", [False, True])
def test_plan_single_product_replays(ctx, name, variant, f32):
    A = {"lap2d": lambda: g.gen_laplace2d(120, 90), "rmat": lambda: g.gen_rmat(11, 12, seed=3, dtype=g.F64),
         "er_heavy": lambda: g.gen_er(6000, 20, seed=4), "band": lambda: g.gen_band(600, 30, seed=5)}[name]()
    if f32:
        A = A.astype(g.F32)
    print(f"[case] plan {name} variant={variant} f32={f32}", flush=True)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    kw = dict(variant=variant, mfma_min_products=16 if variant == 2 else 0)
    plan = ctx.plan(tA, tB, **kw)
    assert_tiles_equal(plan.outputs[0].download(), fresh(ctx, A, A, **kw), keys=CKEYS, label="create ")
    for rep in range(3):
        A2 = with_values(A, 100 + rep)
        ctx.set_values(tA, A2)
        ctx.set_values(tB, A2)
        plan.execute()
        ctx.synchronize()
        _ = fresh(ctx, g.gen_er(200, 3, seed=rep), g.gen_er(200, 3, seed=rep))   # eager work between replays
        assert_tiles_equal(plan.outputs[0].download(), fresh(ctx, A2, A2, **kw), keys=CKEYS, label=f"replay {rep} ")
    plan.destroy()


def test_plan_batch_replays(ctx):
    ps = [g.gen_er(1500, 6, seed=1), g.gen_laplace3d(9, 8, 7), g.gen_er(5000, 24, seed=2),
          g.gen_rmat(9, 8, seed=3, dtype=g.F64), g.gen_band(300, 20, seed=4)]
    tA = [ctx.tiles_from_csr(a) for a in ps]
    tB = [ctx.tiles_from_csr(a, as_b=True) for a in ps]
    plan = ctx.plan(tA, tB)
    for i, a in enumerate(ps):
        assert_tiles_equal(plan.outputs[i].download(), fresh(ctx, a, a), keys=CKEYS, label=f"create {i} ")
    for rep in range(3):
        news = [with_values(a, 10 * rep + i) for i, a in enumerate(ps)]
        for i, a2 in enumerate(news):
            ctx.set_values(tA[i], a2)
            ctx.set_values(tB[i], a2)
        plan.execute()
        ctx.synchronize()
        for i, a2 in enumerate(news):
            assert_tiles_equal(plan.outputs[i].download(), fresh(ctx, a2, a2), keys=CKEYS,
                               label=f"replay {rep} problem {i} ")
    plan.destroy()


def test_set_values_needs_positions(ctx):
    A = g.gen_er(300, 4, seed=1)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    C, _ = ctx.spgemm(tA, tB)
    with pytest.raises(g.TsgError):
        ctx.set_values(C, A)      # outputs keep no value positions
