"""Batched products (tsg_spgemm_batched): every C_i bit-identical to the
single-product call on the same operands, and to the oracle."""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
CKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
         "tile_csr_Val"]


def problems(f32=False):
    r = np.random.default_rng(7)
    ps = []
    for i in range(20):
        kind = i % 4
        if kind == 0:
            A = g.gen_er(int(r.integers(17, 3000)), int(r.integers(1, 20)), seed=i)
            B = A
        elif kind == 1:
            A = g.gen_laplace3d(int(r.integers(2, 14)), int(r.integers(2, 14)), int(r.integers(2, 14)))
            B = A
        elif kind == 2:           # rectangular: A m x k, B k x n
            m, k, n = (int(x) for x in r.integers(5, 700, 3))
            A = g.gen_er_var(m, 0, 9, seed=i)
            A = g.HostCSR(m, k, A.rowptr, A.col % k, A.val)
            A = sort_rows(A)
            Bt = g.gen_er_var(n, 0, 7, seed=100 + i)
            B = g.HostCSR(n, k, Bt.rowptr, Bt.col % k, Bt.val)
            B = sort_rows(B).transpose()
        else:
            A = g.gen_rmat(int(r.integers(6, 10)), 8, seed=i, dtype=g.F64)
            B = A
        ps.append((A, B))
    Z = g.HostCSR(40, 40, np.zeros(41, np.int64), np.zeros(0, np.int32), np.zeros(0))
    ps.insert(5, (Z, Z))
    one = g.HostCSR(3, 3, [0, 1, 1, 1], [2], [2.5])
    ps.insert(11, (one, one))
    if f32:
        ps = [(a.astype(g.F32), b.astype(g.F32)) for a, b in ps]
    return ps


def sort_rows(A):
    """Sorted, duplicate-free columns per row (for the modulo-folded inputs)."""
    S = A.to_scipy().tocsr()
    S.sum_duplicates()
    S.sort_indices()
    return g.HostCSR(A.m, A.n, S.indptr, S.indices, S.data)


@pytest.mark.parametrize("f32", [False, True])
def test_batched_equals_single_calls(ctx, f32):
    ps = problems(f32)
    print(f"[case] batched {len(ps)} problems f32={f32}", flush=True)
    tA = [ctx.tiles_from_csr(a) for a, _ in ps]
    tB = [ctx.tiles_from_csr(b, as_b=True) for _, b in ps]
    Cs, sts = ctx.spgemm_batched(tA, tB)
    assert len(Cs) == len(ps)
    for i, (a, b) in enumerate(ps):
        C1, s1 = ctx.spgemm(tA[i], tB[i])
        got = Cs[i].download()
        assert_tiles_equal(got, C1.download(), keys=CKEYS, label=f"problem {i} ")
        assert sts[i]["numtileC"] == s1["numtileC"] and sts[i]["nnzC"] == s1["nnzC"]
        if i % 5 == 0:
            exp = O.spgemm_tiled(a.astype(g.F64) if f32 else a, b.astype(g.F64) if f32 else b, f32=f32)
            assert_tiles_equal(got, exp, keys=CKEYS, label=f"oracle {i} ")


def test_batched_reference_method_and_learned_path(ctx):
    ps = problems()[:9]
    tA = [ctx.tiles_from_csr(a) for a, _ in ps]
    tB = [ctx.tiles_from_csr(b, as_b=True) for _, b in ps]
    base, _ = ctx.spgemm_batched(tA, tB)
    ref, _ = ctx.spgemm_batched(tA, tB, variant=1)
    for i in range(len(ps)):
        assert_tiles_equal(ref[i].download(), base[i].download(), keys=CKEYS, label=f"ctile {i} ")


def test_batched_with_duplicates_is_exact(ctx):
    rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9], dtype=np.int64)
    col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6], dtype=np.int32)
    D = g.HostCSR(8, 8, rp, col, np.arange(1.0, 10.0))
    E = g.gen_er(300, 6, seed=3)
    ps = [(E, E), (D, D), (E, E)]
    tA = [ctx.tiles_from_csr(a) for a, _ in ps]
    tB = [ctx.tiles_from_csr(b, as_b=True) for _, b in ps]
    Cs, _ = ctx.spgemm_batched(tA, tB)
    for i, (a, b) in enumerate(ps):
        exp = O.spgemm_tiled(a, b)
        assert_tiles_equal(Cs[i].download(), exp, keys=["tile_ptr", "tile_columnidx", "tile_nnz", "mask",
                                                        "tile_csr_Col", "tile_csr_Val"], label=f"{i} ")


@pytest.mark.parametrize("order", ["forward", "reverse", "interleaved"])
def test_batched_outputs_share_the_product_memory(ctx, order):
    """Each C_i's masks, row offsets, values and col bytes are views into the
    batch product's arrays: freeing the C_i in any order (and re-laying one as
    a B operand, which frees its views) returns every byte to the pool, and a
    re-laid C_i still multiplies correctly."""
    print(f"[case] batched views, free order {order}", flush=True)
    ps = problems()[:10]
    tA = [ctx.tiles_from_csr(a) for a, _ in ps]
    tB = [ctx.tiles_from_csr(b, as_b=True) for _, b in ps]
    ctx.synchronize()
    base_in_use, _ = ctx.pool_stats()
    Cs, _ = ctx.spgemm_batched(tA, tB)
    exp = [C1.download() for C1 in (ctx.spgemm(tA[i], tB[i])[0] for i in range(len(ps)))]
    ctx.synchronize()
    # chained use of one output: re-lay C_0 (= A_0 B_0, square ER) as a B operand
    a0, b0 = ps[0]
    Cs[0].build_csc_tiles()
    C2, _ = ctx.spgemm(tA[0], Cs[0])
    AB = O.gustavson_csr(a0, b0)
    chk = O.gustavson_check(a0, g.HostCSR(a0.m, b0.n, *AB), C2.download())
    assert chk["missing"] == 0 and chk["extra"] == 0 and chk["max_scaled"] <= 1e-12
    C2.free()
    idx = list(range(len(Cs)))
    if order == "reverse":
        idx.reverse()
    elif order == "interleaved":
        idx = idx[::2] + idx[1::2]
    for k, i in enumerate(idx):
        if i != 0:
            assert_tiles_equal(Cs[i].download(), exp[i], keys=CKEYS, label=f"problem {i} ")
        Cs[i].free()
    ctx.synchronize()
    in_use, _ = ctx.pool_stats()
    assert in_use == base_in_use, (in_use, base_in_use)


def test_batched_errors_and_empty(ctx):
    Cs, sts = ctx.spgemm_batched([], [])
    assert Cs == [] and sts == []
    A = g.gen_er(100, 3, seed=1)
    t64 = ctx.tiles_from_csr(A)
    b64 = ctx.tiles_from_csr(A, as_b=True)
    b32 = ctx.tiles_from_csr(A.astype(g.F32), as_b=True)
    t32 = ctx.tiles_from_csr(A.astype(g.F32))
    with pytest.raises(g.TsgError):
        ctx.spgemm_batched([t64, t32], [b64, b32])       # mixed value types
    with pytest.raises(g.TsgError):
        ctx.spgemm_batched([t64], [t64])                 # B not laid out as a B operand


def test_uneven_batch_generator_is_deterministic():
    a = g.gen_batch_uneven(count=6, seed=4)
    b = g.gen_batch_uneven(count=6, seed=4)
    assert [n for n, _ in a] == [n for n, _ in b]
    for (_, x), (_, y) in zip(a, b):
        assert np.array_equal(x.col, y.col) and np.array_equal(x.val, y.val)
