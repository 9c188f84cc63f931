"""The benchmark configurations at their full sizes on the GPU.

  * R-MAT scale 20 (1,048,576 rows, edge factor 16, (0.57, 0.19, 0.19, 0.05),
    FP32, seed 1), C = A*A: C holds 1.78e9 tiles and 9.7e9 nonzeros (162 GB
    in HBM), so it is checked without bringing it back whole: nnz(C) and
    numtileC against the oracle's threaded count over the whole product,
    eight sampled tile-row blocks (the heaviest first) against the
    reference's Gustavson check (src/spgemm_raw_v2.cu:253-359: 0 missing,
    0 extra, max |c - c_ref| / sum |a b| <= 1e-5 for FP32), three of them bit
    for bit against the oracle's tiled restatement of Steps 1-3;
  * the uneven batch of 512 products (n_i log-uniform in [256, 65536], half
    7-point Laplacians, half Erdos-Renyi with 2-32 per row, FP64, seed 4) in
    one batched call: every C_i against the Gustavson check (FP64, 1e-12),
    every 16th bit for bit against the tiled restatement;
  * the Galerkin product P^T (A P) on the 27-point 128^3 stencil (FP64):
    integer-valued, so C1 = A P and C2 = P^T C1 (C1 re-laid as a B operand on
    the GPU) must equal scipy's products exactly, and C2's tiles the tiled
    restatement's bit for bit.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
CKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
         "tile_csr_Val"]
NTHREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _rmat_blocks(tilem):
    """Tile-row blocks checked: the heaviest two tile-rows, then one tile-row
    at each eighth of the matrix."""
    return [(0, 2)] + [(i * tilem // 8, i * tilem // 8 + 1) for i in range(1, 8)]


@pytest.fixture(scope="module")
def rmat20():
    A = g.gen_rmat(20, 16, seed=1, dtype=g.F32)
    print(f"[rmat20] nnzA {A.nnz}", flush=True)
    nnz, tiles = O.count_product(A, A, nthreads=NTHREADS)
    print(f"[rmat20] oracle count: nnzC {nnz} numtileC {tiles}", flush=True)
    return {"A": A, "nnz": nnz, "tiles": tiles}


def check_rmat20_blocks(ctx, C, A, tiled_idx=(2, 4, 6), label=""):
    """Downloads the sampled blocks of C, checks them, returns them."""
    tilem = C.s.tilem
    blocks = _rmat_blocks(tilem)
    got = [C.download_rows(lo, hi) for lo, hi in blocks]

    def gus(i):
        lo, hi = blocks[i]
        return O.gustavson_check(A, A, O.embed_rows(got[i], lo, tilem), row_lo=lo * 16, row_hi=hi * 16)

    with cf.ThreadPoolExecutor(NTHREADS) as ex:
        chks = list(ex.map(gus, range(len(blocks))))
    for (lo, hi), d, chk in zip(blocks, got, chks):
        print(f"    {label}block [{lo},{hi}): tiles {d['numtile']} nnz {d['nnz']} missing {chk['missing']} "
              f"extra {chk['extra']} max_scaled {chk['max_scaled']:.2e}", flush=True)
        assert chk["missing"] == 0 and chk["extra"] == 0 and chk["total_gpu"] == d["nnz"]
        assert chk["max_scaled"] <= 1e-5
    exp = O.spgemm_tiled_rows(A, A, [blocks[i] for i in tiled_idx], f32=True) if tiled_idx else []
    for i, e in zip(tiled_idx, exp):
        assert_tiles_equal(got[i], e, keys=CKEYS, label=f"{label}block {blocks[i]} ")
    return got


def test_rmat20_exact(ctx, rmat20):
    A, nnz, tiles = rmat20["A"], rmat20["nnz"], rmat20["tiles"]
    print("[case] R-MAT 20 exact pipeline", flush=True)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    C, st = ctx.spgemm(tA, tB)
    print(f"    numtileC {st['numtileC']} nnzC {st['nnzC']}", flush=True)
    assert st["numtileC"] == tiles and st["nnzC"] == nnz
    rmat20["exact_blocks"] = check_rmat20_blocks(ctx, C, A)
    C.free()
    tA.free()
    tB.free()
    ctx.pool_trim()   # 162 GB of freed C back to the device for the next context


def test_rmat20_any_order(ctx, rmat20):
    """sum_order=1 (heavy-row products added in any order): the same tile set,
    masks and col bytes as the exact pipeline, values within the FP32
    tolerance of the Gustavson recompute on the sampled blocks."""
    A = rmat20["A"]
    print("[case] R-MAT 20 exact pipeline, any-order sums", flush=True)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    C, st = ctx.spgemm(tA, tB, sum_order=1)
    print(f"    numtileC {st['numtileC']} nnzC {st['nnzC']}", flush=True)
    assert st["numtileC"] == rmat20["tiles"] and st["nnzC"] == rmat20["nnz"]
    got = check_rmat20_blocks(ctx, C, A, tiled_idx=(), label="any-order ")
    if "exact_blocks" in rmat20:
        for i, (e, l) in enumerate(zip(rmat20["exact_blocks"], got)):
            assert_tiles_equal(l, e, keys=CKEYS[:-1], label=f"any-order block {i} ")
    C.free()
    tA.free()
    tB.free()
    ctx.pool_trim()


def test_rmat20_learned(rmat20):
    """The learned scheduler on its own configuration: trained on R-MATs of
    scales 17-19 with other seeds and a mixed family (never this product), it
    sizes C's 9.7e9 values without a re-run, and C equals the exact
    pipeline's."""
    A = rmat20["A"]
    print("[case] R-MAT 20 learned scheduler", flush=True)
    c = g.Context(0)
    try:
        fam = [g.gen_rmat(sc, 16, seed=100 + sc, dtype=g.F32) for sc in (17, 18, 19)]
        fam += [g.gen_laplace2d(256, 256), g.gen_er(20000, 6, seed=11), g.gen_er_var(20000, 0, 24, seed=14)]
        for M in fam:
            tA, tB = c.tiles_from_csr(M), c.tiles_from_csr(M, as_b=True)
            c.learned_collect(tA, tB)
            tA.free()
            tB.free()
        loss = c.learned_train(epochs=300, batch=4096, lr=3e-3, seed=0)
        print(f"    trained: loss {loss:.4f} calibration {c.learned_calibration()}", flush=True)
        tA, tB = c.tiles_from_csr(A), c.tiles_from_csr(A, as_b=True)
        C, st = c.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
        print(f"    fallback {st['fallback_used']} capacity {st['pred_nnzC']} nnz {st['nnzC']} "
              f"predicted tiles {st['pred_numtileC']} tiles {st['numtileC']}", flush=True)
        assert st["fallback_used"] == 0
        assert st["numtileC"] == rmat20["tiles"] and st["nnzC"] == rmat20["nnz"]
        assert st["pred_nnzC"] <= 1.5 * st["nnzC"]
        got = check_rmat20_blocks(c, C, A, tiled_idx=(), label="learned ")
        if "exact_blocks" in rmat20:
            for i, (e, l) in enumerate(zip(rmat20["exact_blocks"], got)):
                assert_tiles_equal(l, e, keys=CKEYS, label=f"learned block {i} ")
        C.free()
        tA.free()
        tB.free()
    finally:
        c.close()


def test_batch512_uneven(ctx):
    probs = g.gen_batch_uneven(512, seed=4)
    print(f"[case] batch of {len(probs)} uneven products", flush=True)
    tA = [ctx.tiles_from_csr(A) for _, A in probs]
    tB = [ctx.tiles_from_csr(A, as_b=True) for _, A in probs]
    Cs, sts = ctx.spgemm_batched(tA, tB)
    print(f"    C tiles {sum(s['numtileC'] for s in sts)} nnz {sum(s['nnzC'] for s in sts)}", flush=True)

    def check(i):
        A = probs[i][1]
        return O.gustavson_check(A, A, got[i])

    bad = []
    for lo in range(0, len(probs), 64):       # download in groups, check in threads
        got = {i: Cs[i].download() for i in range(lo, min(lo + 64, len(probs)))}
        with cf.ThreadPoolExecutor(NTHREADS) as ex:
            chks = dict(zip(got, ex.map(check, got)))
        for i, chk in chks.items():
            if not (chk["missing"] == 0 and chk["extra"] == 0 and chk["max_scaled"] <= 1e-12):
                bad.append((i, probs[i][0], chk))
        for i in got:
            if i % 16 == 0:
                A = probs[i][1]
                assert_tiles_equal(got[i], O.spgemm_tiled(A, A), keys=CKEYS, label=f"problem {i} {probs[i][0]} ")
        del got
    assert not bad, f"{len(bad)} products fail the Gustavson check, first {bad[:3]}"
    for C in Cs:
        C.free()


def test_galerkin_128cube(ctx):
    import scipy.sparse as sp
    print("[case] Galerkin P^T (A P), 27-point stencil 128^3", flush=True)
    n = 128
    A = g.gen_stencil27(n, n, n)
    P = g.gen_aggregation(n, n, n)
    Pt = P.transpose()
    tA, tP, tPt = ctx.tiles_from_csr(A), ctx.tiles_from_csr(P, as_b=True), ctx.tiles_from_csr(Pt)
    C1, s1 = ctx.spgemm(tA, tP)
    C1.build_csc_tiles()            # C1 becomes the B operand on the GPU
    C2, s2 = ctx.spgemm(tPt, C1)
    print(f"    C1 tiles {s1['numtileC']} nnz {s1['nnzC']}; C2 tiles {s2['numtileC']} nnz {s2['nnzC']}", flush=True)
    sA, sP = A.to_scipy(), P.to_scipy()
    e1 = (sA @ sP).tocsr()
    e2 = (sP.T.tocsr() @ e1).tocsr()
    e1.sort_indices()
    e2.sort_indices()
    h2 = C2.to_csr()
    assert np.array_equal(h2.rowptr, e2.indptr) and np.array_equal(h2.col, e2.indices)
    assert np.array_equal(h2.val, e2.data)          # integer-valued: exact
    assert (s1["numtileC"], s1["nnzC"]) == (903224, 16387064)
    assert (s2["numtileC"], s2["nnzC"]) == (361000, 6859000)
    # C2's tiles against the tiled restatement of P^T * C1 (C1 from scipy)
    C1h = g.HostCSR(e1.shape[0], e1.shape[1], e1.indptr, e1.indices, e1.data)
    assert_tiles_equal(C2.download(), O.spgemm_tiled(Pt, C1h), keys=CKEYS, label="C2 ")
    for M in (C2, C1, tA, tP, tPt):
        M.free()
