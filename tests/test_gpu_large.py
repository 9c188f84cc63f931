"""Products whose kernels cover more than 2^32 work-items at one thread per
tile lane: 500,000 x 500,000 Erdos-Renyi with 7 entries per row gives C about
3.8e8 tiles (mostly with one entry), so a one-16-lane-group-per-tile launch
needs 6e9 work-items; every such kernel runs a capped grid and strides.
Checked through size-independent properties: nnz(C) and the structure of C
unpacked to CSR (tile2csr on the GPU) equal the oracle's Gustavson product
bit for bit; values within max |c - c_ref| / sum_k |a_ik b_kj| <= 1e-12 (the
Gustavson loop multiplies and adds separately, the kernels fuse)."""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    A = g.gen_er(500_000, 7, seed=11)
    rp, col, val = O.gustavson_csr(A, A)
    Aabs = g.HostCSR(A.m, A.n, A.rowptr, A.col, np.abs(A.val))
    _, _, scale = O.gustavson_csr(Aabs, Aabs)
    return A, rp, col, val, scale


@pytest.mark.parametrize("variant", [1, 0])
def test_more_than_2e32_work_items(ctx, big, variant):
    A, rp, col, val, scale = big
    print(f"[case] large ER 500k x 7, variant {variant}", flush=True)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    C, st = ctx.spgemm(tA, tB, variant=variant)
    print(f"    numtileC {st['numtileC']} nnzC {st['nnzC']}", flush=True)
    assert st["numtileC"] * 16 > 2 ** 32
    assert st["nnzC"] == len(col)
    H = C.to_csr()
    assert np.array_equal(H.rowptr, rp) and np.array_equal(H.col, col)
    err = np.max(np.abs(H.val - val) / np.maximum(scale, 1e-300))
    print(f"    max scaled error {err:.3e}", flush=True)
    assert err <= 1e-12
    C.free()
    tA.free()
    tB.free()
