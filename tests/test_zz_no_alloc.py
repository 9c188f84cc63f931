"""tsg_spgemm_noalloc: the product with no device allocation (work arrays in
the library's static scratch, C in the caller's arrays) against tsg_spgemm's
result on the same operands, bit for bit (structure and values), on calls in a
row -- the second and third on smaller inputs than the first, so stale scratch
contents from the earlier call would show -- and its overflow errors, which
must name the sizes needed.  Heavy tile-rows (R-MAT) take the C-tile kernels
in this mode while tsg_spgemm takes the row-by-row ones: same bits."""
import numpy as np
import pytest

import gpu_spgemm_amd as g

pytestmark = pytest.mark.gpu
KEYS = ["tile_ptr", "tile_rowidx", "tile_columnidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
        "tile_csr_Val"]


def _buffers(ctx, tilem, cap_tiles, cap_nnz, f32):
    """The caller's C arrays, from the library's pool (addresses by field name)."""
    ct, cn = max(cap_tiles, 1), max(cap_nnz, 1)
    return {
        "tile_ptr": ctx.malloc(8 * (tilem + 1)),
        "tile_rowidx": ctx.malloc(4 * ct),
        "tile_columnidx": ctx.malloc(4 * ct),
        "tile_nnz": ctx.malloc(8 * (cap_tiles + 1)),
        "mask": ctx.malloc(2 * 16 * ct),
        "tile_csr_Ptr": ctx.malloc(16 * ct),
        "tile_csr_Val": ctx.malloc((4 if f32 else 8) * cn),
        "tile_csr_Col": ctx.malloc(cn),
    }


def _release(ctx, bufs):
    for p in bufs.values():
        ctx.free(p)


def _download(ctx, s):
    t = g.DeviceTiles(ctx)
    t.s = s
    try:
        return t.download()
    finally:
        t.s = g.TileStruct()   # the arrays are the caller's: never freed by the library


def _run(ctx, A, f32=False, slack=1.0):
    H = A.astype(g.F32) if f32 else A
    tA, tB = ctx.tiles_from_csr(H), ctx.tiles_from_csr(H, as_b=True)
    Ce, se = ctx.spgemm(tA, tB)
    exp = Ce.download()
    cap_t, cap_n = int(se["numtileC"] * slack) + 3, int(se["nnzC"] * slack) + 5
    bufs = _buffers(ctx, tA.s.tilem, cap_t, cap_n, f32)
    cs, sn = ctx.spgemm_noalloc(tA, tB, bufs, cap_t, cap_n)
    assert sn["numtileC"] == se["numtileC"] and sn["nnzC"] == se["nnzC"]
    assert cs.numtile == se["numtileC"] and cs.nnz == se["nnzC"]
    got = _download(ctx, cs)
    for k in KEYS:
        a, b = got[k], exp[k]
        if a.dtype.kind == "f":
            vt = np.uint64 if a.dtype == np.float64 else np.uint32
            a, b = a.view(vt), b.view(vt)
        assert np.array_equal(a, b), f"{k} differs at {np.nonzero(a != b)[0][:5]}"
    Ce.free()
    _release(ctx, bufs)
    return tA, tB, se


def test_noalloc_equals_spgemm_on_calls_in_a_row(ctx):
    cases = [("lap2d_300x300", g.gen_laplace2d(300, 300), False),
             ("er_4000x9", g.gen_er(4000, 9, seed=5), False),      # smaller than the first
             ("rmat11_f32", g.gen_rmat(11, 16, seed=3, dtype=g.F64), True)]
    for name, A, f32 in cases:
        print(f"[case] noalloc {name} f32={f32}", flush=True)
        _run(ctx, A, f32)


def test_noalloc_capacity_errors(ctx):
    A = g.gen_er(2000, 6, seed=9)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    _, se = ctx.spgemm(tA, tB)
    print(f"[case] noalloc overflow tiles {se['numtileC']} nnz {se['nnzC']}", flush=True)
    small = _buffers(ctx, tA.s.tilem, 10, 10, False)
    with pytest.raises(g.TsgError) as e:
        ctx.spgemm_noalloc(tA, tB, small, 10, 10)
    assert e.value.code == -4 and e.value.stats["numtileC"] == se["numtileC"]
    tiles_ok = _buffers(ctx, tA.s.tilem, se["numtileC"], 10, False)
    with pytest.raises(g.TsgError) as e:
        ctx.spgemm_noalloc(tA, tB, tiles_ok, se["numtileC"], 10)
    assert e.value.code == -4 and e.value.stats["nnzC"] == se["nnzC"]
    # exact capacities work after the failures
    exact = _buffers(ctx, tA.s.tilem, se["numtileC"], se["nnzC"], False)
    cs, _ = ctx.spgemm_noalloc(tA, tB, exact, se["numtileC"], se["nnzC"])
    assert cs.nnz == se["nnzC"]
    for b in (small, tiles_ok, exact):
        _release(ctx, b)
