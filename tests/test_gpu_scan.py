"""The device-wide int64 scan (tsg_scan_i64, the library's replacement for the
reference's CUB ExclusiveSum, src/spgemm_raw_v2.cu:131-138) against numpy's
cumsum, at the sizes where its strategy changes: one block (n <= 2048), blocks
that add up the raw partials before them (up to 2048 blocks), and the
recursive scan of the partials beyond that; with data[0] taken as given or as
zero (zero_first), and the total written to a separate device word.  Also:
a context moved onto a caller's stream (tsg_set_stream) computes the same
product as on its own stream."""
import ctypes as C

import numpy as np
import pytest

import gpu_spgemm_amd as g

pytestmark = pytest.mark.gpu
T = 2048  # items per scan block


@pytest.mark.parametrize("n", [1, 2, T, T + 1, T * T, T * T + 1, (T + 1) * T + 7])
@pytest.mark.parametrize("zero_first", [False, True])
def test_scan_matches_cumsum(ctx, n, zero_first):
    print(f"[case] scan n={n} zero_first={zero_first}", flush=True)
    rng = np.random.default_rng(n + (7 if zero_first else 0))
    x = rng.integers(-(1 << 20), 1 << 40, n, dtype=np.int64)
    exp = x.copy()
    if zero_first:
        exp[0] = 0
    exp = np.cumsum(exp)
    d = ctx.to_device(x)
    tot = ctx.to_device(np.full(2, -1, np.int64))
    try:
        ctx.scan_i64(d, n, zero_first, tot)
        got = ctx.to_host(d, n, np.int64)
        t = ctx.to_host(tot, 2, np.int64)
    finally:
        ctx.free(d)
        ctx.free(tot)
    assert np.array_equal(got, exp), f"first mismatch at {np.nonzero(got != exp)[0][:5]}"
    assert t[0] == exp[-1] and t[1] == -1


def test_set_stream_round_trip(ctx):
    hip = C.CDLL("libamdhip64.so.7")   # the runtime libtsg is linked against
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    own = ctx.stream()
    try:
        ctx.set_stream(s.value)
        assert ctx.stream() == s.value
        A = g.gen_er(1500, 5, seed=3)
        tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
        C1, _ = ctx.spgemm(tA, tB)
        d1 = C1.download()
        ctx.set_stream(None)
        assert ctx.stream() == own
        C2, _ = ctx.spgemm(tA, tB)
        d2 = C2.download()
    finally:
        ctx.set_stream(None)
        hip.hipStreamDestroy(s)
    for k in ("tile_ptr", "tile_columnidx", "tile_nnz", "mask", "tile_csr_Col"):
        assert np.array_equal(d1[k], d2[k])
    assert np.array_equal(d1["tile_csr_Val"].view(np.uint64), d2["tile_csr_Val"].view(np.uint64))
