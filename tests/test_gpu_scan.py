"""The device-wide int64 scan (tsg_scan_i64, the library's replacement for the
reference's CUB ExclusiveSum, src/spgemm_raw_v2.cu:131-138) against numpy's
cumsum, at the sizes where its strategy changes: one block (n <= 2048), blocks
that add up the raw partials before them (up to 2048 blocks), and the
recursive scan of the partials beyond that; with data[0] taken as given or as
zero (zero_first), and the total written to a separate device word.  The
context runs on torch's current stream (tsg_set_stream), so torch's copies
and the library's kernels are ordered on one stream."""
import numpy as np
import pytest

import gpu_spgemm_amd as g

pytestmark = pytest.mark.gpu
T = 2048  # items per scan block


@pytest.mark.parametrize("n", [1, 2, T, T + 1, T * T, T * T + 1, (T + 1) * T + 7])
@pytest.mark.parametrize("zero_first", [False, True])
def test_scan_matches_cumsum(ctx, n, zero_first):
    import torch
    print(f"[case] scan n={n} zero_first={zero_first}", flush=True)
    rng = np.random.default_rng(n + (7 if zero_first else 0))
    x = rng.integers(-(1 << 20), 1 << 40, n, dtype=np.int64)
    exp = x.copy()
    if zero_first:
        exp[0] = 0
    exp = np.cumsum(exp)
    s = torch.cuda.current_stream()
    ctx.set_stream(s.cuda_stream)
    try:
        d = torch.from_numpy(x).to("cuda")
        tot = torch.full((2,), -1, dtype=torch.int64, device="cuda")
        ctx.scan_i64(d.data_ptr(), n, zero_first, tot.data_ptr())
        got = d.cpu().numpy()
        t = tot.cpu().numpy()
    finally:
        ctx.synchronize()
        ctx.set_stream(None)
    assert np.array_equal(got, exp), f"first mismatch at {np.nonzero(got != exp)[0][:5]}"
    assert t[0] == exp[-1] and t[1] == -1


def test_set_stream_round_trip(ctx):
    import torch
    own = ctx.stream()
    s = torch.cuda.Stream()
    ctx.set_stream(s.cuda_stream)
    assert ctx.stream() == s.cuda_stream
    A = g.gen_er(1500, 5, seed=3)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    C1, _ = ctx.spgemm(tA, tB)
    d1 = C1.download()
    ctx.set_stream(None)
    assert ctx.stream() == own
    C2, _ = ctx.spgemm(tA, tB)
    d2 = C2.download()
    for k in ("tile_ptr", "tile_columnidx", "tile_nnz", "mask", "tile_csr_Col"):
        assert np.array_equal(d1[k], d2[k])
    assert np.array_equal(d1["tile_csr_Val"].view(np.uint64), d2["tile_csr_Val"].view(np.uint64))
