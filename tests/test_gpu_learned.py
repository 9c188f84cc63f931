"""Learned scheduler on the GPU: the network's kernels against the oracle
(oracle/mlp.py), and the learned product against the exact pipeline.

  * forward pass, MSE gradient, features: GPU (FP32) vs oracle (FP64) on the
    same inputs, relative tolerance 1e-4 (gradient) / 1e-5 (forward, features);
  * one Adam step of tsg_learned_train == oracle adam_step on the oracle's
    gradient of the oracle's features and targets;
  * the learned product is bit-identical to the exact pipeline (structure and
    values), with and without the numeric re-run, FP64 and FP32;
  * what the trained network gains: held-out products whose value arrays are
    sized from the prediction (no host read between the symbolic and numeric
    steps) never need the re-run, with a capacity close to nnz(C) where the
    model-free bound (256 values a C tile) is many times larger; an untrained
    network under-predicts and takes the re-run.
"""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import mlp as M
from oracle import oracle as O
from tests.conftest import assert_tiles_equal

pytestmark = pytest.mark.gpu
CKEYS = ["tile_ptr", "tile_columnidx", "tile_rowidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
         "tile_csr_Val"]


def _train_family():
    return [
        ("er3000x6", lambda: g.gen_er(3000, 6, seed=11)),
        ("er6000x12", lambda: g.gen_er(6000, 12, seed=12)),
        ("er8000x3", lambda: g.gen_er(8000, 3, seed=13)),
        ("ervar5000", lambda: g.gen_er_var(5000, 0, 24, seed=14)),
        ("lap2d_96x80", lambda: g.gen_laplace2d(96, 80)),
        ("lap2d_160x64", lambda: g.gen_laplace2d(160, 64)),
        ("lap3d_20", lambda: g.gen_laplace3d(20, 20, 20)),
        ("st27_14", lambda: g.gen_stencil27(14, 14, 14)),
        ("rmat11", lambda: g.gen_rmat(11, 8, seed=15, dtype=g.F64)),
        ("rmat12", lambda: g.gen_rmat(12, 6, seed=16, dtype=g.F64)),
    ]


def _held_out():
    return [
        ("er5000x9", lambda: g.gen_er(5000, 9, seed=99)),
        ("lap2d_128x100", lambda: g.gen_laplace2d(128, 100)),
        ("st27_12", lambda: g.gen_stencil27(12, 12, 12)),
        ("rmat11b", lambda: g.gen_rmat(11, 8, seed=77, dtype=g.F64)),
    ]


@pytest.fixture(scope="module")
def trained():
    """A context whose network was trained on the training family."""
    c = g.Context(0)
    c.learned_clear()
    for name, mk in _train_family():
        A = mk()
        tA, tB = c.tiles_from_csr(A), c.tiles_from_csr(A, as_b=True)
        c.learned_collect(tA, tB)
        tA.free(); tB.free()
    loss = c.learned_train(epochs=400, batch=2048, lr=3e-3, seed=0)
    print(f"[learned] final loss {loss:.4f} calibration {c.learned_calibration()}", flush=True)
    yield c
    c.close()


def test_forward_matches_oracle(ctx):
    r = np.random.default_rng(0)
    w = (r.normal(0, 0.3, M.NP)).astype(np.float32)
    ctx.learned_set_weights(w)
    assert np.array_equal(ctx.learned_weights(), w)
    X = r.normal(1.0, 1.0, (1000, M.NF)).astype(np.float32)
    Y = ctx.learned_forward(X)
    Yo, _ = M.forward(w.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(Y - Yo)) <= 1e-5 * (1 + np.max(np.abs(Yo)))


@pytest.mark.parametrize("n", [77, 1000])
def test_gradient_matches_oracle(ctx, n):
    r = np.random.default_rng(n)
    w = (r.normal(0, 0.3, M.NP)).astype(np.float32)
    ctx.learned_set_weights(w)
    X = r.normal(1.0, 1.0, (n, M.NF)).astype(np.float32)
    T = r.normal(2.0, 1.0, (n, M.NO)).astype(np.float32)
    gg, lg = ctx.learned_gradient(X, T)
    go, lo = M.gradient(w.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    err = np.max(np.abs(gg - go)) / np.max(np.abs(go))
    print(f"    gradient n={n}: max rel err {err:.2e}, loss {lg:.6f} vs {lo:.6f}", flush=True)
    assert err <= 1e-4
    assert abs(lg - lo) <= 1e-5 * lo


@pytest.mark.parametrize("name", ["er4096", "lap2d", "rmat10", "ervar", "st27"])
def test_features_match_oracle(ctx, name):
    A = {"er4096": lambda: g.gen_er(4096, 8, seed=0), "lap2d": lambda: g.gen_laplace2d(70, 33),
         "rmat10": lambda: g.gen_rmat(10, 16, seed=1, dtype=g.F64), "ervar": lambda: g.gen_er_var(999, 0, 40, seed=2),
         "st27": lambda: g.gen_stencil27(9, 9, 9)}[name]()
    B = A if name != "ervar" else g.gen_er(999, 5, seed=3)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(B, as_b=True)
    Xg = ctx.learned_features(tA, tB)
    Xo = M.features(A, B)
    assert Xg.shape == Xo.shape
    assert np.allclose(Xg, Xo, rtol=1e-5, atol=1e-6)
    # integer features recovered exactly where float32 log1p is exact enough
    assert np.array_equal(np.rint(np.expm1(Xg[:, 0].astype(np.float64))), M.feature_counts(A, B)[:, 0])


def test_adam_step_matches_oracle(ctx):
    A = g.gen_er(2000, 5, seed=21)
    tA, tB = ctx.tiles_from_csr(A), ctx.tiles_from_csr(A, as_b=True)
    r = np.random.default_rng(5)
    w0 = (r.normal(0, 0.2, M.NP)).astype(np.float32)
    ctx.learned_clear()
    ctx.learned_collect(tA, tB)
    ctx.learned_set_weights(w0)
    lr = 1e-3
    ctx.learned_train(epochs=1, batch=1 << 20, lr=lr, seed=0)
    w1 = ctx.learned_weights()
    X = M.features(A, A).astype(np.float64)
    T = M.targets(O.spgemm_tiled(A, A)).astype(np.float64)
    go, _ = M.gradient(w0.astype(np.float64), X, T)
    we, _, _ = M.adam_step(w0.astype(np.float64), 0.0, 0.0, go, 1, lr)
    big = np.abs(go) > 1e-3 * np.max(np.abs(go))
    assert np.max(np.abs(w1 - we)[big]) <= 1e-3 * lr + 1e-7
    assert np.all(np.abs(w1 - w0)[~big] <= lr * 1.001)
    ctx.learned_clear()


def test_learned_without_model_is_an_error():
    c = g.Context(0)
    try:
        A = g.gen_er(200, 3, seed=1)
        tA, tB = c.tiles_from_csr(A), c.tiles_from_csr(A, as_b=True)
        with pytest.raises(g.TsgError):
            c.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
        with pytest.raises(g.TsgError):
            c.learned_train(epochs=1)            # nothing collected
    finally:
        c.close()


def _pair(c, A, f32=False):
    h = A.astype(g.F32) if f32 else A
    return c.tiles_from_csr(h), c.tiles_from_csr(h, as_b=True)


@pytest.mark.parametrize("name", [n for n, _ in _held_out()])
def test_learned_equals_exact_on_held_out(trained, name):
    A = dict(_held_out())[name]()
    tA, tB = _pair(trained, A)
    print(f"[case] learned held-out {name}", flush=True)
    Ce, se = trained.spgemm(tA, tB)
    Cl, sl = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    e, l = Ce.download(), Cl.download()
    assert_tiles_equal(l, e, keys=CKEYS, label="C ")
    print(f"    fallback={sl['fallback_used']} syncs {sl['host_syncs']} vs {se['host_syncs']} "
          f"predicted tiles/tiles {sl['pred_numtileC'] / max(1, se['numtileC']):.3f} "
          f"values capacity/nnz {sl['pred_nnzC'] / max(1, se['nnzC']):.3f}", flush=True)
    assert sl["fallback_used"] == 0 and sl["pred_nnzC"] >= se["nnzC"]
    assert sl["host_syncs"] == 2 and se["host_syncs"] == 2


def test_learned_gain_over_untrained_and_upper_bound(trained):
    """Value capacity the trained network allocates vs nnz(C) and vs the
    model-free exact bound (256 values a C tile), and an untrained network."""
    rows = []
    for name, mk in _held_out():
        A = mk()
        tA, tB = _pair(trained, A)
        _, se = trained.spgemm(tA, tB)
        _, sl = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
        upper = 256 * se["numtileC"]
        rows.append((name, se["nnzC"], sl["pred_nnzC"], upper, sl["fallback_used"]))
        print(f"    {name}: nnz {se['nnzC']} learned capacity {sl['pred_nnzC']} model-free bound {upper} "
              f"fallback {sl['fallback_used']}", flush=True)
    assert all(r[4] == 0 for r in rows)
    learned = sum(r[2] for r in rows)
    exact = sum(r[1] for r in rows)
    upper = sum(r[3] for r in rows)
    assert learned <= 1.5 * exact + 1024 * len(rows)
    assert 4 * learned < upper
    # an untrained network (all weights zero, no calibration margin) predicts
    # nothing: the numeric step re-runs into exact arrays, same result
    c = g.Context(0)
    try:
        c.learned_set_weights(np.zeros(M.NP, np.float32))
        c.learned_calibration([0.0, 0.0])
        A = _held_out()[0][1]()
        tA, tB = _pair(c, A)
        Cu, su = c.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
        Ce, _ = c.spgemm(tA, tB)
        assert_tiles_equal(Cu.download(), Ce.download(), keys=CKEYS, label="C ")
        print(f"    untrained: fallback {su['fallback_used']} capacity {su['pred_nnzC']} nnz {su['nnzC']}",
              flush=True)
        assert su["fallback_used"] == 1
    finally:
        c.close()


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("scale", [0.01, 0.6])
@pytest.mark.parametrize("name", ["er4000x10", "rmat12"])
def test_forced_fallback_is_exact(trained, f32, scale, name):
    """A capacity cut to `scale` of the prediction: the numeric step re-runs
    only for tile-rows (tiles) whose values end past it, after copying the
    values written below it; the result is the exact pipeline's."""
    A = g.gen_er(4000, 10, seed=31) if name == "er4000x10" else g.gen_rmat(12, 16, seed=6, dtype=g.F64)
    tA, tB = _pair(trained, A, f32)
    Ce, se = trained.spgemm(tA, tB)
    Cl, sl = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED, pred_scale=scale)
    print(f"    {name} scale {scale}: capacity {sl['pred_nnzC']} of nnz {se['nnzC']}", flush=True)
    assert sl["fallback_used"] == 1 and sl["pred_nnzC"] < se["nnzC"]
    assert_tiles_equal(Cl.download(), Ce.download(), keys=CKEYS, label="C ")


@pytest.mark.parametrize("f32", [False, True])
def test_learned_matches_oracle(trained, f32):
    A = g.gen_rmat(10, 16, seed=4, dtype=g.F64)
    tA, tB = _pair(trained, A, f32)
    Cl, sl = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    got = Cl.download()
    exp = O.spgemm_tiled(A, A, f32=f32)
    assert_tiles_equal(got, exp, keys=CKEYS, label="C ")
    chk = O.gustavson_check(A, A, got)
    assert chk["missing"] == 0 and chk["extra"] == 0
    assert chk["max_scaled"] <= (1e-5 if f32 else 1e-12)


def test_learned_stays_exact_when_operands_change(trained):
    """Operands freed and re-created between calls (the pool hands the same
    addresses back): every result still equals the exact pipeline's."""
    for seed in range(4):
        A = g.gen_er(3000, 7, seed=100 + seed)
        tA, tB = _pair(trained, A)
        Ce, _ = trained.spgemm(tA, tB)
        for _ in range(2):
            Cl, sl = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
            assert_tiles_equal(Cl.download(), Ce.download(), keys=CKEYS, label=f"seed {seed} ")
            Cl.free()
        Ce.free(); tA.free(); tB.free()


def test_learned_duplicates_and_empty(trained):
    rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9], dtype=np.int64)
    col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6], dtype=np.int32)
    A = g.HostCSR(8, 8, rp, col, np.arange(1.0, 10.0))
    tA, tB = _pair(trained, A)
    Cl, _ = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    exp = O.spgemm_tiled(A, A)
    assert_tiles_equal(Cl.download(), exp, keys=["tile_ptr", "tile_columnidx", "tile_nnz", "mask",
                                                 "tile_csr_Col", "tile_csr_Val"], label="C ")
    Z = g.HostCSR(40, 40, np.zeros(41, np.int64), np.zeros(0, np.int32), np.zeros(0))
    tA, tB = _pair(trained, Z)
    Cz, sz = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    d = Cz.download()
    assert d["numtile"] == 0 and d["nnz"] == 0 and np.all(d["tile_ptr"] == 0)


def test_calibration_matches_oracle(trained):
    """The offsets training left equal the oracle's calibrate on the oracle's
    features and targets of the training family, with the trained weights."""
    w = trained.learned_weights().astype(np.float64)
    Xs, Ts, Fs, ends = [], [], [], []
    for name, mk in _train_family():
        A = mk()
        Xs.append(M.features(A, A))
        Fs.append(M.feature_counts(A, A))
        Ts.append(M.targets(O.spgemm_tiled(A, A)))
        ends.append(sum(len(x) for x in Xs))
    X, T, F = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Fs)
    Y, _ = M.forward(w, X.astype(np.float64))
    off = M.calibrate(Y, T, F, ends)
    got = trained.learned_calibration()
    print(f"    calibration GPU {got} oracle {off}", flush=True)
    assert np.allclose(got, off, atol=1e-4)


@pytest.mark.parametrize("name", ["lap2d_300x200", "er4000x10", "rmat12"])
def test_learned_repeat_has_one_read(trained, name):
    """A repeated product of the same operands: C is sized from the remembered
    capacities (the larger of the predictions and the sizes the first call
    found), every step is queued with device-side sizes and the call makes one
    host read, where the first makes two; both equal the exact pipeline's C.
    pred_scale 0.5 cuts those capacities: the repeat overflows on the device
    and the call re-runs exactly (fallback 1), with the same result."""
    A = {"lap2d_300x200": lambda: g.gen_laplace2d(300, 200), "er4000x10": lambda: g.gen_er(4000, 10, seed=31),
         "rmat12": lambda: g.gen_rmat(12, 16, seed=6, dtype=g.F64)}[name]()
    tA, tB = _pair(trained, A)
    Ce, se = trained.spgemm(tA, tB)
    e = Ce.download()
    C1, s1 = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    C2, s2 = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    C3, s3 = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)
    print(f"[case] learned repeat {name}: syncs {s1['host_syncs']} {s2['host_syncs']} {s3['host_syncs']} "
          f"fallback {s1['fallback_used']} {s2['fallback_used']} {s3['fallback_used']}", flush=True)
    for C, s in ((C1, s1), (C2, s2), (C3, s3)):
        assert s["numtileC"] == se["numtileC"] and s["nnzC"] == se["nnzC"]
        assert_tiles_equal(C.download(), e, keys=CKEYS, label=f"{name} ")
    assert s2["host_syncs"] == 1 and s3["host_syncs"] == 1 and s2["fallback_used"] == 0
    # capacities cut to half: the device flags the overflow, the call re-runs exactly
    Ca, sa = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED, pred_scale=0.5)
    Cb, sb = trained.spgemm(tA, tB, scheduler=g.SCHED_LEARNED, pred_scale=0.5)
    assert sb["fallback_used"] == 1 and sb["host_syncs"] >= 3
    for C in (Ca, Cb):
        assert_tiles_equal(C.download(), e, keys=CKEYS, label=f"{name} cut ")
