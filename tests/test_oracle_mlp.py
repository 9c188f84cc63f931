"""Pins the learned scheduler's oracle (oracle/mlp.py) on the CPU: gradient by
finite differences and closed forms, Adam by its first-step closed form,
features by brute-force set loops and closed forms, targets by the tiled
oracle.  A planted mistake in each function must fail one of these."""
import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import mlp as M
from oracle import oracle as O


def rand_w(seed, scale=0.5):
    return np.random.default_rng(seed).normal(0, scale, M.NP)


def rand_xt(seed, n=37):
    r = np.random.default_rng(seed)
    return r.normal(0, 1, (n, M.NF)), r.normal(0, 1, (n, M.NO))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_finite_differences(seed):
    w = rand_w(seed)
    X, T = rand_xt(seed + 10)
    gr, l0 = M.gradient(w, X, T)
    assert abs(l0 - M.loss(w, X, T)) < 1e-12
    h = 1e-6
    fd = np.empty(M.NP)
    for p in range(M.NP):
        e = np.zeros(M.NP)
        e[p] = h
        fd[p] = (M.loss(w + e, X, T) - M.loss(w - e, X, T)) / (2 * h)
    assert np.max(np.abs(fd - gr)) <= 1e-6 * (1 + np.max(np.abs(gr)))


def test_gradient_closed_form_with_dead_hidden_layers():
    # every hidden unit off: y = b3, only b3 gets a gradient, 2*mean(b3 - t)
    w = np.zeros(M.NP)
    w[M.P_B1:M.P_W2] = -1.0
    w[M.P_B3:] = [0.5, -2.0]
    X, T = rand_xt(3)
    Y, _ = M.forward(w, X)
    assert np.allclose(Y, [0.5, -2.0])
    gr, _ = M.gradient(w, X, T)
    expect = np.zeros(M.NP)
    expect[M.P_B3:] = 2.0 * np.mean(np.array([0.5, -2.0]) - T, axis=0)
    assert np.allclose(gr, expect, atol=1e-14)


def test_forward_single_path():
    # one path x0 -> h1[3] -> h2[5] -> y1: y1 = b3[1] + w3 * relu(w2 * relu(w1*x0 + b1) + b2)
    W1 = np.zeros((M.NF, M.NH)); W1[0, 3] = 2.0
    b1 = np.zeros(M.NH); b1[3] = -1.0
    W2 = np.zeros((M.NH, M.NH)); W2[3, 5] = -3.0
    b2 = np.zeros(M.NH); b2[5] = 10.0
    W3 = np.zeros((M.NH, M.NO)); W3[5, 1] = 0.25
    b3 = np.array([7.0, 1.0])
    w = M.pack(W1, b1, W2, b2, W3, b3)
    X = np.zeros((3, M.NF)); X[:, 0] = [0.0, 2.0, 10.0]
    Y, _ = M.forward(w, X)
    h1 = np.maximum(2.0 * X[:, 0] - 1.0, 0)
    h2 = np.maximum(-3.0 * h1 + 10.0, 0)
    assert np.allclose(Y[:, 0], 7.0)
    assert np.allclose(Y[:, 1], 1.0 + 0.25 * h2)
    assert np.allclose(Y[:, 1], [3.5, 1.25, 1.0])


def test_adam_first_step_is_sign_step():
    w = rand_w(4)
    gr = np.random.default_rng(5).normal(0, 1, M.NP)
    gr[::7] = 0.0
    w1, m, v = M.adam_step(w, np.zeros(M.NP), np.zeros(M.NP), gr, 1, 0.01)
    step = w - w1
    nz = gr != 0
    assert np.allclose(step[nz], 0.01 * np.sign(gr[nz]), rtol=1e-4)  # eps=1e-8 against |g|
    assert np.all(step[~nz] == 0)
    assert np.allclose(m, 0.1 * gr) and np.allclose(v, 0.001 * gr * gr)
    # second step with the same gradient is again lr*sign(g)
    w2, _, _ = M.adam_step(w1, m, v, gr, 2, 0.01)
    assert np.allclose((w1 - w2)[nz], 0.01 * np.sign(gr[nz]), rtol=1e-4)


def brute_features(A, B):
    """Set loops over the nonzeros (a different algorithm from the oracle's)."""
    tilem = (A.m + 15) // 16
    brows = {}
    for i in range(B.m):
        for p in range(B.rowptr[i], B.rowptr[i + 1]):
            brows.setdefault(i // 16, {"tiles": set(), "nnz": 0})
            brows[i // 16]["tiles"].add(int(B.col[p]) // 16)
            brows[i // 16]["nnz"] += 1
    out = []
    for bi in range(tilem):
        ks, pa = set(), 0
        for i in range(16 * bi, min(16 * bi + 16, A.m)):
            for p in range(A.rowptr[i], A.rowptr[i + 1]):
                ks.add(int(A.col[p]) // 16)
                pa += 1
        nb = pb = sumspan = nonempty = maxlen = 0
        cols = set()
        for k in ks:
            r = brows.get(k)
            if not r:
                continue
            nb += len(r["tiles"])
            pb += r["nnz"]
            cols |= r["tiles"]
            sumspan += max(r["tiles"]) - min(r["tiles"]) + 1
            nonempty += 1
            maxlen = max(maxlen, len(r["tiles"]))
        span = max(cols) - min(cols) + 1 if cols else 0
        out.append([len(ks), nb, pa, pb, span, maxlen, sumspan, nonempty])
    return np.array(out, np.int64).reshape(tilem, 8)


@pytest.mark.parametrize("mk", [
    lambda: (g.gen_er(300, 5, seed=1), g.gen_er(300, 5, seed=2)),
    lambda: (g.gen_laplace2d(20, 17), g.gen_laplace2d(20, 17)),
    lambda: (g.gen_er_var(250, 0, 30, seed=3), g.gen_er(250, 3, seed=4)),
    lambda: (g.gen_rmat(8, 6, seed=2, dtype=g.F64), g.gen_rmat(8, 6, seed=3, dtype=g.F64)),
])
def test_features_match_brute_force(mk):
    A, B = mk()
    assert np.array_equal(M.feature_counts(A, B), brute_features(A, B))


def test_features_closed_forms():
    n = 160                      # 10 x 10 tiles
    I = g.HostCSR(n, n, np.arange(n + 1), np.arange(n), np.ones(n))
    F = M.feature_counts(I, I)
    assert np.all(F == [1, 1, 16, 16, 1, 1, 1, 1])
    rp = np.arange(0, n * n + 1, n)
    D = g.HostCSR(n, n, rp, np.tile(np.arange(n), n), np.ones(n * n))
    F = M.feature_counts(D, D)
    assert np.all(F == [10, 100, 16 * n, 10 * 16 * n, 10, 10, 100, 10])
    Z = g.HostCSR(n, n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
    assert np.all(M.feature_counts(Z, D) == 0) and np.all(M.feature_counts(D, Z)[:, 0] == 10)
    assert np.all(M.feature_counts(D, Z)[:, [1, 3, 4, 5, 6, 7]] == 0)


def test_targets_from_tiled_oracle():
    A = g.gen_er(400, 6, seed=7)
    Cd = O.spgemm_tiled(A, A)
    T = np.expm1(M.targets(Cd).astype(np.float64))
    assert T.shape == ((A.m + 15) // 16, 2)
    assert np.isclose(T[:, 0].sum(), Cd["numtile"], rtol=1e-6) and np.isclose(T[:, 1].sum(), Cd["nnz"], rtol=1e-6)
    # per tile-row nonzeros from the Gustavson product; tiles from its distinct tile columns
    rp, col, _ = O.gustavson_csr(A, A)
    rows = np.add.reduceat(np.diff(rp), np.arange(0, A.m, 16))
    assert np.allclose(T[:, 1], rows, rtol=1e-6)
    r = np.repeat(np.arange(A.m), np.diff(rp))
    tiles = np.bincount(np.unique((r // 16) * 10**6 + col // 16) // 10**6, minlength=len(T))
    assert np.allclose(T[:, 0], tiles, rtol=1e-6)


def test_calibrated_capacity_covers_every_training_product():
    """calibrate's defining property: the value capacity of every product it
    saw covers that product's nnz; capacities never exceed 256 a C tile; the
    forced-rerun scale makes one fall short; a perfect model gets the margin."""
    r = np.random.default_rng(11)
    nrow = 600
    F = r.integers(1, 400, (nrow, 8))
    count = np.minimum(r.integers(1, 300, nrow), np.minimum(F[:, 1], F[:, 4]))
    nnz = count * r.integers(1, 200, nrow)
    T = np.log1p(np.stack([count, nnz], 1).astype(np.float64))
    Y = T + r.normal(-0.1, 0.3, T.shape)        # a biased model with 30 % log error
    mat_end = [100, 250, 400, 600]              # four products
    off = M.calibrate(Y, T, F, mat_end)
    _, pn = M.row_predictions(F, Y, count)
    r0 = 0
    for r1 in mat_end:
        tiles = int(count[r0:r1].sum())
        cap = M.value_capacity(pn[r0:r1].sum(), off[1], tiles)
        assert cap >= nnz[r0:r1].sum() and cap <= 256 * tiles
        r0 = r1
    # shrinking (the forced re-run hook) makes some product too small
    short = [M.value_capacity(pn[a:b].sum(), off[1], int(count[a:b].sum()), scale=0.01) < nnz[a:b].sum()
             for a, b in zip([0] + mat_end[:-1], mat_end)]
    assert any(short)
    # a perfect model gets the minimum offset (its rounded-up rows cover the totals)
    assert np.allclose(M.calibrate(T, T, F, mat_end), M.OFF_LO)
    # a planted mistake -- offsets from the best instead of the worst product -- leaves one short
    cnt = np.rint(np.expm1(T))
    best = min(np.log(cnt[a:b, 1].sum() / pn[a:b].sum()) for a, b in zip([0] + mat_end[:-1], mat_end))
    assert any(M.value_capacity(pn[a:b].sum(), best, int(count[a:b].sum())) - M.CAP_EXTRA < nnz[a:b].sum()
               for a, b in zip([0] + mat_end[:-1], mat_end))
