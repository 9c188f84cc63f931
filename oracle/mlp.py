"""Oracle of the learned scheduler: the network (forward pass, mean squared
error, its gradient, one Adam step) and the per-tile-row features and targets,
in numpy float64.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The reference has no learned scheduler (it sizes C with exact count kernels
and host reads, src/spgemm_raw_v2.cu:119-141 and :191-211), so there is no
reference code to follow; what fixes each function is stated beside it and
checked in tests/test_oracle_mlp.py:
  * gradient   -- central finite differences of loss() in float64, and closed
                  forms (zero hidden layers: only b3 moves, by 2*mean(y-t));
  * forward    -- hand-computed single-path networks;
  * adam_step  -- the first step of Adam is lr*sign(g) (bias-corrected moments
                  cancel the (1-beta) factors), and a zero gradient moves nothing;
  * features   -- brute-force set loops over the nonzeros on small matrices,
                  and closed forms (identity, dense);
  * targets    -- tile and nonzero counts of the tiled oracle's C;
  * calibrate  -- its defining property: every training product's capacity
                  expm1(prediction + offset) covers the exact count.

Parameter layout (row-major, x @ W): W1[8][32], b1[32], W2[32][32], b2[32],
W3[32][2], b3[2]; outputs are log1p of (C tiles of the tile-row, C nonzeros
of the tile-row).
"""
import numpy as np

NF, NH, NO = 8, 32, 2
P_W1 = 0
P_B1 = P_W1 + NF * NH
P_W2 = P_B1 + NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NH * NO
NP = P_B3 + NO
OFF_LO, OFF_HI = np.log(1.02), np.log(16.0)
CAP_EXTRA = 1024


def unpack(w):
    w = np.asarray(w, np.float64)
    assert w.size == NP
    return (w[P_W1:P_B1].reshape(NF, NH), w[P_B1:P_W2], w[P_W2:P_B2].reshape(NH, NH), w[P_B2:P_W3],
            w[P_W3:P_B3].reshape(NH, NO), w[P_B3:NP])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(np.float64)


def forward(w, X):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.asarray(X, np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ W3 + b3, (X, z1, h1, z2, h2)


def loss(w, X, T):
    """mean over samples of sum over outputs of (y - t)^2"""
    Y, _ = forward(w, X)
    return float(np.sum((Y - np.asarray(T, np.float64)) ** 2) / len(X))


def gradient(w, X, T):
    """d loss / d w by backpropagation; returns (gradient, loss)."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    Y, (X, z1, h1, z2, h2) = forward(w, X)
    n = len(X)
    E = Y - np.asarray(T, np.float64)
    d3 = 2.0 * E / n
    gW3 = h2.T @ d3
    gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * (z2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    return pack(gW1, gb1, gW2, gb2, gW3, gb3), float(np.sum(E * E) / n)


def adam_step(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def glorot_scale():
    """Glorot-uniform bound of each weight matrix."""
    return {"W1": np.sqrt(6.0 / (NF + NH)), "W2": np.sqrt(6.0 / (NH + NH)), "W3": np.sqrt(6.0 / (NH + NO))}


def feature_counts(A, B):
    """Integer features of every tile-row of A*B (tilem x 8, int64):
    A tiles, B tiles reached (over A's tiles, with multiplicity), A nonzeros,
    B nonzeros reached, span of the B tile columns reached (last - first + 1),
    longest B tile-row reached, summed spans of the B tile-rows reached,
    nonempty B tile-rows reached."""
    tilem = (A.m + 15) // 16
    tileb = (B.m + 15) // 16
    # B tile-rows: tile count, nonzeros, first and last tile column
    brow = np.repeat(np.arange(B.m, dtype=np.int64), np.diff(B.rowptr))
    bkey = np.unique((brow // 16) * (1 << 31) + np.asarray(B.col, np.int64) // 16)
    bk, bj = bkey >> 31, bkey & ((1 << 31) - 1)
    nbk = np.bincount(bk, minlength=tileb).astype(np.int64)
    pbk = np.bincount(brow // 16, minlength=tileb).astype(np.int64) if B.nnz else np.zeros(tileb, np.int64)
    first = np.full(tileb, np.iinfo(np.int64).max, np.int64)
    last = np.full(tileb, -1, np.int64)
    np.minimum.at(first, bk, bj)
    np.maximum.at(last, bk, bj)
    # A tiles
    arow = np.repeat(np.arange(A.m, dtype=np.int64), np.diff(A.rowptr))
    akey = np.unique((arow // 16) * (1 << 31) + np.asarray(A.col, np.int64) // 16)
    ai, ak = akey >> 31, akey & ((1 << 31) - 1)
    F = np.zeros((tilem, 8), np.int64)
    F[:, 0] = np.bincount(ai, minlength=tilem)
    np.add.at(F[:, 1], ai, nbk[ak])
    F[:, 2] = np.bincount(arow // 16, minlength=tilem) if A.nnz else 0
    np.add.at(F[:, 3], ai, pbk[ak])
    ne = nbk[ak] > 0
    lo = np.full(tilem, np.iinfo(np.int64).max, np.int64)
    hi = np.full(tilem, -1, np.int64)
    np.minimum.at(lo, ai[ne], first[ak[ne]])
    np.maximum.at(hi, ai[ne], last[ak[ne]])
    F[:, 4] = np.where(hi >= lo, hi - lo + 1, 0)
    np.maximum.at(F[:, 5], ai, nbk[ak])
    np.add.at(F[:, 6], ai[ne], last[ak[ne]] - first[ak[ne]] + 1)
    np.add.at(F[:, 7], ai, ne.astype(np.int64))
    return F


def features(A, B):
    """log1p of feature_counts, float32 (the network's input)."""
    return np.log1p(feature_counts(A, B).astype(np.float64)).astype(np.float32)


def targets(Cd):
    """Targets of every tile-row from a C tile dict: log1p(tiles), log1p(nonzeros)."""
    tp = np.asarray(Cd["tile_ptr"], np.int64)
    tn = np.asarray(Cd["tile_nnz"], np.int64)
    nc = np.diff(tp)
    nz = tn[tp[1:]] - tn[tp[:-1]]
    return np.log1p(np.stack([nc, nz], 1).astype(np.float64)).astype(np.float32)


def row_predictions(F, Y, count):
    """Per-row predictions as the device sums them (k_learned_cap): raw
    outputs Y -> ceil(max(expm1(y), 0)), tiles clamped to min(B tiles reached,
    tile-column span), nonzeros to 256 per exact C tile (count)."""
    et = np.expm1(np.minimum(np.asarray(Y, np.float64), 40.0))
    pt = np.minimum(np.ceil(np.maximum(et[:, 0], 0)), np.minimum(F[:, 1], F[:, 4]))
    pn = np.minimum(np.ceil(np.maximum(et[:, 1], 0)), 256.0 * np.asarray(count, np.float64))
    return pt, pn


def calibrate(Y, T, F, mat_end):
    """Log-space offsets (tiles, nonzeros): log 1.02 + the largest log(actual /
    predicted total) over the products whose rows end at mat_end, clamped to
    [log 1.02, log 16]; totals from row_predictions and the targets T."""
    cnt = np.rint(np.expm1(np.asarray(T, np.float64)))
    pt, pn = row_predictions(F, Y, cnt[:, 0])
    off = np.full(2, -np.inf)
    r0 = 0
    for r1 in mat_end:
        for o, pred in ((0, pt), (1, pn)):
            act = cnt[r0:r1, o].sum()
            if act > 0:
                off[o] = max(off[o], np.log(act / max(pred[r0:r1].sum(), 1.0)))
        r0 = r1
    return np.clip(off + OFF_LO, OFF_LO, OFF_HI)


def value_capacity(pred_nnz_sum, off_nnz, numtileC, scale=1.0):
    """C's value capacity from the summed row predictions: the calibrated sum
    (+ CAP_EXTRA unless scaled down), never above 256 values a C tile."""
    want = np.ceil(pred_nnz_sum * np.exp(off_nnz) * scale) + (CAP_EXTRA if scale >= 1.0 else 0)
    return int(min(want, 256 * numtileC))
