"""Pins the oracle (CPU restatement of the reference path) before it is trusted.

- conversion (csr2tile / build_csc_tiles / load_mtx, the within-tile CSC of
  build_csc_tiles and build_a_tile_csc) against golden fixtures
  written by the reference's own src/tile_format.h (tests/golden/), and live
  against oracle/_ref when it is built;
- Step 1 against the closed form: the C tile set is the boolean product of
  the tile patterns of A and B;
- Step 2 against the closed form: maskC is the element pattern of the boolean
  product A_pattern * B_pattern, tile_nnzC its per-tile popcount;
- Step 3 against brute-force dense products and the Gustavson recompute, and
  bit-exactly on integer stencils;
- planted mistakes (a transposed operand, a dropped pair, a flipped mask bit)
  must make these checks fail.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from tests.conftest import assert_csc_equal, assert_tiles_equal

GOLD = os.path.join(os.path.dirname(__file__), "golden", "tile_format_golden.npz")
CASES = ["er200", "lap20x13", "rect77x300", "dense40", "sparse50"]


class CSR:
    def __init__(self, m, n, rowptr, col, val):
        self.m, self.n = int(m), int(n)
        self.rowptr = np.asarray(rowptr, dtype=np.int64)
        self.col = np.asarray(col, dtype=np.int32)
        self.val = np.asarray(val, dtype=np.float64)

    def to_scipy(self):
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.m, self.n))


def from_scipy(S):
    S = sp.csr_matrix(S)
    S.sort_indices()
    return CSR(S.shape[0], S.shape[1], S.indptr, S.indices, S.data)


def golden():
    return np.load(GOLD, allow_pickle=False)


def gold_case(z, name):
    A = CSR(z[f"{name}/m"], z[f"{name}/n"], z[f"{name}/rowptr"], z[f"{name}/col"], z[f"{name}/val"])
    exp = {}
    for tag in ("A", "B"):
        d = {}
        for k in z.files:
            if k.startswith(f"{name}/{tag}/"):
                v = z[k]
                d[k.split("/")[-1]] = v if v.ndim else v.item()
        exp[tag] = d
    return A, exp


def rand_csr(m, n, density, seed, vals="uniform"):
    rng = np.random.default_rng(seed)
    S = sp.random(m, n, density=density, format="csr", random_state=rng)
    if vals == "int":
        S.data = rng.integers(-3, 4, S.nnz).astype(np.float64)
        S.data[S.data == 0] = 1.0
    else:
        S.data = rng.uniform(-1, 1, S.nnz)
    return from_scipy(S)


def lap2d(nx, ny):
    ex = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(nx, nx))
    ey = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(ny, ny))
    L = sp.kron(sp.identity(ny), ex) + sp.kron(ey, sp.identity(nx))
    return from_scipy(L)


# ------------------------------------------------------------------ conversion
@pytest.mark.parametrize("name", CASES)
def test_csr2tile_matches_reference_golden(name):
    z = golden()
    A, exp = gold_case(z, name)
    assert_tiles_equal(O.csr2tile(A), exp["A"], label=f"{name} A ")
    assert_tiles_equal(O.csr2tile_b(A), exp["B"], label=f"{name} B ")


@pytest.mark.parametrize("seed", range(4))
def test_csr2tile_matches_reference_live(seed):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    rng = np.random.default_rng(seed)
    m, n = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    A = rand_csr(m, n, float(rng.uniform(0.005, 0.2)), seed)
    assert_tiles_equal(O.csr2tile(A), O.ref_csr2tile(A), label="A ")
    assert_tiles_equal(O.csr2tile_b(A), O.ref_csr2tile(A, as_b=True), label="B ")


def with_duplicates(A, seed):
    """A copy of A in which some entries appear twice (adjacent, as load_mtx
    keeps them: src/tile_format.h:78-89), with a different value."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(A.m), np.diff(A.rowptr))
    dup = rng.random(len(A.col)) < 0.15
    col = np.repeat(A.col, 1 + dup)
    val = np.repeat(A.val, 1 + dup)
    first = np.r_[0, np.cumsum(1 + dup)[:-1]]
    val[first[dup] + 1] = rng.uniform(-1, 1, int(dup.sum()))
    rp = np.zeros(A.m + 1, np.int64)
    np.add.at(rp, rows + 1, 1 + dup)
    return CSR(A.m, A.n, np.cumsum(rp), col, val)


@pytest.mark.parametrize("name", CASES)
def test_tile_csc_matches_reference_golden(name):
    """Within-tile CSC against the reference's own build_csc_tiles (:547-584)
    and build_a_tile_csc (:589-627) output."""
    z = golden()
    A, _ = gold_case(z, name)
    for tag, d in (("A", O.csr2tile(A)), ("B", O.csr2tile_b(A))):
        exp = {k.split("/")[-1]: z[k] for k in z.files if k.startswith(f"{name}/{tag}csc/")}
        assert_csc_equal(O.tile_csc(d), exp, label=f"{name} {tag} ")


@pytest.mark.parametrize("seed", range(4))
def test_tile_csc_matches_reference_live(seed):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    rng = np.random.default_rng(100 + seed)
    m, n = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    A = rand_csr(m, n, float(rng.uniform(0.005, 0.2)), seed)
    if seed % 2:
        A = with_duplicates(A, seed)  # repeated columns keep storage order
    assert_csc_equal(O.tile_csc(O.csr2tile(A)), O.ref_tile_csc(A), label="A ")
    assert_csc_equal(O.tile_csc(O.csr2tile_b(A)), O.ref_tile_csc(A, as_b=True), label="B ")


def csc_vs_transpose(A, csc):
    """Closed form: the within-tile CSC of A's tile (I, J) is the within-tile
    CSR of the transpose's tile (J, I) with rows and columns swapped."""
    a = O.csr2tile(A)
    t = O.csr2tile(from_scipy(A.to_scipy().T))
    where = {}
    for J in range(t["tilem"]):
        for q in range(t["tile_ptr"][J], t["tile_ptr"][J + 1]):
            where[(J, int(t["tile_columnidx"][q]))] = q
    for I in range(a["tilem"]):
        for p in range(a["tile_ptr"][I], a["tile_ptr"][I + 1]):
            q = where[(int(a["tile_columnidx"][p]), I)]
            s0, s1 = a["tile_nnz"][p], a["tile_nnz"][p + 1]
            t0, t1 = t["tile_nnz"][q], t["tile_nnz"][q + 1]
            if not (np.array_equal(csc["tile_csc_Ptr"][p * 16:p * 16 + 16], t["tile_csr_Ptr"][q * 16:q * 16 + 16])
                    and np.array_equal(csc["tile_csc_Row"][s0:s1], t["tile_csr_Col"][t0:t1] & 15)
                    and np.array_equal(csc["tile_csc_Val"][s0:s1], t["tile_csr_Val"][t0:t1])):
                return False
    return True


@pytest.mark.parametrize("shape", [(150, 170, 0.08), (40, 40, 1.0), (97, 33, 0.3)])
def test_tile_csc_is_the_tile_transpose(shape):
    m, n, d = shape
    A = rand_csr(m, n, d, 7)
    assert csc_vs_transpose(A, O.tile_csc(O.csr2tile(A)))


def test_planted_csc_row_order_is_caught():
    """Rows of a column written in descending order (a plausible scatter
    mistake) must fail the transpose check."""
    A = rand_csr(64, 64, 0.3, 3)
    csc = O.tile_csc(O.csr2tile(A))
    tnz = O.csr2tile(A)["tile_nnz"]
    bad = dict(csc)
    row, val = csc["tile_csc_Row"].copy(), csc["tile_csc_Val"].copy()
    t = 0
    ptr = csc["tile_csc_Ptr"][t * 16:t * 16 + 16].astype(np.int64)
    end = np.r_[ptr[1:], tnz[t + 1] - tnz[t]]
    c = int(np.argmax(end - ptr))  # the fullest column of tile 0
    sl = slice(tnz[t] + ptr[c], tnz[t] + end[c])
    row[sl], val[sl] = row[sl][::-1].copy(), val[sl][::-1].copy()
    bad["tile_csc_Row"], bad["tile_csc_Val"] = row, val
    assert end[c] - ptr[c] > 1
    assert not csc_vs_transpose(A, bad)


@pytest.mark.parametrize("seed", range(3))
def test_build_bitmasks_matches_reference_and_tile_pattern_product(seed):
    """build_bitmasks (:628-645): against the reference's own build, and the
    closed form popcount(bmA[I] & bmB[J]) = number of k with A tile (I, k)
    and B tile (k, J), i.e. the (I, J) entry of the tile-pattern product."""
    rng = np.random.default_rng(300 + seed)
    m, k, n = (int(x) for x in rng.integers(20, 400, 3))
    A = rand_csr(m, k, float(rng.uniform(0.005, 0.1)), seed)
    B = rand_csr(k, n, float(rng.uniform(0.005, 0.1)), seed + 10)
    bmA, bmB = O.build_bitmasks(O.csr2tile(A), O.csr2tile_b(B))
    if O.ref_available():
        rA, rB = O.ref_build_bitmasks(A, B)
        assert np.array_equal(bmA, rA) and np.array_equal(bmB, rB)
    def tile_pattern(S, tm, tn):
        S = S.tocoo()
        P = sp.csr_matrix((np.ones(S.nnz), (S.row // 16, S.col // 16)), shape=(tm, tn))
        P.data[:] = 1.0
        return P
    PA = tile_pattern(A.to_scipy(), (m + 15) // 16, (k + 15) // 16)
    PB = tile_pattern(B.to_scipy(), (k + 15) // 16, (n + 15) // 16)
    pairs = (PA @ PB).toarray()
    inter = np.bitwise_and(bmA[:, None, :], bmB[None, :, :])
    pop = np.vectorize(lambda x: bin(int(x)).count("1"))(inter).sum(axis=2)
    assert np.array_equal(pop, pairs.astype(np.int64))


TS_SIZES = (1, 5, 8, 13)


def gold_ts(z, name, tag, ts):
    return {k.split("/")[-1]: (z[k] if z[k].ndim else z[k].item())
            for k in z.files if k.startswith(f"{name}/{tag}ts{ts}/")}


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("ts", TS_SIZES)
def test_csr2tile_ts_matches_reference_golden(name, ts):
    """csr2tile_ts (:225-324) and build_csc_tiles_ts (:327-417) against the
    reference's own output."""
    z = golden()
    A, _ = gold_case(z, name)
    assert_tiles_equal(O.csr2tile_ts(A, ts), gold_ts(z, name, "A", ts), label=f"{name} ts{ts} A ")
    assert_tiles_equal(O.csr2tile_ts(A, ts, as_b=True), gold_ts(z, name, "B", ts), label=f"{name} ts{ts} B ")


@pytest.mark.parametrize("seed", range(4))
def test_csr2tile_ts_matches_reference_live(seed):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    rng = np.random.default_rng(200 + seed)
    m, n = int(rng.integers(1, 250)), int(rng.integers(1, 250))
    A = rand_csr(m, n, float(rng.uniform(0.01, 0.25)), seed)
    if seed % 2:
        A = with_duplicates(A, seed)
    for ts in (2, 3, 7, 11, 15, 16):
        assert_tiles_equal(O.csr2tile_ts(A, ts), O.ref_csr2tile_ts(A, ts), label=f"ts{ts} A ")
        assert_tiles_equal(O.csr2tile_ts(A, ts, as_b=True), O.ref_csr2tile_ts(A, ts, as_b=True), label=f"ts{ts} B ")


def test_csr2tile_ts_closed_forms():
    """ts = 16 is csr2tile / build_csc_tiles; for every ts the tiles hold exactly
    the CSR's entries: row = tile row * ts + (col byte >> 4), column = tile
    column * ts + (col byte & 15), each row's entries in storage order."""
    A = rand_csr(97, 83, 0.15, 11)
    assert_tiles_equal(O.csr2tile_ts(A, 16), O.csr2tile(A))
    assert_tiles_equal(O.csr2tile_ts(A, 16, as_b=True), O.csr2tile_b(A))
    S = A.to_scipy()
    for ts in range(1, 17):
        d = O.csr2tile_ts(A, ts)
        trow = np.repeat(np.arange(d["tilem"]), np.diff(d["tile_ptr"]))
        per_tile = np.diff(d["tile_nnz"])
        r = np.repeat(trow, per_tile) * ts + (d["tile_csr_Col"] >> 4)
        c = np.repeat(d["tile_columnidx"], per_tile) * ts + (d["tile_csr_Col"] & 15)
        order = np.lexsort((c, r))
        R = sp.csr_matrix((d["tile_csr_Val"][order], (r[order], c[order])), shape=S.shape)
        assert (R != S).nnz == 0, f"ts={ts}"
        # masks: bit ts-1-c for every stored column; row offsets: exclusive prefix
        ptr = d["tile_csr_Ptr"].reshape(-1, ts).astype(np.int64)
        assert np.all(ptr[:, 0] == 0) and np.all(np.diff(ptr, axis=1) >= 0)
        pop = np.array([bin(int(x)).count("1") for x in d["mask"]]).reshape(-1, ts)
        assert np.array_equal(pop.sum(axis=1), per_tile)
        assert int(d["mask"].max(initial=0)) < (1 << ts)


@pytest.mark.parametrize("name", ["sym", "pattern", "complex"])
def test_load_mtx_matches_reference_golden(name):
    z = golden()
    p = os.path.join(os.path.dirname(GOLD), f"{name}.mtx")
    m, n, rp, col, val = O.load_mtx(p)
    assert (m, n) == (int(z[f"mtx_{name}/m"]), int(z[f"mtx_{name}/n"]))
    np.testing.assert_array_equal(rp, z[f"mtx_{name}/rowptr"])
    np.testing.assert_array_equal(col, z[f"mtx_{name}/col"])
    np.testing.assert_array_equal(val, z[f"mtx_{name}/val"])


# ------------------------------------------------------------------ closed forms
def tile_pattern(S):
    S = sp.coo_matrix(S)
    tm, tn = (S.shape[0] + 15) // 16, (S.shape[1] + 15) // 16
    P = sp.coo_matrix((np.ones(S.nnz), (S.row // 16, S.col // 16)), shape=(tm, tn)).tocsr()
    P.data[:] = 1
    return P


def closed_form_structure(A, B):
    """C tile set = bool(tilepat(A) @ tilepat(B)); masks = element pattern of bool(A) @ bool(B)."""
    SA, SB = A.to_scipy(), B.to_scipy()
    T = (tile_pattern(SA) @ tile_pattern(SB)).tocsr()
    T.sort_indices()
    PA, PB = SA.copy(), SB.copy()
    PA.data[:] = 1
    PB.data[:] = 1
    E = sp.coo_matrix(PA @ PB)
    return T, E


def check_structure(Cd, A, B):
    T, E = closed_form_structure(A, B)
    np.testing.assert_array_equal(Cd["tile_ptr"], T.indptr)
    np.testing.assert_array_equal(Cd["tile_columnidx"], T.indices)
    # masks from the element pattern
    tid = {}
    for t, (r, c) in enumerate(zip(Cd["tile_rowidx"], Cd["tile_columnidx"])):
        tid[(int(r), int(c))] = t
    mask = np.zeros(len(Cd["tile_columnidx"]) * 16, dtype=np.uint16)
    for r, c in zip(E.row, E.col):
        t = tid[(r // 16, c // 16)]
        mask[t * 16 + r % 16] |= np.uint16(1 << (15 - c % 16))
    np.testing.assert_array_equal(Cd["mask"], mask)
    pc = np.array([bin(int(x)).count("1") for x in mask]).reshape(-1, 16)
    np.testing.assert_array_equal(np.diff(Cd["tile_nnz"]), pc.sum(1))
    excl = np.concatenate([np.zeros((pc.shape[0], 1), int), np.cumsum(pc, 1)[:, :-1]], 1)
    np.testing.assert_array_equal(Cd["tile_csr_Ptr"].reshape(-1, 16), excl)


def tiles_to_dense(Cd):
    M = np.zeros((Cd["m"], Cd["n"]))
    for t in range(len(Cd["tile_columnidx"])):
        tr, tc = Cd["tile_rowidx"][t], Cd["tile_columnidx"][t]
        for e in range(Cd["tile_nnz"][t], Cd["tile_nnz"][t + 1]):
            b = int(Cd["tile_csr_Col"][e])
            M[tr * 16 + (b >> 4), tc * 16 + (b & 15)] += Cd["tile_csr_Val"][e]
    return M


@pytest.mark.parametrize("shape", [(40, 40, 40, 0.2), (33, 70, 50, 0.1), (100, 100, 100, 0.03), (17, 9, 31, 0.5)])
def test_tiled_oracle_closed_forms(shape):
    m, k, n, d = shape
    A, B = rand_csr(m, k, d, 1), rand_csr(k, n, d, 2)
    Cd = O.spgemm_tiled(A, B)
    check_structure(Cd, A, B)
    dense = A.to_scipy().toarray() @ B.to_scipy().toarray()
    np.testing.assert_allclose(tiles_to_dense(Cd), dense, rtol=1e-13, atol=1e-13)
    chk = O.gustavson_check(A, B, Cd)
    assert chk["missing"] == 0 and chk["extra"] == 0 and chk["pass_"] == 1
    assert chk["max_scaled"] <= 1e-15


def test_tiled_oracle_stencil_exact():
    A = lap2d(37, 29)
    Cd = O.spgemm_tiled(A, A)
    check_structure(Cd, A, A)
    exact = (A.to_scipy() @ A.to_scipy()).toarray()
    assert np.array_equal(tiles_to_dense(Cd), exact)   # integer arithmetic: bit-exact


def test_tiled_oracle_f32_matches_f64_within_tolerance():
    A = rand_csr(120, 120, 0.05, 7)
    c64 = O.spgemm_tiled(A, A)
    c32 = O.spgemm_tiled(A, A, f32=True)
    for k in ("tile_ptr", "tile_columnidx", "tile_nnz", "mask", "tile_csr_Col"):
        np.testing.assert_array_equal(c64[k], c32[k])
    chk = O.gustavson_check(A, A, c32)
    assert chk["max_scaled"] < 1e-6 and chk["pass_"] == 1


def test_empty_and_degenerate():
    Z = CSR(20, 20, np.zeros(21), np.zeros(0), np.zeros(0))
    Cd = O.spgemm_tiled(Z, Z)
    assert Cd["numtile"] == 0 and Cd["nnz"] == 0
    A = rand_csr(1, 1, 1.0, 3)
    Cd = O.spgemm_tiled(A, A)
    assert Cd["numtile"] == 1 and Cd["nnz"] == 1
    assert Cd["tile_csr_Val"][0] == A.val[0] * A.val[0]


def test_gustavson_matches_scipy_and_flops():
    A = rand_csr(150, 150, 0.04, 9)
    rp, col, val = O.gustavson_csr(A, A)
    S = A.to_scipy()
    ref = (S @ S).toarray()
    D = sp.csr_matrix((val, col, rp), shape=(150, 150)).toarray()
    np.testing.assert_allclose(D, ref, rtol=1e-14, atol=1e-14)
    P = S.copy()
    P.data[:] = 1
    assert len(col) == (P @ P).nnz        # structural union, no cancellation drop
    expected = sum(int(S.indptr[c + 1] - S.indptr[c]) for c in S.indices)
    assert O.flops(A, A) == expected


# ------------------------------------------------------------------ planted mistakes
def test_planted_transposed_operand_is_caught():
    A, B = rand_csr(60, 60, 0.08, 4), rand_csr(60, 60, 0.08, 5)
    Bt = from_scipy(B.to_scipy().T)          # a transposed operand: same shape, wrong product
    check_structure(O.spgemm_tiled(A, B), A, B)
    with pytest.raises(AssertionError):
        check_structure(O.spgemm_tiled(A, Bt), A, B)


def test_planted_dropped_pair_is_caught():
    A = rand_csr(64, 64, 0.1, 6)
    Cd = O.spgemm_tiled(A, A)
    bad = dict(Cd)
    bad["tile_csr_Val"] = Cd["tile_csr_Val"].copy()
    bad["tile_csr_Val"][len(bad["tile_csr_Val"]) // 2] += 0.25   # one contribution lost/added
    chk = O.gustavson_check(A, A, bad)
    assert chk["max_scaled"] > 1e-3


def test_planted_mask_bit_is_caught():
    A = rand_csr(48, 48, 0.1, 8)
    Cd = O.spgemm_tiled(A, A)
    bad = dict(Cd)
    bad["mask"] = Cd["mask"].copy()
    bad["mask"][np.nonzero(bad["mask"])[0][0]] ^= 1
    with pytest.raises(AssertionError):
        check_structure(bad, A, A)


@pytest.mark.parametrize("case", ["er", "rect", "rmat", "dup", "empty_rows"])
def test_count_product_matches_tiled_steps_and_gustavson(case):
    """ora_count_product (threaded nnz(C) and numtileC straight from the CSRs,
    used where C cannot come back to the host) against Step 1's tile set and
    Step 2's nnz of the tiled restatement and the Gustavson structural union,
    on square and rectangular products (a swapped dimension fails the latter)."""
    import gpu_spgemm_amd as g
    rng = np.random.default_rng(3)
    if case == "er":
        A = g.gen_er(3000, 7, seed=1); B = A
    elif case == "rect":   # 200 x 530 times 530 x 77
        A = sp.random(200, 530, density=0.02, random_state=4, format="csr")
        B = sp.random(530, 77, density=0.05, random_state=5, format="csr")
        A = g.HostCSR(200, 530, A.indptr, A.indices, A.data)
        B = g.HostCSR(530, 77, B.indptr, B.indices, B.data)
    elif case == "rmat":
        A = g.gen_rmat(12, 16, seed=2, dtype=g.F64); B = A
    elif case == "dup":
        rp = np.array([0, 3, 5, 6, 8, 8, 8, 8, 9], dtype=np.int64)
        col = np.array([1, 1, 7, 0, 3, 2, 2, 3, 6], dtype=np.int32)
        A = g.HostCSR(8, 8, rp, col, np.arange(1.0, 10.0)); B = A
    else:
        rp = np.concatenate([[0], np.sort(rng.integers(0, 120, 99)), [120]]).astype(np.int64)
        col = rng.integers(0, 100, 120).astype(np.int32)
        for i in range(100):
            col[rp[i]:rp[i + 1]] = np.sort(col[rp[i]:rp[i + 1]])
        A = g.HostCSR(100, 100, rp, col, np.ones(120)); B = A
    for nt in (1, 5):
        nnz, tiles = O.count_product(A, B, nthreads=nt)
        Cd = O.spgemm_tiled(A, B)
        rp, _, _ = O.gustavson_csr(A, B)
        assert tiles == Cd["numtile"]
        assert nnz == int(rp[-1]) == Cd["nnz"]


def test_cpu_baseline_on_rectangular_products():
    """The CPU baseline (conversion + Gustavson) on an m x k times k x n product
    with m, k, n all different, as in the Galerkin chain's A P and P^T (A P):
    its product count equals the flop count, with and without conversion and
    on several threads."""
    import gpu_spgemm_amd as g
    A = sp.random(700, 1900, density=0.01, random_state=7, format="csr")
    B = sp.random(1900, 260, density=0.02, random_state=8, format="csr")
    A = g.HostCSR(700, 1900, A.indptr, A.indices, A.data)
    B = g.HostCSR(1900, 260, B.indptr, B.indices, B.data)
    f = O.flops(A, B)
    for nt, conv in ((1, True), (3, True), (2, False)):
        prod, gus, ct = O.cpu_baseline(A, B, 0, A.m, nthreads=nt, convert=conv)
        assert prod == f and gus >= 0 and (ct > 0) == conv
