"""CPU-only checks of libtsg: the C ABI loads and exports every entry point
declared in include/tsg.h, and the host-side pieces (generators, MTX reader,
transpose, product count) are right.  No kernel runs here."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import gpu_spgemm_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "tile_format_golden.npz")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tsg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = g.lib()
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_no_device_is_a_clean_error():
    if g.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(g.TsgError) as e:
        g.Context(0)
    assert e.value.code == -7


def test_laplace2d_closed_form():
    A = g.gen_laplace2d(7, 5)
    n = 35
    assert A.m == A.n == n
    # interior 5, edge 4, corner 3 entries per row
    assert A.nnz == 5 * n - 2 * 7 - 2 * 5
    D = A.to_scipy().toarray()
    assert np.all(np.diag(D) == 4)
    assert np.array_equal(D, D.T)
    assert np.all(D.sum(1)[[8, 9, 10]] == 0)  # interior rows sum to zero
    for r in range(n):
        seg = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(seg) > 0)


def test_config_sizes_match_survey():
    # 2048x2048 5-point Laplacian: 20,963,328 nonzeros; 27-point 128^3: 55,742,968
    assert 5 * 2048 * 2048 - 4 * 2048 == 20963328
    A = g.gen_laplace2d(64, 64)
    assert A.nnz == 5 * 64 * 64 - 4 * 64
    S = g.gen_stencil27(8, 8, 8)
    per = lambda k: 3 * k - 2
    assert S.nnz == per(8) ** 3
    assert per(128) ** 3 == 55742968


def test_er_exact_distinct_sorted():
    A = g.gen_er(4096, 8, seed=0)
    assert A.nnz == 4096 * 8
    cols = A.col.reshape(4096, 8)
    assert np.all(np.diff(cols, axis=1) > 0)
    assert A.val.min() >= -1 and A.val.max() <= 1
    B = g.gen_er(4096, 8, seed=0)
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    C = g.gen_er(4096, 8, seed=1)
    assert not np.array_equal(A.col, C.col)


def test_rmat_merged_sorted():
    A = g.gen_rmat(12, 16, seed=1)
    assert A.m == A.n == 4096 and A.val.dtype == np.float32
    for r in range(0, 4096, 97):
        seg = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(seg) > 0)
    assert A.nnz < 16 * 4096           # duplicates merged
    assert A.val.min() > 0 and A.val.max() <= 1
    # power law: row 0 is the heaviest quadrant
    deg = np.diff(A.rowptr)
    assert deg[0] == deg.max()


def test_aggregation_and_stencils():
    P = g.gen_aggregation(4, 4, 4)
    assert P.m == 64 and P.n == 8 and P.nnz == 64
    assert np.all(np.bincount(P.col, minlength=8) == 8)
    L = g.gen_laplace3d(5, 4, 3)
    D = L.to_scipy().toarray()
    assert np.all(np.diag(D) == 6) and np.array_equal(D, D.T)


@pytest.mark.parametrize("name", ["sym", "pattern", "complex"])
def test_load_mtx_matches_reference_golden(name):
    z = np.load(GOLD, allow_pickle=False)
    A = g.load_mtx(os.path.join(ROOT, "tests", "golden", f"{name}.mtx"))
    assert (A.m, A.n) == (int(z[f"mtx_{name}/m"]), int(z[f"mtx_{name}/n"]))
    np.testing.assert_array_equal(A.rowptr, z[f"mtx_{name}/rowptr"])
    np.testing.assert_array_equal(A.col, z[f"mtx_{name}/col"])
    np.testing.assert_array_equal(A.val, z[f"mtx_{name}/val"])


def test_mtx_roundtrip(tmp_path):
    A = g.gen_er(300, 5, seed=3)
    p = str(tmp_path / "a.mtx")
    g.write_mtx(p, A)
    B = g.load_mtx(p)
    assert np.array_equal(A.rowptr, B.rowptr) and np.array_equal(A.col, B.col)
    assert np.array_equal(A.val, B.val)


def test_load_mtx_errors(tmp_path):
    with pytest.raises(g.TsgError) as e:
        g.load_mtx(str(tmp_path / "missing.mtx"))
    assert e.value.code == -6
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n9 9 2.0\n")
    with pytest.raises(g.TsgError):
        g.load_mtx(str(bad))


def test_transpose_and_products():
    A = g.gen_er(500, 6, seed=5)
    T = A.transpose()
    np.testing.assert_array_equal(T.to_scipy().toarray(), A.to_scipy().toarray().T)
    S = A.to_scipy()
    expected = int(sum(S.indptr[c + 1] - S.indptr[c] for c in S.indices))
    assert A.products(A) == expected
    R = g.HostCSR(3, 4, [0, 1, 1, 2], [3, 0], [1.0, 2.0])
    with pytest.raises(g.TsgError):
        R.products(R)
