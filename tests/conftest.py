import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (runs through libtsg's HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    """One libtsg context on device 0.  A GPU test without a GPU fails loudly."""
    import gpu_spgemm_amd as g
    n = g.device_count()
    if n < 1:
        pytest.fail("GPU test selected but no HIP device is visible")
    c = g.Context(0)
    print(f"[tsg] device: {c.device_name()}", flush=True)
    yield c
    c.close()


def assert_tiles_equal(got, exp, keys=None, label=""):
    """Bit-exact comparison of tile dicts (download() / oracle layout)."""
    keys = keys or ["tile_ptr", "tile_columnidx", "tile_nnz", "tile_csr_Ptr", "mask", "tile_csr_Col",
                    "tile_csr_Val", "csc_tile_ptr", "csc_tile_rowidx", "tile_rowidx"]
    for k in ("m", "n", "tilem", "tilen", "nnz", "numtile"):
        assert int(got[k]) == int(exp[k]), f"{label}{k}: {got[k]} != {exp[k]}"
    for k in keys:
        a, b = got.get(k), exp.get(k)
        if a is None or b is None:
            continue
        assert a.shape == b.shape, f"{label}{k}: shape {a.shape} != {b.shape}"
        if a.dtype.kind == "f":
            ok = np.array_equal(a.view(np.uint64 if a.dtype == np.float64 else np.uint32),
                                b.astype(a.dtype).view(np.uint64 if a.dtype == np.float64 else np.uint32))
        else:
            ok = np.array_equal(a, b)
        if not ok:
            bad = np.nonzero(a != b)[0][:5]
            raise AssertionError(f"{label}{k}: first mismatches at {bad}: {a[bad]} vs {b[bad]}")


CSC_KEYS = ("tile_csc_Ptr", "tile_csc_Row", "tile_csc_Val")


def assert_csc_equal(got, exp, label=""):
    """Bit-exact comparison of within-tile CSC arrays (tile_csc_Ptr/Row/Val)."""
    for k in CSC_KEYS:
        a, b = np.asarray(got[k]), np.asarray(exp[k])
        assert a.shape == b.shape, f"{label}{k}: shape {a.shape} != {b.shape}"
        if a.dtype.kind == "f":
            vt = np.uint64 if a.dtype == np.float64 else np.uint32
            a, b = a.view(vt), b.astype(a.dtype).view(vt)
        assert np.array_equal(a, b), f"{label}{k}: first mismatches at {np.nonzero(a != b)[0][:5]}"

