"""Writes the golden fixtures for the tile-format conversion and the MTX reader.

Expected outputs come from the reference's own host code (src/tile_format.h:
load_mtx, csr2tile, build_csc_tiles with its within-tile CSC, build_a_tile_csc,
csr2tile_ts and build_csc_tiles_ts) compiled from the reference checkout into
oracle/_ref/libref.so by oracle/Makefile; inputs are small synthetic matrices
and hand-written .mtx files.  Run in a container that has the reference
checkout:  make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


class CSR:
    def __init__(self, m, n, rowptr, col, val):
        self.m, self.n = m, n
        self.rowptr, self.col, self.val = rowptr.astype(np.int64), col.astype(np.int32), val.astype(np.float64)


def from_dense_pattern(m, n, rows, cols, vals):
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(m + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return CSR(m, n, np.cumsum(rp), cols, vals)


def cases():
    rng = np.random.default_rng(12345)
    out = {}
    # ER-like, ragged last tile-row (200 rows = 12.5 tiles)
    m = n = 200
    rows, cols = [], []
    for r in range(m):
        c = rng.choice(n, 8, replace=False)
        rows += [r] * 8
        cols += list(c)
    rows, cols = np.array(rows), np.array(cols)
    out["er200"] = from_dense_pattern(m, n, rows, cols, rng.uniform(-1, 1, len(rows)))
    # 2-D 5-point Laplacian on a 20 x 13 grid
    nx, ny = 20, 13
    rr, cc, vv = [], [], []
    for y in range(ny):
        for x in range(nx):
            r = y * nx + x
            for dx, dy in ((0, -1), (-1, 0), (0, 0), (1, 0), (0, 1)):
                X, Y = x + dx, y + dy
                if 0 <= X < nx and 0 <= Y < ny:
                    rr.append(r); cc.append(Y * nx + X); vv.append(4.0 if (dx, dy) == (0, 0) else -1.0)
    out["lap20x13"] = from_dense_pattern(nx * ny, nx * ny, np.array(rr), np.array(cc), np.array(vv))
    # rectangular with empty rows and columns
    m, n = 77, 300
    k = 400
    rows = rng.integers(0, m, k); cols = rng.integers(0, n, k)
    key = np.unique(rows * n + cols)
    keep = (key // n) % 5 != 0          # every fifth row left empty
    rows, cols = key[keep] // n, key[keep] % n
    out["rect77x300"] = from_dense_pattern(m, n, rows, cols, rng.uniform(-1, 1, len(rows)))
    # fully dense 40 x 40 (full 16x16 tiles, 256 nnz each)
    m = n = 40
    rows, cols = np.divmod(np.arange(m * n), n)
    out["dense40"] = from_dense_pattern(m, n, rows, cols, rng.uniform(-1, 1, m * n))
    # almost empty
    out["sparse50"] = from_dense_pattern(50, 50, np.array([0, 17, 49]), np.array([49, 17, 0]), np.array([1.0, 2.0, 3.0]))
    return out


TS_SIZES = (1, 5, 8, 13)

MTX = {
    "sym.mtx": "%%MatrixMarket matrix coordinate real symmetric\n% comment\n5 5 6\n1 1 2.5\n2 1 -1\n3 2 0.5\n4 4 7\n5 1 3\n5 5 1\n",
    "pattern.mtx": "%%MatrixMarket matrix coordinate pattern general\n4 6 5\n1 6\n2 2\n4 1\n3 3\n1 2\n",
    "complex.mtx": "%%MatrixMarket matrix coordinate complex general\n3 3 3\n1 1 1.5 9\n3 2 -2 1\n2 3 0.25 -4\n",
}


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: build it with `make -C oracle` next to the reference checkout")
    data = {}
    for name, A in cases().items():
        data[f"{name}/m"] = np.array(A.m); data[f"{name}/n"] = np.array(A.n)
        data[f"{name}/rowptr"] = A.rowptr; data[f"{name}/col"] = A.col; data[f"{name}/val"] = A.val
        for tag, as_b in (("A", False), ("B", True)):
            d = O.ref_csr2tile(A, as_b=as_b)
            for k, v in d.items():
                if isinstance(v, np.ndarray):
                    data[f"{name}/{tag}/{k}"] = v
                elif v is not None:
                    data[f"{name}/{tag}/{k}"] = np.array(v)
            # within-tile CSC: build_csc_tiles (:547-584) for B, build_a_tile_csc (:589-627) for A
            for k, v in O.ref_tile_csc(A, as_b=as_b).items():
                data[f"{name}/{tag}csc/{k}"] = v
            # parameterized tile sizes: csr2tile_ts (:225-324), build_csc_tiles_ts (:327-417)
            for ts in TS_SIZES:
                for k, v in O.ref_csr2tile_ts(A, ts, as_b=as_b).items():
                    if isinstance(v, np.ndarray):
                        data[f"{name}/{tag}ts{ts}/{k}"] = v
                    elif v is not None:
                        data[f"{name}/{tag}ts{ts}/{k}"] = np.array(v)
    for fname, text in MTX.items():
        p = os.path.join(HERE, fname)
        with open(p, "w") as f:
            f.write(text)
        m, n, rp, col, val = O.ref_load_mtx(p)
        key = fname.replace(".mtx", "")
        data[f"mtx_{key}/m"] = np.array(m); data[f"mtx_{key}/n"] = np.array(n)
        data[f"mtx_{key}/rowptr"] = rp; data[f"mtx_{key}/col"] = col; data[f"mtx_{key}/val"] = val
    np.savez_compressed(os.path.join(HERE, "tile_format_golden.npz"), **data)
    print("wrote", len(data), "arrays")


if __name__ == "__main__":
    main()
