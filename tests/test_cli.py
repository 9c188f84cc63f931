"""The spgemm_raw_v2 command-line drop-in (src/spgemm_raw_v2.cu:61-372): flags,
exit codes and stdout lines of the reference, results equal to the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpu_spgemm_amd as g
from oracle import oracle as O

CLI = os.path.join(os.path.dirname(g.LIB_PATH), "spgemm_raw_v2")
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def run(*args, timeout=120):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_built():
    assert os.access(CLI, os.X_OK), "spgemm_raw_v2 missing: run python -m gpu_spgemm_amd.build"


@pytest.mark.parametrize("args", [[], ["--aat", "1"], ["--matrix"], ["--matrix", "x.mtx", "--aat", "one"],
                                  ["--bogus", "1"]])
def test_cli_argument_errors_exit_1(args):
    p = run(*args)
    assert p.returncode == 1 and "Usage" in p.stderr


def parse(out):
    m = re.search(r"Raw Tile SpGEMM v2: tiles_C = (\d+), nnzC = (\d+)", out)
    f = re.search(r"SpGEMM flops = (\d+)", out)
    t = re.search(r"GPU time = ([\d.]+) ms, GFlops = ([\d.]+)", out)
    assert m and f and t, out
    for s in ("Step1 (tile structure)", "Step2 (symbolic+scan)", "Step3 (numeric v2)", "Done."):
        assert s in out
    return int(m.group(1)), int(m.group(2)), int(f.group(1))


@pytest.mark.gpu
@pytest.mark.parametrize("sched", ["plain", "ctile"])
def test_cli_generated_laplacian(sched):
    print(f"[case] cli lap2d 96x80 {sched}", flush=True)
    p = run("--gen", "lap2d:96:80", "--scheduler", sched, "--repeat", "3")
    assert p.returncode == 0, p.stdout + p.stderr
    tiles, nnz, flops = parse(p.stdout)
    A = g.gen_laplace2d(96, 80)
    exp = O.spgemm_tiled(A, A)
    assert (tiles, nnz, flops) == (exp["numtile"], exp["nnz"], O.flops(A, A))


@pytest.mark.gpu
@pytest.mark.parametrize("name,aat", [("sym.mtx", 0), ("sym.mtx", 1), ("pattern.mtx", 1), ("complex.mtx", 0)])
def test_cli_mtx_files(name, aat):
    path = os.path.join(GOLD, name)
    A = g.load_mtx(path)
    if aat == 0 and A.m != A.n:
        pytest.skip("not square")
    print(f"[case] cli {name} aat={aat}", flush=True)
    p = run("--matrix", path, "--aat", str(aat))
    assert p.returncode == 0, p.stdout + p.stderr
    B = A.transpose() if aat else A
    exp = O.spgemm_tiled(A, B)
    tiles, nnz, flops = parse(p.stdout)
    assert (tiles, nnz, flops) == (exp["numtile"], exp["nnz"], O.flops(A, B))
    assert f"A: {A.m} x {A.n}, nnz = {A.nnz}" in p.stdout


@pytest.mark.gpu
def test_cli_missing_file_and_nonsquare(tmp_path):
    p = run("--matrix", str(tmp_path / "absent.mtx"))
    assert p.returncode == 1 and "Cannot open" in p.stdout
    f = tmp_path / "rect.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n3 5 2\n1 1 1.0\n3 5 2.0\n")
    p = run("--matrix", str(f))
    assert p.returncode == 1 and "A^2 requires square" in p.stdout
    p = run("--matrix", str(f), "--aat", "1")
    assert p.returncode == 0, p.stdout + p.stderr
