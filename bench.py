#!/usr/bin/env python3
"""bench.py -- tiled SpGEMM throughput on one MI355X.

Workload (default): C = A*A, A = 5-point Laplacian on a 2048 x 2048 grid
(4,194,304 rows, 20,963,328 nonzeros, values 4/-1), FP64, in 16x16 tiles.
A and B are built on the GPU (CSR -> tiles) before timing; a step is one full
product C = A*B from tiles resident in HBM to C's tiles in HBM (Steps 1-3:
structure, symbolic, numeric, with their scans), as the reference times it
(src/spgemm_raw_v2.cu:111-240).

value = GFLOP/s = 2 * products / step time (products = sum_{(i,k) in A} nnz(B[k,:])).
The JSON line also carries the pipeline's algorithmic HBM GB/s, the roofline
of the dominant kernel (HIP-event duration of the numeric launch on the
library's stream) and the CPU baseline (the oracle's Gustavson loop, the
reference's CPU check, on this host).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(g, name):
    if name == "lap2d_2048":
        A = g.gen_laplace2d(2048, 2048)
        return A, A, {"workload": "laplace2d_5pt_2048x2048_AxA", "rows": A.m, "nnzA": A.nnz, "dtype": "f64"}
    if name == "lap2d_1024":  # 65,536 tile-rows
        A = g.gen_laplace2d(1024, 1024)
        return A, A, {"workload": "laplace2d_5pt_1024x1024_AxA", "rows": A.m, "nnzA": A.nnz, "dtype": "f64"}
    if name == "er4096":
        A = g.gen_er(4096, 8, seed=0)
        return A, A, {"workload": "erdos_renyi_4096_8perrow_AxA", "rows": A.m, "nnzA": A.nnz, "dtype": "f64"}
    if name == "band":
        A = g.gen_band(65536, 48, seed=0)
        return A, A, {"workload": "dense_band_65536_halfwidth48_AxA", "rows": A.m, "nnzA": A.nnz, "dtype": "f64"}
    if name.startswith("rmat"):  # R-MAT, edge factor 16, (0.57, 0.19, 0.19, 0.05), FP32, seed 1
        sc = int(name[4:])
        A = g.gen_rmat(sc, 16, seed=1, dtype=g.F32)
        return A, A, {"workload": f"rmat_scale{sc}_ef16_AxA", "rows": A.m, "nnzA": A.nnz, "dtype": "f32"}
    if name == "stencil27_AP":
        A = g.gen_stencil27(128, 128, 128)
        P = g.gen_aggregation(128, 128, 128)
        return A, P, {"workload": "stencil27_128cube_AxP", "rows": A.m, "nnzA": A.nnz, "dtype": "f64"}
    raise SystemExit(f"unknown workload {name}")


def alg_bytes(A, B, C, s):
    """Bytes of A, B read once and C written once in the reference's layout
    (4-byte offsets): Bytes(A)+Bytes(B)+Bytes(C)."""
    ba = 4 * (A["tilem"] + 1) + 24 * A["numtile"] + 4 + (1 + s) * A["nnz"]
    bb = 4 * (B["tilen"] + 1) + 4 * (B["tilem"] + 1) + 60 * B["numtile"] + 4 + (1 + s) * B["nnz"]
    bc = 4 * (C["tilem"] + 1) + 60 * C["numtile"] + 4 + (1 + s) * C["nnz"]
    return ba + bb + bc


def numeric_bytes(A, B, C, s):
    """Algorithmic bytes of one numeric launch in this library's layout: A and
    B tile headers + payload read once, C's structure read once, C's col bytes
    and values written once."""
    a = 8 * (A["tilem"] + 1) + (4 + 8 + 16) * A["numtile"] + (1 + s) * A["nnz"]
    b = 8 * (B["tilen"] + 1) + (4 + 8 + 16 + 32) * B["numtile"] + s * B["nnz"]
    c = (4 + 4 + 8 + 32 + 16) * C["numtile"] + (1 + s) * C["nnz"]
    return a + b + c


def rows_numeric_bytes(A, B, C, s):
    """Algorithmic bytes of one launch of the row-lane numeric kernel: A's and
    B's scalar CSR read once, C's structure (tile offsets, column indices,
    value offsets, masks, row offsets) read once, C's col bytes and values
    written once."""
    a = 8 * (A["m"] + 1) + (4 + s) * A["nnz"]
    b = 8 * (B["m"] + 1) + (4 + s) * B["nnz"]
    c = 8 * (C["tilem"] + 1) + (4 + 8 + 32 + 16) * C["numtile"] + (1 + s) * C["nnz"]
    return a + b + c


# numeric kernels by their bit in tsg_stats.numeric_kernels
NUMERIC_KERNELS = {1: "k_num_rows", 2: "k_num_staged<VarS>", 4: "k_num_staged<VarW>", 8: "k_num_staged<VarM>",
                   16: "k_numeric_list", 32: "k_rgw_numeric"}


def train_learned(g, ctx, log, rmat_scale=0):
    """Trains the learned scheduler on a family of smaller matrices (never the
    benchmarked product itself), outside any timed region; for an R-MAT
    workload the family adds R-MATs 2-8x smaller with other seeds."""
    fam = [g.gen_laplace2d(256, 256), g.gen_laplace2d(512, 384), g.gen_laplace2d(300, 700),
           g.gen_laplace2d(1024, 96), g.gen_laplace2d(4096, 24),   # wider grids, few rows
           g.gen_laplace3d(40, 40, 40), g.gen_stencil27(32, 32, 32), g.gen_er(20000, 6, seed=11),
           g.gen_er(20000, 12, seed=12), g.gen_er_var(20000, 0, 24, seed=14),
           g.gen_rmat(13, 8, seed=15, dtype=g.F64)]
    if rmat_scale:   # R-MATs 2-8x smaller than the benchmarked one, other seeds
        fam += [g.gen_rmat(sc, 16, seed=100 + sc, dtype=g.F32) for sc in range(rmat_scale - 3, rmat_scale)]
    ctx.learned_clear()
    t0 = time.time()
    for M in fam:
        tA, tB = ctx.tiles_from_csr(M), ctx.tiles_from_csr(M, as_b=True)
        ctx.learned_collect(tA, tB)
        tA.free()
        tB.free()
    loss = ctx.learned_train(epochs=300, batch=4096, lr=3e-3, seed=0)
    log(f"learned scheduler trained on {len(fam)} matrices in {time.time() - t0:.1f} s, "
        f"loss {loss:.4f}, calibration {ctx.learned_calibration().tolist()}")


def copy_probe(mib):
    import torch
    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "copies": 5}), flush=True)


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summaries (lines 'TRAFFIC <workload> <kernel> <bytes>',
    the kernel name may hold spaces), or None."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        for line in open(f):
            p = line.split()
            if len(p) >= 4 and p[0] == "TRAFFIC" and p[1] == workload and " ".join(p[2:-1]) == kernel:
                return int(p[-1])
    return None


def tile_sizes(T):
    return {"tilem": T.s.tilem, "tilen": T.s.tilen, "numtile": T.s.numtile, "nnz": T.s.nnz}


def cpu_baseline_many(pairs, seconds, nthreads=1):
    """cpu_baseline's CPU path over a list of products (A_i, B_i): whole
    products in order until `seconds` have passed (at least one), the rate
    scaled to all of them."""
    from oracle import oracle as O
    total = sum(O.flops(A, B) for A, B in pairs)
    done, t = 0, 0.0
    n = 0
    for A, B in pairs:
        prod, gus, conv = O.cpu_baseline(A, B, 0, A.m, nthreads=nthreads, convert=True)
        done, t, n = done + prod, t + gus + conv, n + 1
        if t >= seconds:
            break
    est = t * total / max(done, 1)
    return {"value": round(2.0 * total / est / 1e9, 4), "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
            "sample": f"the first {n} of {len(pairs)} products ({done} of {total} products, {t:.2f} s): host "
                      f"csr2tile + build_csc_tiles on 1 thread and the Gustavson FP64 loop of the reference's check "
                      f"on {nthreads} thread(s), scaled to all products ({est:.1f} s)"}


def bench_batch(g, args):
    """512 uneven independent products A_i*A_i (FP64): one batched call vs a
    loop of single calls over the same resident operands."""
    t0 = time.time()
    probs = g.gen_batch_uneven(512, seed=4)
    products = sum(A.products(A) for _, A in probs)
    ctx = g.Context(0)
    log(f"batch512: {len(probs)} problems, rows {sum(A.m for _, A in probs)}, "
        f"products {products}, generated in {time.time() - t0:.1f} s")
    tA = [ctx.tiles_from_csr(A) for _, A in probs]
    tB = [ctx.tiles_from_csr(A, as_b=True) for _, A in probs]

    def loop():
        for a, b in zip(tA, tB):
            C, _ = ctx.spgemm(a, b)
            C.free()

    def batched():
        Cs, _ = ctx.spgemm_batched(tA, tB)
        for C in Cs:
            C.free()

    t1 = time.time()
    best, rep = ctx.autotune(tA, tB, reps=1)
    log(f"autotune ({len(rep)} configurations, {time.time() - t1:.1f} s): best var_mask {best['var_mask']} "
        f"rg_density {best['rg_density']} {best['ms']:.1f} ms; " +
        " ".join(f"{r['var_mask']}/{r['rg_density']}:{r['ms']:.0f}{'*' if r['sampled'] else ''}" for r in rep) +
        " (* = timed on every 7th problem only)")
    G = ctx.batch_graph(tA, tB, slack=1.25)
    fallbacks = []

    def graph():  # one graph launch, then the wait for its result row
        G.launch()
        _, sts = G.result()
        fallbacks.append(sts[0]["fallback_used"])

    res = {}
    for label, fn in (("loop", loop), ("batched", batched), ("graph", graph)):
        for _ in range(args.warmup):
            fn()
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.synchronize()
        res[label] = (time.perf_counter() - t1) * 1e3 / args.steps
        log(f"{label}: {res[label]:.3f} ms per batch")
    # same structure, new values: the numeric step of the whole batch as one replayed hipGraph
    plan = ctx.plan(tA, tB)
    plan.execute()
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute()
    ctx.synchronize()
    res["graph_replay_numeric"] = (time.perf_counter() - t1) * 1e3 / args.steps
    log(f"graph replay (values only): {res['graph_replay_numeric']:.3f} ms per batch")
    plan.destroy()
    G.destroy()
    Cs, sts = ctx.spgemm_batched(tA, tB)
    pipe = sum(alg_bytes(tile_sizes(a), tile_sizes(b), tile_sizes(c), 8) for a, b, c in zip(tA, tB, Cs))
    for C in Cs:
        C.free()
    ms = res["batched"]
    out = {
        "metric": "SpGEMM GFLOP/s + HBM GB/s, learned/batched vs plain tiled, 1 MI355X",
        "value": round(2.0 * products / (ms * 1e6), 3), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated on the host, converted to tiles on the GPU before timing)",
        "config": {"workload": "batch512_uneven_AixAi", "problems": len(probs), "products": products,
                   "seed": 4, "scheduler": "batched", "parallelism": "single GPU"},
        "variants_ms_per_step": {k: round(v, 4) for k, v in res.items()},
        "batched_speedup_vs_loop": round(res["loop"] / res["batched"], 3),
        "graph_speedup_vs_loop": round(res["loop"] / res["graph"], 3),
        "graph_fallbacks": int(sum(fallbacks)),
        "autotuned": {"var_mask": best["var_mask"], "rg_density": best["rg_density"],
                      "candidates": len(rep)},
        # the batch runs dozens of kernels: the bound is the whole call's algorithmic bytes
        "roofline": {"bound": "hbm", "kernel": "batched pipeline (all kernels)",
                     "achieved": round(pipe / (ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(pipe / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes_per_launch": pipe, "launch_ms": round(ms, 4)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_many([(A, A) for _, A in probs], args.cpu_seconds)
    ctx.close()
    print(json.dumps(out), flush=True)


def train_galerkin(g, ctx, log):
    """Trains the learned scheduler on smaller Galerkin chains (A P and
    P^T (A P) of 27-point stencils on 24^3-48^3 grids, other than the
    benchmarked 128^3) and a mixed family, outside any timed region."""
    ctx.learned_clear()
    for n in (24, 32, 40, 48):
        A, P = g.gen_stencil27(n, n, n), g.gen_aggregation(n, n, n)
        tA, tP, tPt = ctx.tiles_from_csr(A), ctx.tiles_from_csr(P, as_b=True), ctx.tiles_from_csr(P.transpose())
        ctx.learned_collect(tA, tP)
        C1, _ = ctx.spgemm(tA, tP)
        C1.build_csc_tiles()
        ctx.learned_collect(tPt, C1)
        for T in (C1, tA, tP, tPt):
            T.free()
    for M in (g.gen_laplace3d(40, 40, 40), g.gen_er(20000, 6, seed=11)):
        tA, tB = ctx.tiles_from_csr(M), ctx.tiles_from_csr(M, as_b=True)
        ctx.learned_collect(tA, tB)
        tA.free()
        tB.free()
    loss = ctx.learned_train(epochs=300, batch=4096, lr=3e-3, seed=0)
    log(f"learned scheduler trained on Galerkin chains 24^3-48^3, loss {loss:.4f}, "
        f"calibration {ctx.learned_calibration().tolist()}")


def bench_galerkin(g, args):
    """Galerkin triple product P^T (A P), FP64: A = 27-point stencil on a 128^3
    grid, P = 2x2x2 piecewise-constant aggregation.  A step = C1 = A P, C1
    re-laid as a B operand on the GPU, C2 = P^T C1 (two chained products),
    under the plain and the learned scheduler."""
    n = 128
    A = g.gen_stencil27(n, n, n)
    P = g.gen_aggregation(n, n, n)
    Pt = P.transpose()
    ctx = g.Context(0)
    tA, tP, tPt = ctx.tiles_from_csr(A), ctx.tiles_from_csr(P, as_b=True), ctx.tiles_from_csr(Pt)
    p1 = A.products(P)
    C1, _ = ctx.spgemm(tA, tP)
    c1 = C1.to_csr()          # products of the second product, counted once before timing
    p2 = Pt.products(c1)
    C1.build_csc_tiles()
    C2, _ = ctx.spgemm(tPt, C1)
    pipe = (alg_bytes(tile_sizes(tA), tile_sizes(tP), tile_sizes(C1), 8) +
            alg_bytes(tile_sizes(tPt), tile_sizes(C1), tile_sizes(C2), 8))
    C1.free()
    C2.free()
    log(f"galerkin 128^3: A nnz {A.nnz}, products A*P {p1}, P^T*(AP) {p2}")

    def step(**kw):
        C1, s1 = ctx.spgemm(tA, tP, **kw)
        C1.build_csc_tiles()
        C2, s2 = ctx.spgemm(tPt, C1, **kw)
        C1.free()
        C2.free()
        return s1, s2

    def timed(**kw):
        for _ in range(args.warmup):
            step(**kw)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(**kw)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    ms = timed()
    log(f"chain (plain): {ms:.3f} ms per step")
    train_galerkin(g, ctx, log)
    s1, s2 = step(scheduler=g.SCHED_LEARNED)
    ms_l = timed(scheduler=g.SCHED_LEARNED)
    log(f"chain (learned): {ms_l:.3f} ms per step, re-runs {s1['fallback_used']}/{s2['fallback_used']}, "
        f"value capacity / nnz {s1['pred_nnzC'] / s1['nnzC']:.3f} / {s2['pred_nnzC'] / s2['nnzC']:.3f}")
    out = {
        "metric": "SpGEMM GFLOP/s + HBM GB/s, learned/batched vs plain tiled, 1 MI355X",
        "value": round(2.0 * (p1 + p2) / (ms * 1e6), 3), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated on the host, converted to tiles on the GPU before timing)",
        "config": {"workload": "galerkin_PtAP_stencil27_128cube", "products": p1 + p2,
                   "scheduler": "plain", "parallelism": "single GPU"},
        "learned": {"ms_per_step": round(ms_l, 4), "vs_plain": round(ms / ms_l, 3),
                    "fallback": [s1["fallback_used"], s2["fallback_used"]]},
        "roofline": {"bound": "hbm", "kernel": "chain (both products, all kernels)",
                     "achieved": round(pipe / (ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(pipe / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes_per_launch": pipe, "launch_ms": round(ms, 4)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_many([(A, P), (Pt, c1)], args.cpu_seconds)
    ctx.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(A, B, seconds, nthreads=1):
    """The reference's CPU path restated by the oracle: host csr2tile of A and B
    plus build_csc_tiles of B (src/tile_format.h:93-222, :424-546; always one
    thread, as the reference) and the Gustavson FP64 loop of its check
    (src/spgemm_raw_v2.cu:297-308) on `nthreads` threads over row blocks.  The
    whole product when it finishes within `seconds`, else the conversion once
    and the Gustavson loop on row blocks at 8 evenly spaced starts holding
    about `seconds` worth of products, scaled to the whole product."""
    import numpy as np
    from oracle import oracle as O
    m = A.m
    cum = np.concatenate([[0], np.cumsum(np.diff(B.rowptr)[A.col], dtype=np.int64)])[A.rowptr]
    total = int(cum[-1])

    def block(start, want, stop):  # rows [start, end <= stop) holding about `want` products
        end = int(np.searchsorted(cum, cum[start] + want, side="left"))
        return start, max(start + 1, min(stop, end))

    _, _, conv = O.cpu_baseline(A, B, 0, 0, nthreads=1, convert=True)      # conversion alone
    r0, r1 = block(m // 2, min(total, 20_000_000), m)
    p0, g0, _ = O.cpu_baseline(A, B, r0, r1, nthreads=nthreads, convert=False)
    rate = p0 / max(g0, 1e-6)
    what = (f"host csr2tile(A, B) + build_csc_tiles(B) on 1 thread (src/tile_format.h:93-222, :424-546) and the "
            f"Gustavson FP64 loop of the reference's check (src/spgemm_raw_v2.cu:297-308) on {nthreads} thread(s)")
    if conv + total / rate <= seconds:
        times = []
        prod = 0
        t_end = time.time() + seconds
        while time.time() < t_end or len(times) < 3:
            prod, gus, conv = O.cpu_baseline(A, B, 0, m, nthreads=nthreads, convert=True)
            times.append(gus + conv)
            if len(times) >= 7:
                break
        med = statistics.median(times)
        return {"value": round(2.0 * prod / med / 1e9, 4), "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
                "sample": f"whole product: {what}; median of {len(times)} runs ({med:.3f} s each, of which "
                          f"conversion {conv:.3f} s; {prod} products)"}
    want = rate * max(seconds - conv, 1.0) / 8
    prod, gus, rows = 0, 0.0, 0
    for i in range(8):
        a, b = block(i * m // 8, want, (i + 1) * m // 8)
        p, t, _ = O.cpu_baseline(A, B, a, b, nthreads=nthreads, convert=False)
        prod, gus, rows = prod + p, gus + t, rows + (b - a)
    est = conv + gus * total / prod
    return {"value": round(2.0 * total / est / 1e9, 4), "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
            "sample": f"{what}: conversion of the whole operands ({conv:.2f} s) plus the Gustavson loop on {rows} of "
                      f"{m} rows in 8 blocks at evenly spaced starts ({prod} of {total} products, {gus:.2f} s), "
                      f"scaled to the whole product ({est:.1f} s)"}


def host_threads():
    """Host cores this run may use (at most 16)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1
    return max(1, min(16, n))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lap2d_2048")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-scheduler leg")
    ap.add_argument("--no-batch-leg", action="store_true", help="skip the batched-vs-loop leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB 5 times (the known-size kernel "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here "
              f"(the product is a single-GPU kernel path: shard A's tile-rows over GPUs and run it on each)",
              file=sys.stderr)
        sys.exit(2)

    import gpu_spgemm_amd as g
    if args.workload == "batch512":
        return bench_batch(g, args)
    if args.workload == "galerkin":
        return bench_galerkin(g, args)
    A, B, cfg = make_workload(g, args.workload)
    products = A.products(B)
    log(f"workload {cfg['workload']}: products={products}")
    ctx = g.Context(0)
    log("device", ctx.device_name())
    tA = ctx.tiles_from_csr(A)
    tB = ctx.tiles_from_csr(B, as_b=True)
    s = 4 if cfg["dtype"] == "f32" else 8
    rmat = args.workload.startswith("rmat")

    def measure(label, **kw):
        for _ in range(args.warmup):
            C, st = ctx.spgemm(tA, tB, **kw)
            C.free()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            C, st = ctx.spgemm(tA, tB, **kw)
            C.free()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        br = []
        for _ in range(max(5, min(args.steps, 20))):   # per-step HIP-event breakdown on the library stream
            C, st = ctx.spgemm(tA, tB, timing=True, **kw)
            br.append(st)
            Cd = {"tilem": C.s.tilem, "tilen": C.s.tilen, "numtile": C.s.numtile, "nnz": C.s.nnz,
                  "numeric_kernels": st["numeric_kernels"]}
            C.free()
        ev = {k: statistics.mean(x[k] for x in br) for k in ("ms_step1", "ms_step2", "ms_step3")}
        log(f"{label}: {ms:.4f} ms/step  events: structure {ev['ms_step1']:.4f}  symbolic {ev['ms_step2']:.4f}"
            f"  numeric {ev['ms_step3']:.4f} ms")
        return ms, ev, Cd

    Ad = {"m": tA.s.m, "tilem": tA.s.tilem, "tilen": tA.s.tilen, "numtile": tA.s.numtile, "nnz": tA.s.nnz}
    Bd = {"m": tB.s.m, "tilem": tB.s.tilem, "tilen": tB.s.tilen, "numtile": tB.s.numtile, "nnz": tB.s.nnz}
    runs = {}
    if not rmat:   # R-MAT compares the learned scheduler with the exact binned pipeline
        runs["ctile"] = measure("reference C-tile method", variant=1)
    runs["plain"] = measure("plain (binned, exact sizes)", variant=0)
    if rmat:   # heavy sparse rows: the row-by-row numeric step with products added in any order
        runs["any_order"] = measure("plain, any-order sums (sum_order=1)", variant=0, sum_order=1)
    if args.workload == "band":     # full tiles: matrix cores for the dense pairs
        runs["mfma"] = measure("C-tile + matrix cores", variant=2)
    learned = None
    if not args.no_learned:
        train_learned(g, ctx, log, int(args.workload[4:]) if rmat else 0)
        t0 = time.perf_counter()
        C0, st = ctx.spgemm(tA, tB, scheduler=g.SCHED_LEARNED)   # the network evaluated on this product
        ctx.synchronize()
        first_ms = (time.perf_counter() - t0) * 1e3
        C0.free()
        log(f"learned: first call {first_ms:.3f} ms, fallback {st['fallback_used']}, predicted tiles "
            f"{st['pred_numtileC']} (tiles {st['numtileC']}), value capacity {st['pred_nnzC']} (nnz {st['nnzC']})")
        runs["learned"] = measure("learned (predictions reused)", scheduler=g.SCHED_LEARNED)
        learned = {"ms_per_step": round(runs["learned"][0], 5),
                   "vs_plain": round(runs["plain"][0] / runs["learned"][0], 3),
                   "fallback": st["fallback_used"], "first_call_ms": round(first_ms, 4),
                   "value_capacity_over_nnz": round(st["pred_nnzC"] / max(1, st["nnzC"]), 4),
                   "predicted_tiles_over_tiles": round(st["pred_numtileC"] / max(1, st["numtileC"]), 4)}
    # the headline is the exact plain pipeline; the other schedulers are reported beside it
    ms, ev, Cd = runs["plain"]
    step1, step2, step3 = ev["ms_step1"], ev["ms_step2"], ev["ms_step3"]
    gflops = 2.0 * products / (ms * 1e6)
    pipe_bytes = alg_bytes(Ad, Bd, Cd, s)
    # the numeric kernels the library launched (tsg_stats.numeric_kernels): when
    # one kernel is the whole numeric step, its HIP-event time is step3
    names = [NUMERIC_KERNELS[b] for b in sorted(NUMERIC_KERNELS) if Cd["numeric_kernels"] & b]
    vt = "double" if s == 8 else "float"
    if len(names) == 1:
        kname = names[0].replace("<", f"<{vt},") if "<" in names[0] else f"{names[0]}<{vt}>"
    elif not names:  # the C-tile kernels (dense tiles take the matrix-core ones)
        kname = "numeric step (C-tile kernels)"
    else:
        kname = "numeric step (" + ", ".join(names) + ")"
    nbytes = rows_numeric_bytes(Ad, Bd, Cd, s) if names == ["k_num_rows"] else numeric_bytes(Ad, Bd, Cd, s)
    achieved = nbytes / (step3 * 1e-3) / 1e9
    out = {
        "metric": "SpGEMM GFLOP/s + HBM GB/s, learned/batched vs plain tiled, 1 MI355X",
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": cfg["dtype"],
        "data": "synthetic (generated on the host, converted to tiles on the GPU before timing)",
        "config": {**cfg, "scheduler": "plain", "products": products,
                   "numtileC": Cd["numtile"], "nnzC": Cd["nnz"], "parallelism": "single GPU"},
        "pipeline_alg_GBps": round(pipe_bytes / (ms * 1e-3) / 1e9, 2),
        "pipeline_alg_bytes": pipe_bytes,
        "step_ms": {"structure": round(step1, 5), "symbolic": round(step2, 5), "numeric": round(step3, 5)},
        "variants_ms_per_step": {k: round(v[0], 5) for k, v in runs.items()},
        "variants_gflops": {k: round(2.0 * products / (v[0] * 1e6), 3) for k, v in runs.items()},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(cfg["workload"], kname), "alg_bytes_per_launch": nbytes,
                     "launch_ms": round(step3, 5)},
        "cpu_baseline": None,
    }
    if learned:
        out["learned"] = learned
        out["learned_vs_plain"] = learned["vs_plain"]
        out["learned_fallback"] = learned["fallback"]
    if not args.no_batch_leg and not rmat:
        out["batched"] = batch_leg(g, ctx, args)
        out["batched_vs_loop"] = out["batched"]["vs_loop"]
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(A, B, args.cpu_seconds)
        out["cpu_baseline_parallel"] = cpu_baseline(A, B, args.cpu_seconds, nthreads=host_threads())
    ctx.close()
    print(json.dumps(out), flush=True)


def batch_leg(g, ctx, args, count=64, hi=16384):
    """Batched vs a loop of single calls on a fixed small uneven batch (the
    batch generator's first `count` problems at seed 4 with n_i up to `hi`),
    same warmup and steps as the headline."""
    probs = g.gen_batch_uneven(count, seed=4, hi=hi)
    products = sum(A.products(A) for _, A in probs)
    tA = [ctx.tiles_from_csr(A) for _, A in probs]
    tB = [ctx.tiles_from_csr(A, as_b=True) for _, A in probs]

    def loop():
        for a, b in zip(tA, tB):
            C, _ = ctx.spgemm(a, b)
            C.free()

    def batched():
        Cs, _ = ctx.spgemm_batched(tA, tB)
        for C in Cs:
            C.free()

    G = ctx.batch_graph(tA, tB, slack=1.25)
    fallbacks = []

    def graph():  # one graph launch, then the wait for its result row
        G.launch()
        _, sts = G.result()
        fallbacks.append(sts[0]["fallback_used"])

    res = {}
    for label, fn in (("loop", loop), ("batched", batched), ("graph", graph)):
        for _ in range(args.warmup):
            fn()
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.synchronize()
        res[label] = (time.perf_counter() - t1) * 1e3 / args.steps
    log(f"batch of {count} (n_i <= {hi}): loop {res['loop']:.3f} ms, batched {res['batched']:.3f} ms, "
        f"graph {res['graph']:.3f} ms")
    G.destroy()
    for t in tA + tB:
        t.free()
    return {"problems": count, "max_rows": hi, "seed": 4, "products": products,
            "ms_loop": round(res["loop"], 4), "ms_batched": round(res["batched"], 4),
            "ms_graph": round(res["graph"], 4),
            "vs_loop": round(res["loop"] / res["batched"], 3),
            "graph_vs_loop": round(res["loop"] / res["graph"], 3),
            "graph_vs_batched": round(res["batched"] / res["graph"], 3),
            "graph_fallbacks": int(sum(fallbacks)),
            "batched_gflops": round(2.0 * products / (res["batched"] * 1e6), 3),
            "graph_gflops": round(2.0 * products / (res["graph"] * 1e6), 3)}


if __name__ == "__main__":
    main()
