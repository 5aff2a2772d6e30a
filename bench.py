"""Benchmark of the CSR analytics path on one MI355X.

Headline workload: PageRank on directed R-MAT scale 22 (4,194,304 vertices,
edge factor 16, a/b/c/d = .57/.19/.19/.05, self-loops dropped, seed 7),
damping 0.85, dangling mass spread uniformly, x0 = 1/N, stop at L1 residual
< 1e-7 or 100 iterations.  One step = one full solve with the graph resident
in HBM.  value = GTEPS = E * iterations / step time.

The JSON line also carries the roofline of the dominant kernel (pr_gather:
algorithmic bytes per launch / mean launch time from HIP events recorded on
the launch stream inside the timed region), the CPU baseline (the oracle's
FP64 OpenMP power iteration on the host cores, same graph) and secondary
measurements for SSSP and Brandes in "extra".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pagerank_bytes(V, E):
    # row_ptr 4(V+1) + col 4E + gathered contrib 4E + x_old 4V + 1/outdeg 4V + x_new 4V
    return 8 * E + 16 * V + 4


def run_pagerank(gb, torch, np, args):
    t0 = time.time()
    h = gb.gen_rmat(22, 16, seed=7, weights="none")
    r = h.transpose()
    log(f"generated R-MAT 22: V={h.V} E={h.E} in {time.time() - t0:.1f}s")
    g = gb.Graph.from_host(h, reverse=r)
    V, E = h.V, h.E
    kw = dict(damping=0.85, tol=1e-7, max_iter=100, items_per_thread=args.ipt)
    for _ in range(args.warmup):
        gb.pagerank(g, use_graph=False, time_kernels=True, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter()
    gms, gl, iters = 0.0, 0, []
    for _ in range(args.steps):
        x, res = gb.pagerank(g, use_graph=False, time_kernels=True, **kw)
        gms += res["gather_ms"]
        gl += res["gather_launches"]
        iters.append(res["iterations"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    step_s = el / args.steps
    it = iters[-1]
    gteps = E * it / step_s / 1e9
    avg_gather_s = gms / max(gl, 1) / 1e3
    B = pagerank_bytes(V, E)
    achieved = B / avg_gather_s / 1e9
    # hipGraph-captured form, same solve (reported beside the headline)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ng = max(1, min(args.steps, 5))
    for _ in range(ng):
        xg, resg = gb.pagerank(g, use_graph=True, **kw)
    torch.cuda.synchronize()
    graph_step_s = (time.perf_counter() - t) / ng
    out = dict(V=V, E=E, iterations=it, residual=res["residual"], step_s=step_s,
               gteps=gteps, gather_ms_avg=avg_gather_s * 1e3, gather_launches=gl,
               tiles=res["tiles"], bytes_per_launch=B, achieved_gbs=achieved,
               graph_step_ms=graph_step_s * 1e3, graph_gteps=E * resg["iterations"] / graph_step_s / 1e9)
    x_host = x.cpu().numpy().astype(np.float64)
    return h, r, out, x_host


def cpu_baseline(oracle, np, h, r, x_gpu, budget_s):
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    od = h.outdeg()
    t = time.perf_counter()
    oracle.pagerank_cfg(r.row_ptr, r.col, od, 0.85, 1e-7, 2)
    per_it = (time.perf_counter() - t) / 2
    full = per_it * 100 <= budget_s
    k = 100 if full else max(2, int(budget_s / per_it))
    t = time.perf_counter()
    xr, it, rr = oracle.pagerank_cfg(r.row_ptr, r.col, od, 0.85, 1e-7, k)
    el = time.perf_counter() - t
    res = dict(value=h.E * it / el / 1e9, unit="GTEPS", cores=threads, kind="port",
               sample=(f"oracle FP64 power iteration (OpenMP, {threads} threads) on the same "
                       f"R-MAT 22 graph, {it} iterations" + (" (full solve)" if full else " (bounded sample)")),
               seconds=el)
    err = float(np.abs(x_gpu - xr).sum()) if full else None
    return res, err, it


def run_sssp_c1(gb, oracle, torch, np):
    h = gb.gen_rmat(16, 16, seed=42, weights="uniform", wlo=1, whi=255)
    srcs = gb.pick_sources(h, 8, 42)
    g = gb.Graph.from_host(h, need_reverse=False)
    gb.sssp(g, srcs)  # warm
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 5
    for _ in range(reps):
        d, st = gb.sssp(g, srcs, stats=True)
    torch.cuda.synchronize()
    s = (time.perf_counter() - t) / reps
    dd = gb.u32(d)
    bad = sum(int((dd[i] != oracle.dijkstra(h.row_ptr, h.col, h.w, int(x))).sum())
              for i, x in enumerate(srcs))
    return dict(workload="sssp_rmat16_8src", V=h.V, E=h.E, ms=s * 1e3,
                gteps=st["traversed"] / s / 1e9, buckets=st["buckets"], phases=st["phases"],
                mismatches_vs_dijkstra=bad)


def run_bc_c3(gb, oracle, torch, np):
    h = gb.gen_rmat(20, 16, seed=11, weights="none", symmetric=True)
    srcs = gb.pick_sources(h, 256, 11)
    g = gb.Graph.from_host(h, need_reverse=False)
    gb.bc(g, srcs[:8])
    torch.cuda.synchronize()
    t = time.perf_counter()
    bc, st = gb.bc(g, srcs, stats=True)
    torch.cuda.synchronize()
    s = time.perf_counter() - t
    # error vs the FP64 oracle on 4 sources (the full 256 would take minutes on the host)
    sub = srcs[:4]
    bsub = gb.bc(g, sub).cpu().numpy()
    ref = oracle.brandes_bfs(h.row_ptr, h.col, sub)
    rel = float(np.max(np.abs(bsub - ref) / np.maximum(np.abs(ref), 1e-300)))
    M = h.E
    B = 40 * M + 48 * h.V + 8
    return dict(workload="bc_rmat20_sym_256src", V=h.V, arcs=M, ms=s * 1e3,
                gteps=256 * M / s / 1e9, max_depth=st["max_depth"],
                achieved_gbs=256 * B / s / 1e9, max_rel_err_vs_oracle_4src=rel)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ipt", type=int, default=0, help="merge-path items per thread (0=tuned)")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--only", choices=["sssp", "bc"], default=None,
                    help="time one secondary workload alone (diagnostics; no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import graphbench_amd as gb
    import oracle

    gb.init(0)
    if args.only:
        fn = {"sssp": run_sssp_c1, "bc": run_bc_c3}[args.only]
        print(json.dumps(fn(gb, oracle, torch, np)), flush=True)
        return
    log("pagerank R-MAT 22")
    h, r, pr, x = run_pagerank(gb, torch, np, args)
    log(f"pagerank: {pr['step_s']*1e3:.2f} ms/step, {pr['gteps']:.1f} GTEPS, gather "
        f"{pr['gather_ms_avg']:.3f} ms, {pr['achieved_gbs']:.0f} GB/s")
    cpu, err, cpu_it = (None, None, None)
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu, err, cpu_it = cpu_baseline(oracle, np, h, r, x, args.cpu_budget)
        log(f"cpu: {cpu['value']:.3f} GTEPS, L1 vs FP64 oracle {err}")
    extra = {"pagerank_graph_mode": {"ms_per_step": pr["graph_step_ms"], "gteps": pr["graph_gteps"]},
             "pagerank_l1_vs_fp64_oracle": err, "pagerank_iterations": pr["iterations"],
             "pagerank_residual": pr["residual"], "oracle_iterations": cpu_it}
    if not args.no_extra:
        log("sssp config 1")
        extra["sssp_c1"] = run_sssp_c1(gb, oracle, torch, np)
        log("bc config 3")
        extra["bc_c3"] = run_bc_c3(gb, oracle, torch, np)
    line = {
        "metric": "GTEPS + % HBM peak (SSSP/PR/BC); learned-delta+batched vs plain, same GPU",
        "value": pr["gteps"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pr["step_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic R-MAT (generated on the host, resident in HBM)",
        "config": {"workload": "pagerank_rmat22", "V": pr["V"], "E": pr["E"], "edge_factor": 16,
                   "seed": 7, "damping": 0.85, "tol": 1e-7, "max_iter": 100,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": pr["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": pr["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "pr_gather", "bytes_per_launch": pr["bytes_per_launch"],
                     "mean_launch_ms": pr["gather_ms_avg"]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
