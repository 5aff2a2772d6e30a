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
# HBM bytes per pr_gather launch measured by rocprofv3 (FETCH_SIZE and
# WRITE_SIZE, scaled by a copy of known size); written with the profile summary
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pr_gather_rmat22_traffic.json")


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
    kw = dict(damping=0.85, tol=1e-7, max_iter=100, items_per_thread=args.ipt,
              relabel=args.relabel, row_block=args.row_block)
    # first call: builds the per-graph caches (1/outdeg, out-degree relabelling,
    # merge-path tiles) and autotunes the gather shape; reported, not timed
    torch.cuda.synchronize()
    t = time.perf_counter()
    gb.pagerank(g, use_graph=False, time_kernels=True, **kw)
    torch.cuda.synchronize()
    prep_ms = (time.perf_counter() - t) * 1e3
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
    # hipGraph-captured form, same solve (reported beside the headline), after
    # one untimed call of its own
    gb.pagerank(g, use_graph=True, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ng = max(1, min(args.steps, 5))
    for _ in range(ng):
        xg, resg = gb.pagerank(g, use_graph=True, **kw)
    torch.cuda.synchronize()
    graph_step_s = (time.perf_counter() - t) / ng
    out = dict(V=V, E=E, iterations=it, residual=res["residual"], step_s=step_s,
               first_call_ms=prep_ms, items_per_thread=res["items_per_thread"],
               hot=res["hot"], relabeled=res["relabeled"], row_block=res["row_block"],
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
                relaxations=st["relax_edges"], mismatches_vs_dijkstra=bad)


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


def _timed(torch, fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


def run_learned_c5(gb, oracle, torch, np, cpu_budget):
    """R-MAT 24 (seed 99, weights 50/50 uniform [1,255] / log-uniform
    [1,65535]), 32 SSSP sources + PageRank to 1e-7: the plain GPU form (one
    source per solve, delta = max_w / avg_degree; PageRank from 1/N) against
    the learned + batched form (one 32-source solve with the predicted delta
    and light threshold; PageRank from the predicted warm start).  Model
    inference is inside the learned timings."""
    from graphbench_amd import learned
    models = learned.LearnedModels.load()
    t0 = time.time()
    h = gb.gen_rmat(24, 16, seed=99, weights="mix", wlo=1, whi=255)
    r = h.transpose()
    log(f"c5: generated R-MAT 24: V={h.V} E={h.E} in {time.time() - t0:.1f}s")
    srcs = gb.pick_sources(h, 32, 99)
    g = gb.Graph.from_host(h, reverse=r)
    V, E = h.V, h.E
    plain = learned.plain_delta(h)
    gb.sssp(g, srcs[:1], delta=plain, light_max=plain)  # warm-up
    d_plain = torch.empty((len(srcs), V), dtype=torch.int32, device="cuda")

    def plain_sssp():
        tr = 0
        for i, s in enumerate(srcs):
            _, st = gb.sssp(g, [s], delta=plain, light_max=plain, out=d_plain[i:i + 1], stats=True)
            tr += st["traversed"]
        return tr
    traversed, t_sp = _timed(torch, plain_sssp)

    def learned_sssp():
        dl, lm = models.predict_delta(g, len(srcs))
        out, st = gb.sssp(g, srcs, delta=dl, light_max=lm, stats=True)
        return out, st, dl, lm
    (d_learn, st_l, dl, lm), t_sl = _timed(torch, learned_sssp)
    exact = bool(torch.equal(d_plain, d_learn))
    del d_plain, d_learn
    pr_kw = dict(damping=0.85, tol=1e-7, max_iter=100)
    gb.pagerank(g, max_iter=3)  # warm-up (autotunes the tile size)
    (x_p, res_p), t_pp = _timed(torch, lambda: gb.pagerank(g, **pr_kw))

    def learned_pr():
        x0 = models.warm_start(g)
        return gb.pagerank(g, x0=x0, **pr_kw)
    (x_w, res_w), t_pw = _timed(torch, learned_pr)
    xp = x_p.cpu().numpy().astype(np.float64)
    xw = x_w.cpu().numpy().astype(np.float64)
    l1 = {"plain_vs_warm": float(np.abs(xp - xw).sum())}
    oracle.set_threads(min(16, os.cpu_count() or 1))
    tq = time.perf_counter()
    oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), 0.85, 1e-10, 1)
    if (time.perf_counter() - tq) * 40 <= 2 * cpu_budget:
        ref, _, _ = oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), 0.85, 1e-10, 150)
        l1["plain_vs_fp64_oracle"] = float(np.abs(xp - ref).sum())
        l1["warm_vs_fp64_oracle"] = float(np.abs(xw - ref).sum())
    work_pr = E * res_p["iterations"]  # the plain solve's edge work, credited to both forms
    plain_s = t_sp + t_pp
    learned_s = t_sl + t_pw
    return dict(
        workload="learned_rmat24_mix_32src_plus_pagerank", V=V, E=E,
        sssp={"plain_ms": t_sp * 1e3, "learned_batched_ms": t_sl * 1e3,
              "plain_delta": plain, "learned_delta": dl, "learned_light_max": lm,
              "plain_gteps": traversed / t_sp / 1e9, "learned_batched_gteps": traversed / t_sl / 1e9,
              "speedup": t_sp / t_sl, "distances_bit_identical": exact,
              "buckets": st_l["buckets"], "phases": st_l["phases"],
              "relaxations": st_l["relax_edges"], "traversed": traversed},
        pagerank={"plain_ms": t_pp * 1e3, "warm_ms": t_pw * 1e3,
                  "plain_iterations": res_p["iterations"], "warm_iterations": res_w["iterations"],
                  "plain_gteps": work_pr / t_pp / 1e9, "warm_effective_gteps": work_pr / t_pw / 1e9,
                  "speedup": t_pp / t_pw, "l1": l1},
        combined={"plain_gteps": (traversed + work_pr) / plain_s / 1e9,
                  "learned_batched_gteps": (traversed + work_pr) / learned_s / 1e9,
                  "speedup": plain_s / learned_s})


def c4_specs(n=512, seed=2024):
    """512 directed graphs, V = round(2^(10 + 10 U)) (log-uniform in [1024,
    1048576]); even indices R-MAT with edge factor U{8..32}, odd indices
    Erdos-Renyi with average degree U{8..32}; weights uniform [1, 1023]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        V = int(round(2.0 ** (10 + 10 * rng.random())))
        deg = int(rng.integers(8, 33))
        out.append(("rmat" if i % 2 == 0 else "er", V, deg, 5000 + i))
    return out


def run_batched_c4(gb, oracle, torch, np):
    """512 uneven graphs, one source each: one gb_sssp_multi solve with a
    learned delta per graph, against 512 plain single-graph solves (delta =
    max_w / avg_degree).  Distances of the two forms must be identical."""
    from graphbench_amd import learned
    models = learned.LearnedModels.load()
    t0 = time.time()
    hs = []
    for kind, V, deg, seed in c4_specs():
        if kind == "rmat":
            sc = max(1, (V - 1).bit_length())
            hs.append(gb.gen_rmat(sc, deg, seed=seed, weights="uniform", wlo=1, whi=1023, nverts=V))
        else:
            hs.append(gb.gen_er(V, deg, seed=seed, weights="uniform", wlo=1, whi=1023))
    sumV = sum(h.V for h in hs)
    sumE = sum(h.E for h in hs)
    log(f"c4: generated 512 graphs, sum V={sumV} sum E={sumE} in {time.time() - t0:.1f}s")
    gs = [gb.Graph.from_host(h, need_reverse=False) for h in hs]
    srcs = [int(gb.pick_sources(h, 1, 7 + i)[0]) for i, h in enumerate(hs)]
    outs_p = [torch.empty(h.V, dtype=torch.int32, device="cuda") for h in hs]
    outs_l = [torch.empty(h.V, dtype=torch.int32, device="cuda") for h in hs]
    gb.sssp(gs[0], [srcs[0]], out=outs_p[0].view(1, -1))  # warm-up

    def plain():
        tr = 0
        for g, s, o in zip(gs, srcs, outs_p):
            _, st = gb.sssp(g, [s], out=o.view(1, -1), stats=True)
            tr += st["traversed"]
        return tr
    traversed, t_p = _timed(torch, plain)

    def batched():
        dl = models.predict_deltas(gs, 1)
        _, st = gb.sssp_multi(gs, srcs, deltas=[d for d, _ in dl], light_max=[l for _, l in dl],
                              outs=outs_l, stats=True)
        return st
    st_l, t_l = _timed(torch, batched)
    exact = all(bool(torch.equal(a, b)) for a, b in zip(outs_p, outs_l))
    small = sorted(range(len(hs)), key=lambda i: hs[i].E)[:4]
    bad = sum(int((gb.u32(outs_l[i]) != oracle.dijkstra(hs[i].row_ptr, hs[i].col, hs[i].w,
                                                          srcs[i])).sum()) for i in small)
    B = sum(12 * h.E + 12 * h.V + 4 for h in hs)
    return dict(workload="sssp_512_uneven_graphs", graphs=len(hs), sum_V=sumV, sum_E=sumE,
                plain_ms=t_p * 1e3, batched_learned_ms=t_l * 1e3,
                plain_gteps=traversed / t_p / 1e9, batched_learned_gteps=traversed / t_l / 1e9,
                speedup=t_p / t_l, distances_bit_identical=exact,
                mismatches_vs_dijkstra_4_smallest=bad, buckets=st_l["buckets"],
                phases=st_l["phases"], relaxations=st_l["relax_edges"],
                achieved_gbs=B / t_l / 1e9)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ipt", type=int, default=0, help="merge-path items per thread (0=tuned)")
    ap.add_argument("--relabel", type=int, default=0,
                    help="PageRank gather in out-degree order: 0 auto, 1 on, -1 off")
    ap.add_argument("--row-block", type=int, default=0,
                    help="PageRank gather form: 0 tuned, 1024/2048/4096 row blocks, -1 merge path")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--only", choices=["sssp", "bc", "c4", "c5"], default=None,
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
        if args.only == "c5":
            print(json.dumps(run_learned_c5(gb, oracle, torch, np, args.cpu_budget)), flush=True)
            return
        fn = {"sssp": run_sssp_c1, "bc": run_bc_c3, "c4": run_batched_c4}[args.only]
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
        log("512 uneven graphs")
        extra["batched_c4"] = run_batched_c4(gb, oracle, torch, np)
        log("learned + batched vs plain, R-MAT 24")
        extra["learned_c5"] = run_learned_c5(gb, oracle, torch, np, args.cpu_budget)
    traffic = {}
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            traffic = json.load(f)
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
                     "unit": "GB/s", "frac": pr["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic.get("traffic_bytes_per_launch"),
                     "traffic_source": traffic.get("source"),
                     "kernel": traffic.get("kernel", "pr_gather"),
                     "bytes_per_launch": pr["bytes_per_launch"],
                     "mean_launch_ms": pr["gather_ms_avg"],
                     "shape": {"items_per_thread": pr["items_per_thread"], "hot": pr["hot"],
                               "row_block": pr["row_block"], "relabeled": pr["relabeled"]},
                     "first_call_ms": pr["first_call_ms"]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
