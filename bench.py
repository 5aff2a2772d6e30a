"""Benchmark of the CSR analytics path on one MI355X.

Headline workload: PageRank on directed R-MAT scale 22 (4,194,304 vertices,
edge factor 16, a/b/c/d = .57/.19/.19/.05, self-loops dropped, seed 7),
damping 0.85, dangling mass spread uniformly, x0 = 1/N, stop at L1 residual
< 1e-7 or 100 iterations.  One step = one full solve with the graph resident
in HBM.  value = GTEPS = E * iterations / step time.

The JSON line also carries the roofline of the dominant kernel (pr_gather:
algorithmic bytes per launch / mean launch time from HIP events recorded on
the launch stream inside the timed region), the CPU baseline (the oracle's
FP64 OpenMP power iteration on the host cores, same graph, median of 5) and,
in "extra", the other four workloads (graphbench_amd/workloads.py), each with
a CPU baseline of the oracle on a bounded sample of the same work.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per pr_gather launch measured by rocprofv3 (FETCH_SIZE and
# WRITE_SIZE, scaled by a copy of known size); written with the profile summary
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pr_gather_rmat22_traffic.json")
# HBM-side bytes of one learned + batched 32-source solve at R-MAT 24, from
# copy-calibrated FETCH_SIZE / WRITE_SIZE runs (rocprofv3 --pmc)
SSSP_TRAFFIC_JSON = os.path.join(ROOT, "profiles", "sssp_rmat24_32src_traffic.json")
# HBM-side bytes of the timed 256-source Brandes solve at R-MAT 20, likewise
BC_TRAFFIC_JSON = os.path.join(ROOT, "profiles", "bc_rmat20_256src_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host cores this process may run on: logical CPUs in its affinity
    mask, the physical cores among them, the CPU model, and the thread count
    the CPU baselines use (one per physical core, within OMP_NUM_THREADS when
    the environment sets it)."""
    aff = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in aff:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "core_id") as f1, open(base + "physical_package_id") as f2:
                cores.add((f2.read().strip(), f1.read().strip()))
        except OSError:
            cores.add(("cpu", c))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = len(cores)
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return dict(nproc=os.cpu_count(), affinity_cpus=len(aff), physical_cores=len(cores),
                model=model, omp_num_threads=omp or None, threads=max(1, threads))


def median_time(fn, runs=5, warmup=1):
    """Median wall time of `runs` calls after `warmup` untimed ones; returns
    (median seconds, all times, last result)."""
    import numpy as np
    out = None
    for _ in range(warmup):
        out = fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), ts, out


def cpu_record(value, unit, cpu, threads, sample, runs, **kw):
    rec = dict(value=value, unit=unit, cores=threads, kind="port", sample=sample,
               timing=f"median of {runs} runs after 1 untimed run" if runs > 1 else "one timed run",
               host={k: cpu[k] for k in ("nproc", "affinity_cpus", "physical_cores", "model",
                                         "omp_num_threads")})
    rec.update(kw)
    return rec


def cpu_sssp(oracle, np, graphs, threads, runs, warmup):
    """The oracle's delta-stepping (run_sspath_arr restated), one (graph,
    source, delta) job per host thread; returns (median s, traversed edges)."""
    def one(job):
        h, s, delta = job
        return oracle.delta_stepping(h.row_ptr, h.col, h.w, int(s), int(delta))

    def run():
        with ThreadPoolExecutor(threads) as ex:
            return list(ex.map(one, graphs))
    med, ts, outs = median_time(run, runs, warmup)
    trav = 0
    for (h, _, _), d in zip(graphs, outs):
        trav += int(h.outdeg()[d != oracle.UINF].astype(np.int64).sum())
    return med, trav


# (hub chunk, log2 chunk width, log2 bin width) of the binned PageRank shapes
PB_SHAPES = {1: (8192, 13, 13), 2: (16384, 14, 13), 3: (8192, 13, 12), 4: (16384, 13, 13),
             5: (20480, 13, 13), 6: (16384, 12, 13)}


def gather_kernel_name(pr):
    """The gather kernel(s) of the form the solve ran, as rocprofv3 names them."""
    if pr["binned"]:
        hub, csl, bsl = PB_SHAPES[pr["binned"]]
        return f"pb_scatter_kernel<{csl}> + pb_accum_kernel<{hub}, {bsl}>"
    if pr["row_block"]:
        if pr["hot"]:
            return f"pr_rowblock_hot_kernel<{pr['row_block']}, {pr['hot']}>"
        return f"pr_rowblock_kernel<{pr['row_block']}>"
    return f"pr_gather_kernel<{pr['items_per_thread']}, {pr['hot']}>"


def pagerank_bytes(V, E):
    # row_ptr 4(V+1) + col 4E + gathered contrib 4E + x_old 4V + 1/outdeg 4V + x_new 4V
    return 8 * E + 16 * V + 4


def run_pagerank(gb, torch, np, args):
    from graphbench_amd import workloads
    t0 = time.time()
    h = workloads.c2()
    r = h.transpose()
    log(f"generated R-MAT 22: V={h.V} E={h.E} in {time.time() - t0:.1f}s")
    g = gb.Graph.from_host(h, reverse=r)
    V, E = h.V, h.E
    kw = dict(damping=0.85, tol=1e-7, max_iter=100, items_per_thread=args.ipt,
              relabel=args.relabel, row_block=args.row_block, binned=args.binned)
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
               binned=res["binned"], gteps=gteps, gather_ms_avg=avg_gather_s * 1e3, gather_launches=gl,
               tiles=res["tiles"], bytes_per_launch=B, achieved_gbs=achieved,
               graph_step_ms=graph_step_s * 1e3, graph_gteps=E * resg["iterations"] / graph_step_s / 1e9)
    x_host = x.cpu().numpy().astype(np.float64)
    return h, r, out, x_host


def cpu_baseline(oracle, np, h, r, x_gpu, budget_s, cpu):
    """The oracle's FP64 power iteration (OpenMP over rows, deterministic
    chunked reductions) on the same graph: a full solve when 6 of them fit the
    budget (median of 5 after 1 warm-up), else a bounded number of
    iterations."""
    threads = cpu["threads"]
    oracle.set_threads(threads)
    od = h.outdeg()
    t = time.perf_counter()
    oracle.pagerank_cfg(r.row_ptr, r.col, od, 0.85, 1e-7, 2)
    per_it = (time.perf_counter() - t) / 2
    full = per_it * 100 * 6 <= budget_s
    k = 100 if full else max(2, int(budget_s / 6 / per_it))
    med, ts, out = median_time(lambda: oracle.pagerank_cfg(r.row_ptr, r.col, od, 0.85, 1e-7, k))
    xr, it, rr = out
    res = cpu_record(h.E * it / med / 1e9, "GTEPS", cpu, threads,
                     (f"oracle FP64 power iteration (OpenMP, {threads} threads) on the same "
                      f"R-MAT 22 graph, {it} iterations" + (" (full solve)" if full else " (bounded sample)")),
                     5, seconds=med, runs_s=ts)
    err = float(np.abs(x_gpu - xr).sum()) if full else None
    return res, err, it


def run_sssp_c1(gb, oracle, torch, np, cpu, with_cpu):
    from graphbench_amd import learned, workloads
    h, srcs = workloads.c1()
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
    out = dict(workload="sssp_rmat16_8src", V=h.V, E=h.E, ms=s * 1e3,
               gteps=st["traversed"] / s / 1e9, buckets=st["buckets"], phases=st["phases"],
               relaxations=st["relax_edges"], mismatches_vs_dijkstra=bad)
    if with_cpu:
        plain = learned.plain_delta(h)
        threads = min(cpu["threads"], len(srcs))
        med, trav = cpu_sssp(oracle, np, [(h, x, plain) for x in srcs], threads, 5, 1)
        out["cpu"] = cpu_record(trav / med / 1e9, "GTEPS", cpu, threads,
                                f"oracle delta-stepping (delta {plain}), the same 8 sources, "
                                f"one source per thread", 5, seconds=med)
    return out


def run_bc_c3(gb, oracle, torch, np, cpu, with_cpu):
    from graphbench_amd import workloads
    h, srcs = workloads.c3()
    g = gb.Graph.from_host(h, need_reverse=False)
    # warm-up with the timed call's sources: the first batch of each size runs
    # the launch-shape autotuner
    gb.bc(g, srcs)
    torch.cuda.synchronize()
    t = time.perf_counter()
    bc, st = gb.bc(g, srcs, stats=True)
    torch.cuda.synchronize()
    s = time.perf_counter() - t
    # error vs the FP64 oracle on 4 sources (the 256-source comparison is a GPU test)
    sub = srcs[:4]
    bsub = gb.bc(g, sub).cpu().numpy()
    oracle.set_threads(cpu["threads"])
    ref = oracle.brandes_bfs(h.row_ptr, h.col, sub)
    rel = float(np.max(np.abs(bsub - ref) / np.maximum(np.abs(ref), 1e-300)))
    M = h.E
    B = 40 * M + 48 * h.V + 8
    out = dict(workload="bc_rmat20_sym_256src", V=h.V, arcs=M, ms=s * 1e3,
               gteps=256 * M / s / 1e9, max_depth=st["max_depth"],
               # 256 x the one-source byte count over the time: what 256
               # separate single-source solves would have to move, not a
               # bandwidth (64 sources share each col and mask read)
               per_source_equivalent_gbs=256 * B / s / 1e9, max_rel_err_vs_oracle_4src=rel)
    if os.path.exists(BC_TRAFFIC_JSON):
        with open(BC_TRAFFIC_JSON) as f:
            tr = json.load(f)
        # measured HBM-side bytes of this solve over its time: the bandwidth
        out["traffic"] = {"bytes_per_solve": tr["traffic_bytes_per_solve"],
                          "gbs": tr["traffic_bytes_per_solve"] / s / 1e9,
                          "frac": tr["traffic_bytes_per_solve"] / s / 1e9 / HBM_PEAK_GBS,
                          "source": tr["source"]}
    if with_cpu:
        threads = cpu["threads"]
        ns = min(len(srcs), 2 * threads)
        sample = srcs[:ns]
        oracle.set_threads(threads)
        med, _, _ = median_time(lambda: oracle.brandes_bfs(h.row_ptr, h.col, sample), 5, 1)
        out["cpu"] = cpu_record(ns * M / med / 1e9, "GTEPS", cpu, threads,
                                f"oracle Brandes (FP64), the first {ns} of the 256 sources, "
                                f"one source per thread", 5, seconds=med)
    return out


def _timed(torch, fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


def sssp_batch_bytes(V, E, k):
    """Algorithmic bytes of one batched solve of k sources sharing the CSR:
    row_ptr 4(V+1), col + w 8E, and per source one dist access per edge (4E)
    plus dist init and the source-side read (8V)."""
    return 4 * (V + 1) + 8 * E + k * (4 * E + 8 * V)


def run_learned_c5(gb, oracle, torch, np, cpu, with_cpu):
    """R-MAT 24 (seed 99, weights 50/50 uniform [1,255] / log-uniform
    [1,65535]), 32 SSSP sources + PageRank to 1e-7, three forms on the same
    GPU: plain (one source per solve with the per-source engine, delta =
    max_w / avg_degree; PageRank from 1/N), batched with the plain delta (one
    32-source solve of the batched engine), and learned + batched (the same
    solve with the predicted delta; PageRank from the predicted warm start).
    Model inference is inside the learned timings."""
    from graphbench_amd import learned, workloads
    models = learned.LearnedModels.load()
    t0 = time.time()
    h, srcs = workloads.c5()
    r = h.transpose()
    log(f"c5: generated R-MAT 24: V={h.V} E={h.E} in {time.time() - t0:.1f}s")
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
    # warm-up of the batched engine with the timed call's sources and delta:
    # its shape autotuner runs on the first solve of each (group width, delta)
    gb.sssp(g, srcs, delta=plain, light_max=plain)
    (d_bp, st_bp), t_sbp = _timed(torch, lambda: gb.sssp(g, srcs, delta=plain, light_max=plain,
                                                          stats=True))
    exact_bp = bool(torch.equal(d_plain, d_bp))
    del d_bp

    def learned_sssp():
        dl, lm = models.predict_delta(g, len(srcs))
        out, st = gb.sssp(g, srcs, delta=dl, light_max=lm, stats=True)
        return out, st, dl, lm
    learned_sssp()  # warm-up (the shape autotuner for the predicted delta)
    (d_learn, st_l, dl, lm), t_sl = _timed(torch, learned_sssp)
    exact = bool(torch.equal(d_plain, d_learn))
    del d_plain, d_learn
    B32 = sssp_batch_bytes(V, E, len(srcs))
    sssp_traffic = {}
    if os.path.exists(SSSP_TRAFFIC_JSON):
        with open(SSSP_TRAFFIC_JSON) as f:
            sssp_traffic = json.load(f)
    pr_kw = dict(damping=0.85, tol=1e-7, max_iter=100)
    gb.pagerank(g, max_iter=3)  # warm-up (autotunes the gather)
    (x_p, res_p), t_pp = _timed(torch, lambda: gb.pagerank(g, **pr_kw))

    def learned_pr():
        x0 = models.warm_start(g)
        return gb.pagerank(g, x0=x0, **pr_kw)
    (x_w, res_w), t_pw = _timed(torch, learned_pr)
    xp = x_p.cpu().numpy().astype(np.float64)
    xw = x_w.cpu().numpy().astype(np.float64)
    l1 = {"plain_vs_warm": float(np.abs(xp - xw).sum())}
    oracle.set_threads(cpu["threads"])
    t_cpu_pr = None
    tq = time.perf_counter()
    ref, it_ref, _ = oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), 0.85, 1e-7, 100)
    t_cpu_pr = time.perf_counter() - tq
    l1["plain_vs_fp64_oracle"] = float(np.abs(xp - ref).sum())
    l1["warm_vs_fp64_oracle"] = float(np.abs(xw - ref).sum())
    work_pr = E * res_p["iterations"]  # the plain solve's edge work, credited to both forms
    plain_s = t_sp + t_pp
    learned_s = t_sl + t_pw
    out = dict(
        workload="learned_rmat24_mix_32src_plus_pagerank", V=V, E=E,
        sssp={"plain_ms": t_sp * 1e3, "batched_plain_delta_ms": t_sbp * 1e3,
              "learned_batched_ms": t_sl * 1e3,
              "plain_delta": plain, "learned_delta": dl, "learned_light_max": lm,
              "plain_gteps": traversed / t_sp / 1e9,
              "batched_plain_delta_gteps": traversed / t_sbp / 1e9,
              "learned_batched_gteps": traversed / t_sl / 1e9,
              "speedup": t_sp / t_sl, "batching_speedup": t_sp / t_sbp,
              "learned_over_batched_plain": t_sbp / t_sl,
              "distances_bit_identical": exact and exact_bp,
              "buckets": st_l["buckets"], "phases": st_l["phases"],
              "relaxations": st_l["relax_edges"], "traversed": traversed,
              "batched_plain_delta": {"buckets": st_bp["buckets"], "phases": st_bp["phases"],
                                      "relaxations": st_bp["relax_edges"]},
              "roofline": {"bytes": B32, "achieved_gbs": B32 / t_sl / 1e9,
                           "frac": B32 / t_sl / 1e9 / HBM_PEAK_GBS,
                           "traffic": sssp_traffic.get("traffic_bytes_per_solve"),
                           "traffic_over_bytes": (sssp_traffic["traffic_bytes_per_solve"] / B32
                                                  if sssp_traffic else None),
                           "traffic_gbs": (sssp_traffic["traffic_bytes_per_solve"] / t_sl / 1e9
                                           if sssp_traffic else None),
                           "traffic_source": sssp_traffic.get("source"),
                           "note": "B(32) = 4(V+1) + 8E + 32(4E + 8V) over the learned + "
                                   "batched solve time; traffic = HBM-side bytes of one such "
                                   "solve from the profile named"}},
        pagerank={"plain_ms": t_pp * 1e3, "warm_ms": t_pw * 1e3,
                  "plain_iterations": res_p["iterations"], "warm_iterations": res_w["iterations"],
                  "plain_gteps": work_pr / t_pp / 1e9, "warm_effective_gteps": work_pr / t_pw / 1e9,
                  "speedup": t_pp / t_pw, "l1": l1},
        combined={"plain_gteps": (traversed + work_pr) / plain_s / 1e9,
                  "learned_batched_gteps": (traversed + work_pr) / learned_s / 1e9,
                  "speedup": plain_s / learned_s})
    if with_cpu:
        threads = cpu["threads"]
        ns = min(len(srcs), threads)
        med, trav = cpu_sssp(oracle, np, [(h, x, plain) for x in srcs[:ns]], threads, 1, 0)
        out["cpu"] = {
            "sssp": cpu_record(trav / med / 1e9, "GTEPS", cpu, threads,
                               f"oracle delta-stepping (delta {plain}), the first {ns} of the 32 "
                               f"sources, one source per thread", 1, seconds=med),
            "pagerank": cpu_record(E * it_ref / t_cpu_pr / 1e9, "GTEPS", cpu, threads,
                                   f"oracle FP64 power iteration, full solve ({it_ref} iterations)",
                                   1, seconds=t_cpu_pr)}
    return out


def run_batched_c4(gb, oracle, torch, np, cpu, with_cpu):
    """512 uneven graphs, one source each, three forms: 512 plain
    single-graph solves (delta = max_w / avg_degree), one gb_sssp_multi solve
    with the plain delta per graph, and one with a learned delta per graph.
    Distances of the three forms must be identical."""
    from graphbench_amd import learned, workloads
    models = learned.LearnedModels.load()
    t0 = time.time()
    hs, srcs = workloads.c4()
    sumV = sum(h.V for h in hs)
    sumE = sum(h.E for h in hs)
    log(f"c4: generated 512 graphs, sum V={sumV} sum E={sumE} in {time.time() - t0:.1f}s")
    gs = [gb.Graph.from_host(h, need_reverse=False) for h in hs]
    outs_p = [torch.empty(h.V, dtype=torch.int32, device="cuda") for h in hs]
    outs_b = [torch.empty(h.V, dtype=torch.int32, device="cuda") for h in hs]
    outs_l = [torch.empty(h.V, dtype=torch.int32, device="cuda") for h in hs]
    gb.sssp(gs[0], [srcs[0]], out=outs_p[0].view(1, -1))  # warm-up

    def plain():
        tr = 0
        for g, s, o in zip(gs, srcs, outs_p):
            _, st = gb.sssp(g, [s], out=o.view(1, -1), stats=True)
            tr += st["traversed"]
        return tr
    traversed, t_p = _timed(torch, plain)
    # each lockstep form once untimed: the first solve per graph and light
    # threshold builds that graph's light/heavy split (cached on the handle)
    _, t_b0 = _timed(torch, lambda: gb.sssp_multi(gs, srcs, outs=outs_b))
    st_b, t_b = _timed(torch, lambda: gb.sssp_multi(gs, srcs, outs=outs_b, stats=True)[1])

    def batched():
        dl = models.predict_deltas(gs, 1)
        _, st = gb.sssp_multi(gs, srcs, deltas=[d for d, _ in dl], light_max=[l for _, l in dl],
                              outs=outs_l, stats=True)
        return st
    _, t_l0 = _timed(torch, batched)
    st_l, t_l = _timed(torch, batched)
    exact = all(bool(torch.equal(a, b)) and bool(torch.equal(a, c))
                for a, b, c in zip(outs_p, outs_b, outs_l))
    small = sorted(range(len(hs)), key=lambda i: hs[i].E)[:4]
    bad = sum(int((gb.u32(outs_l[i]) != oracle.dijkstra(hs[i].row_ptr, hs[i].col, hs[i].w,
                                                          srcs[i])).sum()) for i in small)
    B = sum(12 * h.E + 12 * h.V + 4 for h in hs)
    out = dict(workload="sssp_512_uneven_graphs", graphs=len(hs), sum_V=sumV, sum_E=sumE,
               plain_ms=t_p * 1e3, batched_plain_delta_ms=t_b * 1e3, batched_learned_ms=t_l * 1e3,
               plain_gteps=traversed / t_p / 1e9, batched_plain_delta_gteps=traversed / t_b / 1e9,
               batched_learned_gteps=traversed / t_l / 1e9,
               first_call_ms={"batched_plain_delta": t_b0 * 1e3, "batched_learned": t_l0 * 1e3},
               speedup=t_p / t_l, batching_speedup=t_p / t_b, learned_over_batched_plain=t_b / t_l,
               distances_bit_identical=exact,
               mismatches_vs_dijkstra_4_smallest=bad, buckets=st_l["buckets"],
               phases=st_l["phases"], relaxations=st_l["relax_edges"],
               batched_plain_delta_stats={"buckets": st_b["buckets"], "phases": st_b["phases"],
                                          "relaxations": st_b["relax_edges"]},
               achieved_gbs=B / t_l / 1e9, frac=B / t_l / 1e9 / HBM_PEAK_GBS)
    if with_cpu:
        threads = cpu["threads"]
        jobs = [(hs[i], srcs[i], learned.plain_delta(hs[i])) for i in range(0, len(hs), 8)]
        med, trav = cpu_sssp(oracle, np, jobs, threads, 5, 1)
        out["cpu"] = cpu_record(trav / med / 1e9, "GTEPS", cpu, threads,
                                f"oracle delta-stepping (plain delta per graph) on every 8th "
                                f"graph ({len(jobs)} graphs, sum E = "
                                f"{sum(j[0].E for j in jobs)}), one graph per thread", 5,
                                seconds=med)
    return out


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
    ap.add_argument("--binned", type=int, default=0,
                    help="PageRank binned form: 0 tuned, -1 never, 1..6 that (hub, chunk, bin) shape")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of host time for the headline CPU baseline (6 solves)")
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
    cpu = host_cpu()
    with_cpu = not args.no_cpu
    if args.only:
        fn = {"sssp": run_sssp_c1, "bc": run_bc_c3, "c4": run_batched_c4,
              "c5": run_learned_c5}[args.only]
        print(json.dumps(fn(gb, oracle, torch, np, cpu, with_cpu)), flush=True)
        return
    log("pagerank R-MAT 22")
    h, r, pr, x = run_pagerank(gb, torch, np, args)
    log(f"pagerank: {pr['step_s']*1e3:.2f} ms/step, {pr['gteps']:.1f} GTEPS, gather "
        f"{pr['gather_ms_avg']:.3f} ms, {pr['achieved_gbs']:.0f} GB/s")
    cpu_pr, err, cpu_it = (None, None, None)
    if with_cpu:
        log(f"cpu baseline (oracle, {cpu['threads']} threads: {cpu['model']})")
        cpu_pr, err, cpu_it = cpu_baseline(oracle, np, h, r, x, args.cpu_budget, cpu)
        log(f"cpu: {cpu_pr['value']:.3f} GTEPS, L1 vs FP64 oracle {err}")
    extra = {"pagerank_graph_mode": {"ms_per_step": pr["graph_step_ms"], "gteps": pr["graph_gteps"]},
             "pagerank_l1_vs_fp64_oracle": err, "pagerank_iterations": pr["iterations"],
             "pagerank_residual": pr["residual"], "oracle_iterations": cpu_it}
    if not args.no_extra:
        del h, r, x
        log("sssp config 1")
        extra["sssp_c1"] = run_sssp_c1(gb, oracle, torch, np, cpu, with_cpu)
        log("bc config 3")
        extra["bc_c3"] = run_bc_c3(gb, oracle, torch, np, cpu, with_cpu)
        log("512 uneven graphs")
        extra["batched_c4"] = run_batched_c4(gb, oracle, torch, np, cpu, with_cpu)
        log("learned + batched vs plain, R-MAT 24")
        extra["learned_c5"] = run_learned_c5(gb, oracle, torch, np, cpu, with_cpu)
    traffic = {}
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            traffic = json.load(f)
    ran = gather_kernel_name(pr)
    # the counters describe one kernel: a profile of another one is not used
    if traffic.get("kernel") != ran:
        traffic = {"stale": f"{TRAFFIC_JSON} profiles {traffic.get('kernel')}; this run used {ran}"}
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
                     # what the gather's counters show bounds it (the profile
                     # named in traffic_source)
                     "limiter": traffic.get("limiter"),
                     "traffic": traffic.get("traffic_bytes_per_launch"),
                     "traffic_source": traffic.get("source", traffic.get("stale")),
                     "kernel": ran,
                     "bytes_per_launch": pr["bytes_per_launch"],
                     "mean_launch_ms": pr["gather_ms_avg"],
                     "shape": {"items_per_thread": pr["items_per_thread"], "hot": pr["hot"],
                               "row_block": pr["row_block"], "binned": pr["binned"],
                               "relabeled": pr["relabeled"]},
                     "first_call_ms": pr["first_call_ms"]},
        "cpu_baseline": cpu_pr,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
