"""Parity at the benchmark sizes, through size-independent properties.

* SSSP, R-MAT 20 (V = 1,048,576, E = 16.8 M) with the 50/50 uniform /
  log-uniform weight mix and 32 sources: the batched engine equals the
  per-source engine bit for bit; for every source and edge d[v] <= d[u] + w
  (no edge can still relax), every reached vertex other than the source has
  a tight in-edge d[u] + w == d[v], and unreached vertices have no reached
  in-neighbour; two sources are compared with Dijkstra (oracle).
* PageRank, R-MAT 22 (the headline graph): L1 to the FP64 oracle <= 1e-6 and
  the ranks sum to 1.
* Brandes, R-MAT 20 symmetrised with 32 sources: relative error to the
  oracle (the CPU oracle takes a few seconds for 32 sources).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[fullsize] {msg}", flush=True)


def test_sssp_rmat20_mix_32_sources(gpu, monkeypatch):
    import torch
    gb = gpu
    h = gb.gen_rmat(20, 16, seed=99, weights="mix", wlo=1, whi=255)
    srcs = gb.pick_sources(h, 32, 5)
    g = gb.Graph.from_host(h, need_reverse=False)
    _say(f"sssp rmat20 V={h.V} E={h.E}")
    d = gb.sssp(g, srcs, delta=1 << 16, light_max=1 << 15)
    monkeypatch.setenv("GB_SSSP_PER_SOURCE", "1")
    d_ps = gb.sssp(g, srcs)
    monkeypatch.delenv("GB_SSSP_PER_SOURCE")
    assert torch.equal(d, d_ps)
    # properties on the GPU: distances as int64, INF as a large value
    INF = 0xFFFFFFFF
    dd = d.to(torch.int64) & 0xFFFFFFFF
    rp = torch.from_numpy(h.row_ptr.astype(np.int64)).cuda()
    src_of_edge = torch.repeat_interleave(torch.arange(h.V, device="cuda"), rp[1:] - rp[:-1])
    col = torch.from_numpy(h.col.astype(np.int64)).cuda()
    w = torch.from_numpy(h.w.astype(np.int64)).cuda()
    for t in range(len(srcs)):
        du = dd[t][src_of_edge]
        dv = dd[t][col]
        reached_u = du != INF
        # no edge can relax
        assert bool(((dv <= du + w) | ~reached_u).all()), t
        # unreached targets have no reached source
        assert bool((~reached_u | (dv != INF)).all()), t
        # every reached non-source vertex has a tight in-edge
        tight = reached_u & (du + w == dv)
        has_tight = torch.zeros(h.V, dtype=torch.bool, device="cuda")
        has_tight[col[tight]] = True
        reached = dd[t] != INF
        reached[int(srcs[t])] = False
        assert bool((has_tight | ~reached).all()), t
    for t in (0, 31):
        ref = oracle.dijkstra(h.row_ptr, h.col, h.w, int(srcs[t]))
        assert np.array_equal(gb.u32(d[t]), ref), t


def test_pagerank_rmat22_headline_graph(gpu):
    gb = gpu
    h = gb.gen_rmat(22, 16, seed=7, weights="none")
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    _say(f"pagerank rmat22 V={h.V} E={h.E}")
    x, res = gb.pagerank(g, damping=0.85, tol=1e-7, max_iter=100)
    xs = x.cpu().numpy().astype(np.float64)
    oracle.set_threads(min(16, oracle.get_threads()))
    ref, it, _ = oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), 0.85, 1e-7, 100)
    assert res["converged"] and abs(res["iterations"] - it) <= 1
    assert np.abs(xs - ref).sum() <= 1e-6
    assert abs(xs.sum() - 1.0) < 1e-5


def test_bc_rmat20_symmetrised_32_sources(gpu):
    gb = gpu
    h = gb.gen_rmat(20, 16, seed=11, weights="none", symmetric=True)
    srcs = gb.pick_sources(h, 32, 11)
    g = gb.Graph.from_host(h, need_reverse=False)
    _say(f"bc rmat20 sym V={h.V} arcs={h.E}")
    bc = gb.bc(g, srcs).cpu().numpy()
    oracle.set_threads(min(16, oracle.get_threads()))
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    rel = np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-300)
    assert float(rel.max()) <= 1e-9
