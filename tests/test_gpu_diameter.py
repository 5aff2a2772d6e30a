"""BFS, Brandes and SSSP on high-diameter graphs, against the oracle.

R-MAT graphs are a few levels deep; these are hundreds to tens of thousands:
* a 320 x 320 grid (102,400 vertices, 4-neighbour, symmetric; diameter 638);
* a path of 20,000 vertices (symmetric; depth up to 19,999), where the
  Brandes level pool of 64 sources must grow to hold up to 64 levels per
  vertex.
Bars: BFS levels and SSSP distances bit-exact; BC relative error <= 1e-9.
Reference: btwcentral/csr_bc.c:16-130 (Brandes), sssp/csr_sspath_array.c:
149-363 (delta-stepping).
"""
import numpy as np
import pytest

import oracle
from sssp_props import DeviceCSR, check_distances

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


def _say(msg):
    print(f"[diameter] {msg}", flush=True)


def grid_graph(gb, n, seed, weighted):
    idx = np.arange(n * n).reshape(n, n)
    a = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    b = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    w = None
    if weighted:
        rng = np.random.default_rng(seed)
        wh = rng.integers(1, 256, len(a)).astype(np.uint32)
        w = np.concatenate([wh, wh])
    return gb.HostCSR.from_edges(n * n, src, dst, w)


def path_graph(gb, n, seed, weighted):
    a = np.arange(n - 1)
    b = a + 1
    w = None
    if weighted:
        wh = np.random.default_rng(seed).integers(1, 256, n - 1).astype(np.uint32)
        w = np.concatenate([wh, wh])
    return gb.HostCSR.from_edges(n, np.concatenate([a, b]), np.concatenate([b, a]), w)


@pytest.fixture(scope="module")
def grid(gpu):
    return grid_graph(gpu, 320, 7, weighted=True)


def test_grid_bfs_levels(gpu, grid):
    gb = gpu
    h = grid
    srcs = np.random.default_rng(1).choice(h.V, 70, replace=False).astype(np.int32)
    _say(f"grid bfs V={h.V} E={h.E} sources={len(srcs)}")
    g = gb.Graph.from_host(h, need_reverse=False)
    lv = gb.bfs(g, srcs).cpu().numpy()
    for i, s in enumerate(srcs):
        assert np.array_equal(lv[i], oracle.bfs_levels(h.row_ptr, h.col, int(s))), i
    assert lv.max() >= 300


def test_grid_bc(gpu, grid):
    gb = gpu
    h = grid
    srcs = np.random.default_rng(2).choice(h.V, 96, replace=False).astype(np.int32)
    _say(f"grid bc V={h.V} sources={len(srcs)}")
    g = gb.Graph.from_host(h, need_reverse=False)
    bc, st = gb.bc(g, srcs, stats=True)
    bc = bc.cpu().numpy()
    oracle.set_threads(min(16, oracle.get_threads()))
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    rel = np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-300)
    _say(f"grid bc depth {st['max_depth']} max rel {rel.max():.2e}")
    assert st["max_depth"] >= 300
    assert float(rel.max()) <= BC_RTOL


def test_grid_sssp(gpu, grid, monkeypatch):
    import torch
    gb = gpu
    h = grid
    srcs = np.random.default_rng(3).choice(h.V, 40, replace=False).astype(np.int32)
    _say(f"grid sssp sources={len(srcs)}")
    g = gb.Graph.from_host(h, need_reverse=False)
    dg = DeviceCSR(torch, h)
    for scan in ("0", "1"):
        monkeypatch.setenv("GB_MSSSP_SCAN", scan)
        d = gb.sssp(g, srcs)
        for t in range(len(srcs)):
            check_distances(torch, dg, d[t], int(srcs[t]), f"grid source {t} scan {scan}")
        for t in (0, 39):
            assert np.array_equal(gb.u32(d[t]), oracle.dijkstra(h.row_ptr, h.col, h.w, int(srcs[t])))
    d1 = gb.sssp(g, srcs[:1])  # the per-source engine
    assert np.array_equal(gb.u32(d1[0]), oracle.dijkstra(h.row_ptr, h.col, h.w, int(srcs[0])))


def test_path_bc_and_bfs_pool_growth(gpu):
    """128 sources on a 20,000-vertex path: every vertex lies at up to 64
    depths per group, so the level pool must grow several times."""
    gb = gpu
    h = path_graph(gb, 20000, 0, weighted=False)
    srcs = np.random.default_rng(4).choice(h.V, 128, replace=False).astype(np.int32)
    srcs[0] = 0  # one end: depth 19,999
    _say(f"path bc V={h.V} sources={len(srcs)}")
    g = gb.Graph.from_host(h, need_reverse=False)
    bc, st = gb.bc(g, srcs, stats=True)
    assert st["max_depth"] == h.V - 1
    oracle.set_threads(min(16, oracle.get_threads()))
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    bc = bc.cpu().numpy()
    rel = np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-300)
    assert float(rel.max()) <= BC_RTOL
    # closed form for one source at an end: each interior vertex i lies on
    # the shortest paths to every vertex beyond it
    one = gb.bc(g, [0]).cpu().numpy()
    n = h.V
    assert np.array_equal(one, np.array([0.0] + [float(n - 1 - i) for i in range(1, n - 1)] + [0.0]))
    lv = gb.bfs(g, srcs[:8]).cpu().numpy()
    for i in range(8):
        assert np.array_equal(lv[i], np.abs(np.arange(n) - srcs[i]))


def test_path_sssp(gpu):
    gb = gpu
    h = path_graph(gb, 20000, 5, weighted=True)
    g = gb.Graph.from_host(h, need_reverse=False)
    srcs = np.array([0, 19999, 7000, 12345, 50, 19000, 333, 10000], np.int32)
    _say("path sssp")
    d = gb.u32(gb.sssp(g, srcs))
    for i, s in enumerate(srcs):
        assert np.array_equal(d[i], oracle.dijkstra(h.row_ptr, h.col, h.w, int(s))), i
    d1 = gb.u32(gb.sssp(g, srcs[:1]))
    assert np.array_equal(d1[0], oracle.dijkstra(h.row_ptr, h.col, h.w, 0))
