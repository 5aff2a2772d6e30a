"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Bars: SSSP distances and BFS levels bit-exact; configured PageRank within
1e-6 L1 of the FP64 oracle; reference-semantics PageRank bit-exact with the
FP64 restatement of pagerank_mpi.c; Brandes BC within 1e-9 relative.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9
PR_L1 = 1e-6


def _say(msg):
    print(f"[case] {msg}", flush=True)


def small_graph(seed, V=400, E=3000, wlo=1, whi=255, self_loops=False):
    import graphbench_amd as gb
    rng = np.random.default_rng(seed)
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    if not self_loops:
        m = src != dst
        src, dst = src[m], dst[m]
    w = rng.integers(wlo, whi + 1, len(src)).astype(np.uint32)
    return gb.HostCSR.from_edges(V, src, dst, w)


# --------------------------------------------------------------------- SSSP

@pytest.mark.parametrize("seed,delta,light", [(1, 0, 0), (2, 16, 16), (3, 1, 1), (4, 300, 300),
                                              (5, 16, 4), (6, 8, 64)])
def test_sssp_small_bit_exact(gpu, seed, delta, light):
    gb = gpu
    h = small_graph(seed)
    _say(f"sssp small seed={seed} delta={delta} light={light}")
    g = gb.Graph.from_host(h, need_reverse=False)
    srcs = [0, 7, 123, 399]
    d = gb.u32(gb.sssp(g, srcs, delta=delta, light_max=light))
    for i, s in enumerate(srcs):
        ref = oracle.dijkstra(h.row_ptr, h.col, h.w, s)
        assert np.array_equal(d[i], ref), f"source {s}"


@pytest.mark.parametrize("scan", ["0", "1"])
@pytest.mark.parametrize("nsrc", [2, 9, 17, 33, 64, 100])
def test_sssp_batched_engine_matches_per_source(gpu, nsrc, scan, monkeypatch):
    """Several sources run in the batched engine (groups of up to 64 sources,
    widths 8..64), in both forms (appends inside the step, or the compaction
    scan between phases); distances equal the per-source engine's and
    Dijkstra's, with duplicate sources and sources without out-edges
    included."""
    gb = gpu
    monkeypatch.setenv("GB_MSSSP_SCAN", scan)
    h = gb.gen_rmat(12, 8, seed=50 + nsrc, weights="uniform", wlo=1, whi=300)
    g = gb.Graph.from_host(h, need_reverse=False)
    rng = np.random.default_rng(nsrc)
    srcs = rng.integers(0, h.V, nsrc).astype(np.int32)
    srcs[-1] = srcs[0]  # a duplicate
    zero_out = np.flatnonzero(np.diff(h.row_ptr) == 0)
    if len(zero_out) and nsrc > 2:
        srcs[1] = zero_out[0]
    _say(f"sssp batched nsrc={nsrc}")
    for delta in (0, 64, 1 << 12):
        d, st = gb.sssp(g, srcs, delta=delta, stats=True)
        d = gb.u32(d)
        monkeypatch.setenv("GB_SSSP_PER_SOURCE", "1")
        ref_engine = gb.u32(gb.sssp(g, srcs, delta=delta))
        monkeypatch.delenv("GB_SSSP_PER_SOURCE")
        assert np.array_equal(d, ref_engine)
        for i in (0, 1, nsrc // 2, nsrc - 1):
            ref = oracle.dijkstra(h.row_ptr, h.col, h.w, int(srcs[i]))
            assert np.array_equal(d[i], ref), (delta, i)
        assert st["reached"] == int((d != 0xFFFFFFFF).sum())


@pytest.mark.parametrize("chunk", ["32", "64"])
@pytest.mark.parametrize("grid", ["4", "16"])
def test_sssp_batched_shapes_exact(gpu, grid, chunk, monkeypatch):
    """Every launch shape the autotuner may pick for the batched engine (form,
    step blocks per CU, edges per queue entry) gives the same distances; the
    autotuned call (first call on a new handle and delta, then its cached
    shape) too."""
    gb = gpu
    h = gb.gen_rmat(12, 8, seed=61, weights="uniform", wlo=1, whi=500)
    srcs = gb.pick_sources(h, 40, 62)
    ref = None
    monkeypatch.setenv("GB_MSSSP_TUNE", "0")
    monkeypatch.setenv("GB_MSSSP_GRID", grid)
    monkeypatch.setenv("GB_MSSSP_CHUNK", chunk)
    for scan in ("0", "1"):
        monkeypatch.setenv("GB_MSSSP_SCAN", scan)
        _say(f"sssp shape grid={grid} chunk={chunk} scan={scan}")
        g = gb.Graph.from_host(h, need_reverse=False)
        d = gb.u32(gb.sssp(g, srcs, delta=200))
        if ref is None:
            ref = np.stack([oracle.dijkstra(h.row_ptr, h.col, h.w, int(x)) for x in srcs])
        assert np.array_equal(d, ref)
    for k in ("GB_MSSSP_TUNE", "GB_MSSSP_GRID", "GB_MSSSP_CHUNK", "GB_MSSSP_SCAN"):
        monkeypatch.delenv(k)
    g = gb.Graph.from_host(h, need_reverse=False)
    for _ in range(2):  # tuned, then cached
        d, st = gb.sssp(g, srcs, delta=200, stats=True)
        assert np.array_equal(gb.u32(d), ref)
        assert st["reached"] == int((ref != 0xFFFFFFFF).sum())


@pytest.mark.parametrize("alpha", ["4", "64"])
@pytest.mark.parametrize("grid", ["4", "16"])
def test_bc_shapes_exact(gpu, grid, alpha, monkeypatch):
    """Brandes in every launch shape the autotuner may pick (grid blocks per CU,
    push/pull ratio alpha) matches the oracle; so does the autotuned call."""
    gb = gpu
    h = gb.gen_rmat(12, 8, seed=63, weights="none", symmetric=True)
    srcs = gb.pick_sources(h, 70, 64)
    oracle.set_threads(min(16, oracle.get_threads()))
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    monkeypatch.setenv("GB_BC_TUNE", "0")
    monkeypatch.setenv("GB_BC_GRID", grid)
    monkeypatch.setenv("GB_BC_ALPHA", alpha)
    _say(f"bc shape grid={grid} alpha={alpha}")
    g = gb.Graph.from_host(h, need_reverse=False)
    bc = gb.bc(g, srcs).cpu().numpy()
    assert float(np.max(np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-300))) <= 1e-9
    lv = gb.bfs(g, srcs[:3]).cpu().numpy()
    for i in range(3):
        assert np.array_equal(lv[i], oracle.bfs_levels(h.row_ptr, h.col, int(srcs[i])))
    for k in ("GB_BC_TUNE", "GB_BC_GRID", "GB_BC_ALPHA"):
        monkeypatch.delenv(k)
    g = gb.Graph.from_host(h, need_reverse=False)
    for _ in range(2):  # tuned on the first batch, then cached
        bc = gb.bc(g, srcs).cpu().numpy()
        assert float(np.max(np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-300))) <= 1e-9


@pytest.mark.parametrize("nsrc", [1, 6])
def test_sssp_relaxation_outcome_counters(gpu, nsrc, monkeypatch):
    """Relaxation outcomes (the reference's REV_STATS totals): every reached
    vertex other than a source was improved at least once, improvements are a
    subset of relaxations, and small buckets send improvements to later
    buckets while one bucket covering every distance sends none there."""
    gb = gpu
    h = gb.gen_rmat(12, 8, seed=77, weights="uniform", wlo=1, whi=200)
    g = gb.Graph.from_host(h, need_reverse=False)
    srcs = gb.pick_sources(h, nsrc, 78)
    for per_source in (False, True):
        if per_source:
            monkeypatch.setenv("GB_SSSP_PER_SOURCE", "1")
        for delta, expect_far in ((8, True), (1 << 30, False)):
            _, st = gb.sssp(g, srcs, delta=delta, stats=True)
            imp = st["improved_near"] + st["improved_far"]
            assert st["reached"] - nsrc <= imp <= st["relax_edges"], st
            assert (st["improved_far"] > 0) == expect_far, st
        monkeypatch.delenv("GB_SSSP_PER_SOURCE", raising=False)


@pytest.mark.parametrize("scan", ["0", "1"])
@pytest.mark.parametrize("nsrc", [1, 6, 40])
def test_sssp_counters_exact_on_trees(gpu, nsrc, scan, monkeypatch):
    """The relaxation outcome totals (REV_STATS, csr_sspath_array.c:738-759)
    pinned exactly on out-trees, where every reached vertex has one in-edge:
    each (source, edge) out of a reached vertex is relaxed exactly once, each
    reached non-source vertex is improved exactly once, and the improvement
    stays in the expanding vertex's bucket (near) iff (d[u] + w) // delta ==
    d[u] // delta.  Both engines (the batched one in both forms), several
    deltas and light thresholds."""
    gb = gpu
    monkeypatch.setenv("GB_MSSSP_SCAN", scan)
    rng = np.random.default_rng(900 + nsrc)
    V = 3000
    parent = np.array([rng.integers(0, i) for i in range(1, V)])
    child = np.arange(1, V)
    w = rng.integers(1, 301, V - 1).astype(np.uint32)
    h = gb.HostCSR.from_edges(V, parent, child, w)
    g = gb.Graph.from_host(h, need_reverse=False)
    srcs = np.concatenate([[0], rng.choice(np.arange(1, V), nsrc - 1, replace=False)]).astype(np.int32)
    src_of_edge = np.repeat(np.arange(V), np.diff(h.row_ptr))
    _say(f"sssp counters on a tree, nsrc={nsrc}")
    for per_source in (False, True):
        if per_source:
            monkeypatch.setenv("GB_SSSP_PER_SOURCE", "1")
        for delta, light in ((1, 1), (50, 50), (200, 60), (1 << 20, 1 << 20)):
            d, st = gb.sssp(g, srcs, delta=delta, light_max=light, stats=True)
            d = gb.u32(d).astype(np.int64)
            relax = near = far = reached = 0
            for i in range(nsrc):
                du = d[i][src_of_edge]
                ok = du != 0xFFFFFFFF
                nd = du[ok] + h.w[ok]
                relax += int(ok.sum())
                n_near = int((nd // delta == du[ok] // delta).sum())
                near += n_near
                far += int(ok.sum()) - n_near
                reached += int((d[i] != 0xFFFFFFFF).sum())
            tag = (per_source, delta, light)
            assert st["reached"] == reached, tag
            assert st["traversed"] == relax, tag
            assert st["relax_edges"] == relax, (tag, st)
            assert st["improved_near"] + st["improved_far"] == reached - nsrc, (tag, st)
            assert (st["improved_near"], st["improved_far"]) == (near, far), (tag, st)
        monkeypatch.delenv("GB_SSSP_PER_SOURCE", raising=False)


def test_sssp_batched_overflow_reported(gpu):
    gb = gpu
    h = gb.HostCSR.from_edges(4, [0, 1, 2], [1, 2, 3], [0xFFFFFFF0, 0xFFFFFFF0, 5])
    g = gb.Graph.from_host(h, need_reverse=False)
    with pytest.raises(gb.GBError) as ei:
        gb.sssp(g, [0, 1], delta=1 << 30)
    assert ei.value.status == 4


def test_sssp_config1_rmat16(gpu):
    """R-MAT scale 16, edge factor 16, w in [1,255], seed 42, 8 sources, plain delta."""
    gb = gpu
    h = gb.gen_rmat(16, 16, seed=42, weights="uniform", wlo=1, whi=255)
    srcs = gb.pick_sources(h, 8, 42)
    _say(f"sssp config1 V={h.V} E={h.E}")
    g = gb.Graph.from_host(h, need_reverse=False)
    d, st = gb.sssp(g, srcs, stats=True)
    d = gb.u32(d)
    for i, s in enumerate(srcs):
        ref = oracle.dijkstra(h.row_ptr, h.col, h.w, int(s))
        assert np.array_equal(d[i], ref), f"source {s}"
    deg = h.outdeg().astype(np.int64)
    reached = d != oracle.UINF
    assert st["reached"] == int(reached.sum())
    assert st["traversed"] == int(sum(deg[r].sum() for r in reached))


def test_sssp_edge_cases(gpu):
    gb = gpu
    _say("sssp edge cases")
    # isolated source, unreachable vertices, zero-weight edges, parallel edges
    V = 10
    src = [0, 0, 1, 1, 2, 5, 5, 6]
    dst = [1, 1, 2, 3, 3, 6, 7, 5]
    w = [5, 3, 0, 7, 0, 1, 2, 0]
    h = gb.HostCSR.from_edges(V, src, dst, w)
    g = gb.Graph.from_host(h, need_reverse=False)
    for s in range(V):
        d = gb.u32(gb.sssp(g, [s], delta=2))[0]
        assert np.array_equal(d, oracle.dijkstra(h.row_ptr, h.col, h.w, s)), s
    # single vertex, no edges
    h1 = gb.HostCSR(np.zeros(2, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint32))
    g1 = gb.Graph.from_host(h1, need_reverse=False)
    d1 = gb.u32(gb.sssp(g1, [0], delta=1))
    assert d1.tolist() == [[0]]


def test_sssp_hub_chunks(gpu):
    """A star with 5000 out-edges (more than one 1024-edge chunk) plus a tail."""
    gb = gpu
    _say("sssp hub chunks")
    rng = np.random.default_rng(9)
    V = 6000
    src = np.concatenate([np.zeros(5000, np.int64), rng.integers(1, V, 20000)])
    dst = np.concatenate([np.arange(1, 5001), rng.integers(0, V, 20000)])
    m = src != dst
    w = rng.integers(1, 100, m.sum()).astype(np.uint32)
    h = gb.HostCSR.from_edges(V, src[m], dst[m], w)
    g = gb.Graph.from_host(h, need_reverse=False)
    for delta, lm in [(10, 10), (50, 5), (1000, 1000)]:
        d = gb.u32(gb.sssp(g, [0, 17], delta=delta, light_max=lm))
        assert np.array_equal(d[0], oracle.dijkstra(h.row_ptr, h.col, h.w, 0))
        assert np.array_equal(d[1], oracle.dijkstra(h.row_ptr, h.col, h.w, 17))


def test_sssp_f64_matches_dijkstra(gpu):
    import torch
    gb = gpu
    _say("sssp f64")
    h = small_graph(11, V=300, E=2000)
    rng = np.random.default_rng(3)
    w64 = rng.random(h.E) * 0.9 + 0.1
    g = gb.Graph.from_host(h, need_reverse=False, w64=w64)
    for s in [0, 5, 77]:
        d = gb.sssp_f64(g, [s], delta=0.25).cpu().numpy()[0]
        ref = oracle.dijkstra(h.row_ptr, h.col, w64, s)
        assert np.array_equal(d, ref), s


def test_sssp_multi_graphs(gpu):
    gb = gpu
    _say("sssp multi graphs")
    hs = [gb.gen_rmat(10, 8, seed=100 + i, weights="uniform", wlo=1, whi=1023) for i in range(3)]
    hs += [gb.gen_er(2000 + 500 * i, 8 + i, seed=200 + i, weights="uniform", wlo=1, whi=1023)
           for i in range(3)]
    gs = [gb.Graph.from_host(h, need_reverse=False) for h in hs]
    srcs = [int(gb.pick_sources(h, 1, 5 + i)[0]) for i, h in enumerate(hs)]
    outs = gb.sssp_multi(gs, srcs, deltas=[0, 64, 200, 0, 16, 1000])
    for h, s, o in zip(hs, srcs, outs):
        assert np.array_equal(gb.u32(o), oracle.dijkstra(h.row_ptr, h.col, h.w, s))


def test_sssp_multi_repeated_handles(gpu):
    """One handle twice: the same light threshold shares the handle's cached
    split; two thresholds that both need a split are refused (a handle holds
    one split), and the handle still solves afterwards."""
    gb = gpu
    _say("sssp multi repeated handles")
    h = gb.gen_rmat(11, 8, seed=321, weights="uniform", wlo=1, whi=1023)
    g = gb.Graph.from_host(h, need_reverse=False)
    s = gb.pick_sources(h, 2, 9)
    srcs = [int(s[0]), int(s[1])]
    outs = gb.sssp_multi([g, g], srcs, deltas=[64, 64], light_max=[64, 64])
    for x, o in zip(srcs, outs):
        assert np.array_equal(gb.u32(o), oracle.dijkstra(h.row_ptr, h.col, h.w, x))
    with pytest.raises(gb.GBError) as ei:
        gb.sssp_multi([g, g], srcs, deltas=[64, 128], light_max=[64, 128])
    assert ei.value.status == 1 and "light threshold" in str(ei.value)
    outs = gb.sssp_multi([g], srcs[:1], deltas=[128], light_max=[128])
    assert np.array_equal(gb.u32(outs[0]), oracle.dijkstra(h.row_ptr, h.col, h.w, srcs[0]))


def test_sssp_overflow_reported(gpu):
    gb = gpu
    _say("sssp overflow")
    V = 4
    h = gb.HostCSR.from_edges(V, [0, 1, 2], [1, 2, 3], [0xFFFFFFF0, 0xFFFFFFF0, 5])
    g = gb.Graph.from_host(h, need_reverse=False)
    with pytest.raises(gb.GBError) as ei:
        gb.sssp(g, [0], delta=1 << 30)
    assert ei.value.status == 4


# --------------------------------------------------------------------- BFS / BC

def test_bfs_levels_bit_exact(gpu):
    gb = gpu
    h = gb.gen_rmat(14, 8, seed=3, weights="none", symmetric=True)
    srcs = gb.pick_sources(h, 5, 1)
    _say(f"bfs V={h.V} E={h.E}")
    g = gb.Graph.from_host(h, need_reverse=False)
    lv = gb.bfs(g, srcs).cpu().numpy()
    for i, s in enumerate(srcs):
        assert np.array_equal(lv[i], oracle.bfs_levels(h.row_ptr, h.col, int(s)))


@pytest.mark.parametrize("batch", [0, 3])
def test_bc_rmat_undirected(gpu, batch):
    gb = gpu
    h = gb.gen_rmat(12, 8, seed=11, weights="none", symmetric=True)
    srcs = gb.pick_sources(h, 16, 11)
    _say(f"bc V={h.V} E={h.E} batch={batch}")
    g = gb.Graph.from_host(h, need_reverse=False)
    bc = gb.bc(g, srcs, batch=batch).cpu().numpy()
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    assert np.allclose(bc, ref, rtol=BC_RTOL, atol=1e-12)
    # the backward pass sums (1 + delta[w]) / sigma[w] over successors and
    # multiplies by sigma[v] once, where the oracle multiplies and divides per
    # successor as csr_bc.c does: a few ulps apart
    np.testing.assert_allclose(bc, ref, rtol=1e-12, atol=1e-12)


def test_bc_and_bfs_hub_split(gpu, monkeypatch):
    """Force the hub path (every vertex above 64 edges is split across waves)."""
    gb = gpu
    monkeypatch.setenv("GB_BC_HUB", "64")
    h = gb.gen_rmat(12, 16, seed=21, weights="none", symmetric=True)
    assert h.outdeg().max() > 64 * 4
    srcs = gb.pick_sources(h, 70, 2)
    _say("bc hub split")
    g = gb.Graph.from_host(h, need_reverse=False)
    bc = gb.bc(g, srcs).cpu().numpy()
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    assert np.allclose(bc, ref, rtol=BC_RTOL, atol=1e-12)
    lv = gb.bfs(g, srcs[:5]).cpu().numpy()
    for i, s in enumerate(srcs[:5]):
        assert np.array_equal(lv[i], oracle.bfs_levels(h.row_ptr, h.col, int(s)))
    # directed, distinct in/out hubs
    hd = gb.gen_rmat(11, 16, seed=22, weights="none")
    r = hd.transpose()
    gd = gb.Graph.from_host(hd, reverse=r)
    s2 = gb.pick_sources(hd, 40, 3)
    bcd = gb.bc(gd, s2).cpu().numpy()
    refd = oracle.brandes_bfs(hd.row_ptr, hd.col, s2, r.row_ptr, r.col)
    assert np.allclose(bcd, refd, rtol=BC_RTOL, atol=1e-12)


def test_bc_many_sources_and_duplicates(gpu):
    """More than 64 sources (several lane groups), a repeated source, batch split."""
    gb = gpu
    h = gb.gen_rmat(11, 8, seed=12, weights="none", symmetric=True)
    srcs = list(gb.pick_sources(h, 150, 3)) + [int(gb.pick_sources(h, 1, 3)[0])]
    _say(f"bc many sources n={len(srcs)}")
    g = gb.Graph.from_host(h, need_reverse=False)
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs)
    first = None
    for batch in (0, 64, 100):
        bc = gb.bc(g, srcs, batch=batch).cpu().numpy()
        assert np.allclose(bc, ref, rtol=BC_RTOL, atol=1e-12), batch
        np.testing.assert_allclose(bc, ref, rtol=1e-12, atol=1e-12)
        # lanes sum in CSR order and sources accumulate in order: the batch
        # split changes no bit
        if first is None:
            first = bc
        assert np.array_equal(bc, first), batch


def test_bfs_many_sources(gpu):
    gb = gpu
    h = gb.gen_rmat(12, 8, seed=4, weights="none")
    srcs = list(gb.pick_sources(h, 130, 9)) + [0, 0]
    _say("bfs many sources")
    g = gb.Graph.from_host(h)
    lv = gb.bfs(g, srcs).cpu().numpy()
    for i, s in enumerate(srcs):
        assert np.array_equal(lv[i], oracle.bfs_levels(h.row_ptr, h.col, int(s))), i


def test_bc_directed(gpu):
    gb = gpu
    h = small_graph(21, V=300, E=1500)
    r = h.transpose()
    srcs = list(range(0, 300, 7))
    _say("bc directed")
    g = gb.Graph.from_host(h, reverse=r)
    bc = gb.bc(g, srcs).cpu().numpy()
    ref = oracle.brandes_bfs(h.row_ptr, h.col, srcs, r.row_ptr, r.col)
    assert np.allclose(bc, ref, rtol=BC_RTOL, atol=1e-12)
    np.testing.assert_allclose(bc, ref, rtol=1e-12, atol=1e-12)


def test_bc_weighted(gpu):
    gb = gpu
    h = small_graph(31, V=250, E=1500, wlo=1, whi=20)
    r = h.transpose()
    _say("bc weighted")
    w64 = h.w.astype(np.float64) / 8.0
    rw64 = r.w.astype(np.float64) / 8.0
    g = gb.Graph.from_host(h, reverse=r, w64=w64, rw64=rw64)
    srcs = list(range(0, 250, 5))
    bc = gb.bc(g, srcs, weighted=True).cpu().numpy()
    ref = oracle.brandes_weighted(h.row_ptr, h.col, w64, r.row_ptr, r.col, rw64, srcs)
    assert np.allclose(bc, ref, rtol=BC_RTOL, atol=1e-12)


def test_bc_weighted_zero_weights(gpu):
    """A zero-weight edge off any cycle is an ordinary tight edge; a cycle of
    zero-weight edges has no Brandes order and is refused, not divided by a
    zero path count."""
    gb = gpu
    _say("bc weighted zero weights")

    def graph(src, dst, w, V):
        h = gb.HostCSR.from_edges(V, src, dst, np.ones(len(src), np.uint32))
        order = np.lexsort((np.asarray(dst), np.asarray(src)))
        w64 = np.asarray(w, np.float64)[order]
        r = h.transpose()
        rorder = np.lexsort((np.asarray(src)[order], np.asarray(dst)[order]))
        rw64 = w64[rorder]
        return h, r, w64, rw64

    # 0 -> 1 (0.0), 1 -> 2 (1.0), 0 -> 3 (0.5), 3 -> 2 (0.5): a DAG
    h, r, w64, rw64 = graph([0, 1, 0, 3], [1, 2, 3, 2], [0.0, 1.0, 0.5, 0.5], 4)
    g = gb.Graph.from_host(h, reverse=r, w64=w64, rw64=rw64)
    bc = gb.bc(g, [0, 1, 2, 3], weighted=True).cpu().numpy()
    ref = oracle.brandes_weighted(h.row_ptr, h.col, w64, r.row_ptr, r.col, rw64, [0, 1, 2, 3])
    assert np.allclose(bc, ref, rtol=BC_RTOL, atol=1e-12)
    # 0 -> 1 (1.0), 1 -> 2 (0.0), 2 -> 1 (0.0): 1 and 2 lie on a tight cycle
    h, r, w64, rw64 = graph([0, 1, 2], [1, 2, 1], [1.0, 0.0, 0.0], 3)
    g = gb.Graph.from_host(h, reverse=r, w64=w64, rw64=rw64)
    with pytest.raises(gb.GBError, match="cycle of tight edges"):
        gb.bc(g, [0], weighted=True)


# --------------------------------------------------------------------- PageRank

def _pr_case(gb, h, **kw):
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    x, res = gb.pagerank(g, **kw)
    x = x.cpu().numpy().astype(np.float64)
    ref, it, rr = oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), kw.get("damping", 0.85),
                                      kw.get("tol", 1e-7), kw.get("max_iter", 100))
    return x, res, ref, it


@pytest.mark.parametrize("ipt", [0, 4, 8, 12, 16])
def test_pagerank_rmat(gpu, ipt):
    gb = gpu
    h = gb.gen_rmat(15, 16, seed=7, weights="none")
    _say(f"pagerank rmat15 ipt={ipt}")
    x, res, ref, it = _pr_case(gb, h, items_per_thread=ipt)
    l1 = np.abs(x - ref).sum()
    assert l1 <= PR_L1, l1
    assert abs(res["iterations"] - it) <= 1
    assert abs(x.sum() - 1.0) < 1e-5


@pytest.mark.parametrize("relabel", [-1, 1])
def test_pagerank_relabel_matches_oracle(gpu, relabel):
    """The out-degree relabelled gather gives the oracle's ranks in the
    caller's ids, with and without a warm start, on skewed and small graphs."""
    gb = gpu
    for h in (gb.gen_rmat(14, 16, seed=21, weights="none"),
              gb.gen_er(3000, 5, seed=22, weights="none"),
              gb.HostCSR.from_edges(6, [0, 1, 2, 3], [1, 2, 0, 0])):
        _say(f"pagerank relabel={relabel} V={h.V} E={h.E}")
        x, res, ref, it = _pr_case(gb, h, relabel=relabel)
        assert np.abs(x - ref).sum() <= PR_L1
        r = h.transpose()
        g = gb.Graph.from_host(h, reverse=r)
        x0 = gb.to_device(np.asarray(ref, np.float32))
        xw, rw = gb.pagerank(g, x0=x0, relabel=relabel)
        assert rw["iterations"] <= 2
        assert np.abs(xw.cpu().numpy().astype(np.float64) - ref).sum() <= PR_L1


def test_pagerank_bad_items_per_thread_rejected(gpu):
    gb = gpu
    h = gb.gen_rmat(8, 8, seed=3, weights="none")
    g = gb.Graph.from_host(h, reverse=h.transpose())
    with pytest.raises(gb.GBError):
        gb.pagerank(g, items_per_thread=5)


def test_pagerank_hubs_span_tiles(gpu):
    """In-degree hubs far larger than a tile exercise the cross-tile fix-up."""
    gb = gpu
    rng = np.random.default_rng(5)
    V = 3000
    src = np.concatenate([rng.integers(0, V, 40000), rng.integers(0, V, 5000)])
    dst = np.concatenate([np.full(40000, 17), rng.integers(0, V, 5000)])
    m = src != dst
    h = gb.HostCSR.from_edges(V, src[m], dst[m])
    _say("pagerank hubs")
    for ipt in (4, 8):
        x, res, ref, it = _pr_case(gb, h, items_per_thread=ipt)
        assert np.abs(x - ref).sum() <= PR_L1


@pytest.mark.parametrize("hot", [0, 4096, 8192])
@pytest.mark.parametrize("rb", [1024, 2048, 4096])
def test_pagerank_row_blocks_match_oracle(gpu, rb, hot, monkeypatch):
    """The row-block form: rows summed in LDS blocks, rows longer than a block
    split into chunks joined by the fix-up.  Cases: skewed R-MAT (relabelled
    and not), an in-degree hub of 40000 edges, runs of rows without in-edges,
    an edgeless graph, a cycle (closed form: uniform), and graph replay.  With
    a hot set (relabelled graphs), several row blocks share one launch block
    and the persistent grid walks them in rounds: R-MAT 18 takes several
    rounds and ends with launch blocks holding fewer row blocks.  The fix-up
    runs in one block when partials and long rows are few, else in many:
    R-MAT 18 at NB 1024 without a hot set, and 4200 long rows at NB 1024,
    take the multi-block form."""
    gb = gpu
    monkeypatch.setenv("GB_PR_HOT", str(hot))
    rng = np.random.default_rng(rb)
    V = 3000
    src = np.concatenate([rng.integers(0, V, 40000), rng.integers(0, V, 5000)])
    dst = np.concatenate([np.full(40000, 17), rng.integers(0, V // 3, 5000)])
    m = src != dst
    # the hub graph converges slowly: its FP32 residual floor sits near the
    # tolerance, so only its ranks are compared (as for the merge path)
    graphs = [(gb.gen_rmat(14, 16, seed=21, weights="none"), 1, True),
              (gb.gen_rmat(13, 8, seed=23, weights="none"), -1, True),
              (gb.HostCSR.from_edges(V, src[m], dst[m]), -1, False)]
    if hot or rb == 1024:
        # (rb 1024, no hot set: more than 16384 partials, the multi-block fix-up)
        graphs.append((gb.gen_rmat(18, 16, seed=5, weights="none"), 1, True))
    if rb == 1024:
        # 4200 long rows (more than the one-block fix-up takes)
        n = 4200
        hs = rng.integers(0, 3 * n, n * 1030)
        hd = np.repeat(np.arange(n), 1030)
        graphs.append((gb.HostCSR.from_edges(3 * n, hs[hs != hd], hd[hs != hd]), -1, True))
    for h, rl, check_it in graphs:
        _say(f"pagerank row blocks rb={rb} hot={hot} V={h.V} E={h.E} relabel={rl}")
        x, res, ref, it = _pr_case(gb, h, row_block=rb, relabel=rl)
        assert res["row_block"] == rb and res["items_per_thread"] == 0
        assert res["hot"] == (hot if rl == 1 else 0)
        assert np.abs(x - ref).sum() <= PR_L1
        assert not check_it or abs(res["iterations"] - it) <= 1
    h = gb.HostCSR(np.zeros(6, np.int32), np.zeros(0, np.int32))
    x, res, ref, it = _pr_case(gb, h, row_block=rb)
    assert np.allclose(x, 0.2, atol=1e-7)
    Vc = 50
    h = gb.HostCSR.from_edges(Vc, np.arange(Vc), (np.arange(Vc) + 1) % Vc)
    x, res, ref, it = _pr_case(gb, h, row_block=rb)
    assert np.abs(x - 1.0 / Vc).sum() < 1e-6 and res["iterations"] == 1
    h = gb.gen_rmat(12, 8, seed=9, weights="none")
    g = gb.Graph.from_host(h, reverse=h.transpose())
    xs = [gb.pagerank(g, use_graph=ug, row_block=rb)[0].cpu().numpy() for ug in (True, False)]
    assert np.array_equal(xs[0], xs[1])


def test_pagerank_row_block_args_rejected(gpu):
    gb = gpu
    h = gb.gen_rmat(8, 8, seed=3, weights="none")
    g = gb.Graph.from_host(h, reverse=h.transpose())
    for kw in (dict(row_block=3000), dict(row_block=-2), dict(row_block=2048, items_per_thread=8)):
        with pytest.raises(gb.GBError):
            gb.pagerank(g, **kw)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6])
def test_pagerank_binned_matches_oracle(gpu, shape, monkeypatch):
    """The binned form: contributions streamed into (bin, chunk) segments and
    summed per bin in fixed point.  Cases: skewed R-MAT (relabelled and not),
    an in-degree hub of 40000 edges, an ER graph whose V is no multiple of the
    chunk or bin width, runs of rows without in-edges, an edgeless graph and a
    cycle (closed form: uniform).  Small pieces (4096 entries) split the hub's
    bin and the R-MAT hub bins over several blocks, whose sums meet in the
    fix-up; the default pieces keep most bins in one block."""
    gb = gpu
    rng = np.random.default_rng(40 + shape)
    V = 3000
    src = np.concatenate([rng.integers(0, V, 40000), rng.integers(0, V, 5000)])
    dst = np.concatenate([np.full(40000, 17), rng.integers(0, V // 3, 5000)])
    m = src != dst
    graphs = [(gb.gen_rmat(14, 16, seed=21, weights="none"), 1, True),
              (gb.gen_rmat(15, 8, seed=23, weights="none"), -1, True),
              (gb.gen_er(20011, 6, seed=24, weights="none"), -1, True),
              (gb.HostCSR.from_edges(V, src[m], dst[m]), -1, False)]
    for piece in ("4096", None):
        if piece:
            monkeypatch.setenv("GB_PR_PB_PIECE", piece)
        else:
            monkeypatch.delenv("GB_PR_PB_PIECE", raising=False)
        for h, rl, check_it in graphs:
            _say(f"pagerank binned shape={shape} piece={piece} V={h.V} E={h.E} relabel={rl}")
            x, res, ref, it = _pr_case(gb, h, binned=shape, relabel=rl)
            assert res["binned"] == shape and res["row_block"] == 0
            assert res["items_per_thread"] == 0 and res["hot"] == 0
            assert np.abs(x - ref).sum() <= PR_L1
            assert not check_it or abs(res["iterations"] - it) <= 1
    h = gb.HostCSR(np.zeros(6, np.int32), np.zeros(0, np.int32))
    x, res, ref, it = _pr_case(gb, h, binned=shape)
    assert np.allclose(x, 0.2, atol=1e-7)
    Vc = 50
    h = gb.HostCSR.from_edges(Vc, np.arange(Vc), (np.arange(Vc) + 1) % Vc)
    x, res, ref, it = _pr_case(gb, h, binned=shape)
    assert np.abs(x - 1.0 / Vc).sum() < 1e-6 and res["iterations"] == 1


def test_pagerank_binned_identical_runs_and_scaled_start(gpu, monkeypatch):
    """Fixed-point sums make the binned form's ranks independent of the order
    in which blocks add: graph replay, eager launches and repeated solves give
    identical bits, also with split bins.  A start of mass 1000 (fixed-point
    shift 40 instead of 50) converges to the same ranks."""
    gb = gpu
    monkeypatch.setenv("GB_PR_PB_PIECE", "4096")
    h = gb.gen_rmat(14, 16, seed=9, weights="none")
    g = gb.Graph.from_host(h, reverse=h.transpose())
    _say("pagerank binned identical runs")
    xs = [gb.pagerank(g, use_graph=ug, binned=1)[0].cpu().numpy() for ug in (True, False, True)]
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[1], xs[2])
    ref, it, _ = oracle.pagerank_cfg(h.transpose().row_ptr, h.transpose().col, h.outdeg(),
                                     0.85, 1e-10, 300)
    x0 = gb.to_device(np.full(h.V, 1000.0 / h.V, np.float32))
    # from mass 1000 the error falls by 0.85 an iteration: about 150 to 1e-7
    x, res = gb.pagerank(g, x0=x0, binned=1, tol=1e-8, max_iter=300)
    assert res["binned"] == 1
    assert np.abs(x.cpu().numpy().astype(np.float64) - ref).sum() <= PR_L1


def test_pagerank_binned_args_rejected(gpu):
    gb = gpu
    h = gb.gen_rmat(8, 8, seed=3, weights="none")
    g = gb.Graph.from_host(h, reverse=h.transpose())
    for kw in (dict(binned=7), dict(binned=-2), dict(binned=1, row_block=2048),
               dict(binned=1, items_per_thread=8)):
        with pytest.raises(gb.GBError):
            gb.pagerank(g, **kw)


def test_pagerank_graph_cache_across_handles_and_forms(gpu):
    """The instantiated hipGraph is kept between calls and replayed only for
    the same handle, form and buffers: alternating handles and forms, a handle
    destroyed and another created, and a start vector each give the eager
    result bit for bit."""
    gb = gpu
    hs = [gb.gen_rmat(12, 8, seed=31, weights="none"), gb.gen_er(5000, 6, seed=32, weights="none")]
    gs = [gb.Graph.from_host(h, reverse=h.transpose()) for h in hs]
    _say("pagerank graph cache")
    forms = [dict(binned=1), dict(row_block=2048), dict(items_per_thread=8), dict(binned=4)]
    eager = {}
    for i, g in enumerate(gs):
        for j, f in enumerate(forms):
            eager[i, j] = gb.pagerank(g, use_graph=False, **f)[0].cpu().numpy()
    for _ in range(2):
        for j, f in enumerate(forms):
            for i, g in enumerate(gs):
                x = gb.pagerank(g, use_graph=True, **f)[0].cpu().numpy()
                assert np.array_equal(x, eager[i, j]), (i, f)
    del gs[0]
    import gc
    gc.collect()
    g3 = gb.Graph.from_host(hs[0], reverse=hs[0].transpose())
    for j, f in enumerate(forms):
        assert np.array_equal(gb.pagerank(g3, use_graph=True, **f)[0].cpu().numpy(), eager[0, j])
    x0 = gb.to_device(np.asarray(eager[0, 0], np.float32))
    xa = gb.pagerank(g3, use_graph=True, x0=x0, binned=1)[0].cpu().numpy()
    xb = gb.pagerank(g3, use_graph=False, x0=x0, binned=1)[0].cpu().numpy()
    assert np.array_equal(xa, xb)


def test_pagerank_graph_vs_eager(gpu):
    gb = gpu
    h = gb.gen_rmat(12, 8, seed=9, weights="none")
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    _say("pagerank graph vs eager")
    xs = []
    for ug in (True, False, True):
        x, res = gb.pagerank(g, use_graph=ug)
        xs.append(x.cpu().numpy())
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[1], xs[2])


def test_pagerank_edge_cases(gpu):
    gb = gpu
    _say("pagerank edge cases")
    # all vertices dangling (no edges): stays uniform
    h = gb.HostCSR(np.zeros(6, np.int32), np.zeros(0, np.int32))
    x, res, ref, it = _pr_case(gb, h)
    assert np.allclose(x, 0.2, atol=1e-7)
    # directed cycle: uniform fixed point
    V = 50
    h = gb.HostCSR.from_edges(V, np.arange(V), (np.arange(V) + 1) % V)
    x, res, ref, it = _pr_case(gb, h)
    assert np.abs(x - 1.0 / V).sum() < 1e-6 and res["iterations"] == 1


def test_pagerank_warm_start(gpu):
    import torch
    gb = gpu
    h = gb.gen_rmat(13, 16, seed=4, weights="none")
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    _say("pagerank warm start")
    x, res = gb.pagerank(g, tol=1e-7)
    x2, res2 = gb.pagerank(g, tol=1e-7, x0=x)
    assert res2["iterations"] <= 2
    ref, it, _ = oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), 0.85, 1e-7, 100)
    assert np.abs(x2.cpu().numpy() - ref).sum() <= PR_L1


def test_pagerank_ref_bit_exact(gpu):
    import torch
    gb = gpu
    h = gb.gen_rmat(11, 8, seed=2, weights="none")
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    od = h.outdeg().astype(np.float64)
    with np.errstate(divide="ignore"):
        recip = 1.0 / od
    _say("pagerank ref")
    x, it = gb.pagerank_ref(g, torch.from_numpy(recip).cuda())
    xr, itr = oracle.pagerank_ref(r.row_ptr, r.col, recip)
    assert it == itr
    assert np.array_equal(x.cpu().numpy(), xr)


# --------------------------------------------------------------------- split

def test_merge_path_split(gpu):
    import torch
    gb = gpu
    h = gb.gen_rmat(12, 16, seed=1, weights="none")
    rp = torch.from_numpy(h.row_ptr).cuda()
    _say("merge path split")
    for items in (7, 256, 2048):
        nt = (h.V + h.E + items - 1) // items
        tr, te = gb.merge_path_split(rp, items, nt)
        rr, re = oracle.merge_path_split(h.row_ptr, items, nt)
        assert np.array_equal(tr.cpu().numpy(), rr) and np.array_equal(te.cpu().numpy(), re)


def test_edge_balanced_parts(gpu):
    import torch
    gb = gpu
    h = gb.gen_rmat(12, 16, seed=1, weights="none")
    rp = torch.from_numpy(h.row_ptr).cuda()
    _say("edge balanced parts")
    for p in (1, 2, 3, 8, 64):
        d = gb.edge_balanced_parts(rp, p).cpu().numpy()
        ref = oracle.comp_distrib(h.outdeg(), h.E, p)
        assert np.array_equal(d, ref), p


def test_csr_from_edges_and_transpose(gpu):
    """Device CSR construction: rows in source order, a row's edges in input
    order (a stable sort, like the reference's scatter in file order); the
    transpose equals the host transpose; weights ride along; bad ids fail."""
    gb = gpu
    import torch
    rng = np.random.default_rng(3)
    V, E = 5000, 60000
    src = rng.integers(0, V, E).astype(np.int32)
    dst = rng.integers(0, V, E).astype(np.int32)
    w = rng.integers(1, 1000, E).astype(np.uint32)
    w64 = rng.random(E)
    d = gb.to_device
    rp, col, wo = gb.csr_from_edges(V, d(src), d(dst), d(w))
    order = np.argsort(src, kind="stable")
    ref_rp = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=V))]).astype(np.int32)
    assert np.array_equal(rp.cpu().numpy(), ref_rp)
    assert np.array_equal(col.cpu().numpy(), dst[order])
    assert np.array_equal(gb.u32(wo), w[order])
    _, _, wf = gb.csr_from_edges(V, d(src), d(dst), torch.from_numpy(w64).cuda())
    assert np.array_equal(wf.cpu().numpy(), w64[order])
    h = gb.gen_rmat(13, 8, seed=4, weights="uniform", wlo=1, whi=50)
    r = h.transpose()
    trp, tcol, tw = gb.csr_transpose(d(h.row_ptr), d(h.col), d(h.w))
    assert np.array_equal(trp.cpu().numpy(), r.row_ptr)
    assert np.array_equal(tcol.cpu().numpy(), r.col)
    assert np.array_equal(gb.u32(tw), r.w)
    erp, ecol, _ = gb.csr_from_edges(7, d(np.zeros(0, np.int32)), d(np.zeros(0, np.int32)))
    assert np.array_equal(erp.cpu().numpy(), np.zeros(8, np.int32)) and ecol.numel() == 0
    bad = src.copy()
    bad[10] = V
    with pytest.raises(gb.GBError):
        gb.csr_from_edges(V, d(bad), d(dst))


@pytest.mark.parametrize("E,V", [(0, 1), (1, 1), (300000, 700), (300000, 1 << 30)])
def test_edges_unique_vs_oracle(gpu, E, V):
    """Device dedup (remove_duplicate's first-occurrence rule) against the
    oracle, whose own rule is pinned on the reference's output files."""
    gb = gpu
    rng = np.random.default_rng(E + 17)
    src = rng.integers(0, V, E).astype(np.int32)
    dst = rng.integers(0, V, E).astype(np.int32)
    _say(f"edges_unique E={E} V={V}")
    so, do = gb.edges_unique(gb.to_device(src), gb.to_device(dst))
    ref = oracle.remove_duplicate(np.stack([src, dst], axis=1).astype(np.int64))
    assert np.array_equal(so.cpu().numpy(), ref[:, 0]) and np.array_equal(do.cpu().numpy(), ref[:, 1])


def test_edges_unique_rejects_negative_ids(gpu):
    gb = gpu
    src = np.array([0, 1, -1], np.int32)
    dst = np.array([1, 2, 3], np.int32)
    with pytest.raises(gb.GBError):
        gb.edges_unique(gb.to_device(src), gb.to_device(dst))


@pytest.mark.parametrize("E,V,p", [(0, 10, 3), (1, 10, 1), (3_000_000, 1 << 20, 2),
                                   (3_000_000, 1 << 20, 64), (1_000_000, 1 << 16, 91),
                                   (200_000, 1 << 16, 4096)])
def test_edge_parts_vs_oracle(gpu, E, V, p):
    """Owners of both endpoints and the cross-part counts of chunk_files.
    p = 91 and 4096 exceed the per-block LDS counts (global atomics)."""
    gb = gpu
    rng = np.random.default_rng(E + p)
    _say(f"edge_parts E={E} V={V} p={p}")
    cuts = np.sort(rng.choice(np.arange(V - 1), p - 1, replace=False)) if p > 1 else []
    distrib = np.append(np.asarray(cuts, np.int64), V - 1)
    src = rng.integers(0, V, E).astype(np.int32)
    dst = rng.integers(0, V, E).astype(np.int32)
    sp, rp, ub = gb.edge_parts(gb.to_device(src), gb.to_device(dst), distrib)
    osp, orp, oub = oracle.edge_parts(src, dst, distrib)
    assert np.array_equal(sp.cpu().numpy(), osp)
    assert np.array_equal(rp.cpu().numpy(), orp)
    assert np.array_equal(ub.cpu().numpy(), oub)


def test_edge_parts_unordered_distrib_and_bad_p(gpu):
    """A distrib with parts left at 0 (comp_distrib_csr cut fewer parts than
    num_nodes): bin_search's answers as the oracle gives them; p outside
    [1, 4096] is rejected."""
    gb = gpu
    rng = np.random.default_rng(3)
    distrib = np.array([40, 0, 0, 99], np.int64)
    src = rng.integers(0, 100, 5000).astype(np.int32)
    dst = rng.integers(0, 100, 5000).astype(np.int32)
    sp, rp, ub = gb.edge_parts(gb.to_device(src), gb.to_device(dst), distrib)
    osp, orp, oub = oracle.edge_parts(src, dst, distrib)
    assert np.array_equal(sp.cpu().numpy(), osp) and np.array_equal(rp.cpu().numpy(), orp)
    assert np.array_equal(ub.cpu().numpy(), oub)
    for bad in (np.zeros(0, np.int64), np.arange(4097, dtype=np.int64)):
        with pytest.raises(gb.GBError):
            gb.edge_parts(gb.to_device(src), gb.to_device(dst), bad)


def _hub_graph(gb):
    """Empty rows, a 5000-edge hub spanning many edge tiles, short rows."""
    rng = np.random.default_rng(5)
    V = 3000
    src = np.concatenate([np.full(5000, 7), rng.integers(100, V, 4000)])
    dst = rng.integers(0, V, src.size)
    w = rng.integers(1, 300, src.size)
    return gb.HostCSR.from_edges(V, src, dst, w)


@pytest.mark.parametrize("case,light", [("small", 100), ("rmat", 0), ("rmat", 60),
                                        ("rmat", 255), ("hub", 150)])
def test_separate(gpu, case, light):
    gb = gpu
    if case == "small":
        h = small_graph(41)
    elif case == "rmat":
        h = gb.gen_rmat(14, 16, seed=3, weights="uniform", wlo=1, whi=255)
    else:
        h = _hub_graph(gb)
    g = gb.Graph.from_host(h, need_reverse=False)
    _say(f"separate {case} light<={light}")
    col, w, le = gb.separate(g, light)
    (lrp, lcol, lw), (hrp, hcol, hw) = oracle.separate_csr(h.row_ptr, h.col, h.w, light)
    col, w, le = col.cpu().numpy(), w.cpu().numpy().view(np.uint32), le.cpu().numpy()
    for v in range(h.V):
        a, b, c = h.row_ptr[v], le[v], h.row_ptr[v + 1]
        assert np.array_equal(col[a:b], lcol[lrp[v]:lrp[v + 1]])
        assert np.array_equal(w[a:b], lw[lrp[v]:lrp[v + 1]])
        assert np.array_equal(col[b:c], hcol[hrp[v]:hrp[v + 1]])
        assert np.array_equal(w[b:c], hw[hrp[v]:hrp[v + 1]])
