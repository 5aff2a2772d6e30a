"""Pins for the CPU oracle (oracle/oracle.c), no GPU.

Each oracle function is checked against (a) golden vectors produced by the
reference's own code (tests/golden/*.json, made by tests/golden/make_golden.py
from the reference build), and (b) checks the mathematics fixes: scipy/networkx,
closed forms, brute force, invariants.  A plausible slip (wrong sign, index or
operand, a dropped term) fails at least one of them.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)["cases"]


def csr(V, src, dst, w=None):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    o = np.argsort(src, kind="stable")
    rp = np.zeros(V + 1, np.int64)
    np.add.at(rp, src + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    ww = None if w is None else np.asarray(w)[o]
    return rp, dst[o].astype(np.int32), ww


def transpose(V, rp, col, w=None):
    src = np.repeat(np.arange(V), np.diff(rp))
    return csr(V, col, src, w)


# ------------------------------------------------------------------ golden

def first_parallel_weights(rp, col, w):
    """Weight of every edge replaced by that of the first parallel edge (u,v) in
    CSR order: what relax() uses, through find_edge_idx (csr_sspath.c:94-96, 124-134)."""
    out = w.copy()
    for u in range(len(rp) - 1):
        first = {}
        for e in range(rp[u], rp[u + 1]):
            first.setdefault(int(col[e]), w[e])
            out[e] = first[int(col[e])]
    return out


@pytest.mark.parametrize("case", load("sssp_list.json"), ids=lambda c: f"V{c['V']}s{c['source']}")
def test_sssp_vs_reference_run_sspath(case):
    """Distances of the reference's run_sspath (csr_sspath.c:38-83), bit-exact.

    With parallel edges of different weights the reference relaxes with the
    first parallel edge's weight (defect A10); its output is then exactly the
    oracle's on those weights, and the plain oracle gives the true distances."""
    rp, col, w = csr(case["V"], case["src"], case["dst"], np.array(case["w"]))
    ref = np.array(case["dist"])
    s = case["source"]
    wf = first_parallel_weights(rp, col, w)
    assert np.array_equal(oracle.dijkstra(rp, col, wf, s), ref)
    assert np.array_equal(oracle.delta_stepping(rp, col, wf, s, case["delta"]), ref)
    if np.array_equal(wf, w):  # defect cannot fire: the true distances match
        assert np.array_equal(oracle.dijkstra(rp, col, w, s), ref)
    if all(float(x).is_integer() for x in case["w"]):
        d32 = oracle.delta_stepping(rp, col, wf.astype(np.uint32), s, int(case["delta"]))
        assert np.array_equal(np.where(d32 == oracle.UINF, np.inf, d32.astype(float)), ref)


@pytest.mark.parametrize("case", load("sssp_list.json")[:15], ids=lambda c: f"V{c['V']}s{c['source']}")
def test_reference_predecessors_are_tight_edges_with_duplicates(case):
    """run_sspath's predecessor lists hold exactly the tight in-edges' sources,
    with repeats (the duplication defect); the deduplicated sets match."""
    V = case["V"]
    rp, col, w = csr(V, case["src"], case["dst"], np.array(case["w"]))
    w = first_parallel_weights(rp, col, w)
    d = np.array(case["dist"])
    tight = [set() for _ in range(V)]
    for u in range(V):
        for e in range(rp[u], rp[u + 1]):
            v = col[e]
            if np.isfinite(d[u]) and d[u] + w[e] == d[v]:
                tight[v].add(u)
    for v in range(V):
        if v == case["source"]:
            continue
        assert set(case["pred_raw"][v]) == tight[v], v


@pytest.mark.parametrize("case", load("distrib.json"), ids=lambda c: f"V{c['V']}p{c['p']}t{c['type']}")
def test_comp_distrib_and_bin_search_vs_reference(case):
    deg = np.array(case["send_deg"] if case["type"] == 1 else case["recv_deg"])
    E = int(np.sum(case["send_deg"]))
    d = oracle.comp_distrib(deg, E, case["p"])
    assert d.tolist() == case["distrib"]
    parts = [oracle.bin_search(np.array(case["distrib"]), v) for v in range(case["V"])]
    assert parts == case["part"]


@pytest.mark.parametrize("case", load("separate.json"), ids=lambda c: f"V{c['V']}")
def test_separate_vs_reference(case):
    V = case["V"]
    rp, col, w = csr(V, case["src"], case["dst"], np.array(case["w"]))
    (lrp, lcol, lw), (hrp, hcol, hw) = oracle.separate_csr(rp, col, w, case["delta"])
    assert lrp[1:].tolist() == case["light_idx"] and hrp[1:].tolist() == case["heavy_idx"]
    assert lcol.tolist() == case["light_col"] and hcol.tolist() == case["heavy_col"]
    assert lw.tolist() == case["light_w"] and hw.tolist() == case["heavy_w"]


@pytest.mark.parametrize("case", load("remove_duplicate.json"),
                         ids=lambda c: f"{c['name']}-n{c['num_nodes']}")
def test_remove_duplicate_vs_reference(case):
    """The oracle's kept pairs equal the pairs in the reference's output file
    (after its header: g_v_num, count, 0, g_v_num for one node; count otherwise)."""
    out = np.array(case["output"], np.int64)
    hdr = 4 if case["num_nodes"] == 1 else 1
    kept = oracle.remove_duplicate(case["pairs"])
    assert (out[1] if hdr == 4 else out[0]) == len(kept)
    assert np.array_equal(out[hdr:].reshape(-1, 2), kept)
    if hdr == 4:
        assert out[0] == out[3] == case["g_v_num"] and out[2] == 0


def test_remove_duplicate_brute_force():
    """Against a dictionary of first occurrences on random small inputs."""
    rng = np.random.default_rng(9)
    for E in (0, 1, 2, 50, 777):
        p = rng.integers(0, 12, (E, 2))
        seen, ref = set(), []
        for a, b in p.tolist():
            if (a, b) not in seen:
                seen.add((a, b))
                ref.append((a, b))
        assert oracle.remove_duplicate(p).tolist() == [list(x) for x in ref]


@pytest.mark.parametrize("case", load("chunk.json"), ids=lambda c: f"{c['name']}-n{c['num_nodes']}")
def test_chunk_files_vs_reference(case):
    """The oracle's chunking step (degrees, distrib, write_recip, chunk_files)
    writes exactly the files the reference CLI -k wrote."""
    got = oracle.chunk_files(case["text"], "g.txt", case["num_nodes"])
    assert sorted(got) == sorted(case["files"])
    for name, text in case["files"].items():
        assert got[name] == text, name


def test_edge_parts_lower_bound_and_counts():
    """bin_search over an ascending distrib is the first part whose last
    vertex is >= x (lower bound); the cross-part counts are a 2-D histogram
    that skips the diagonal and sums to the number of cut edges."""
    rng = np.random.default_rng(17)
    for V, E, p in ((10, 0, 3), (1000, 5000, 1), (1000, 5000, 2), (5000, 20000, 7),
                    (300, 4000, 64)):
        cuts = np.sort(rng.choice(np.arange(V - 1), p - 1, replace=False)) if p > 1 else []
        distrib = np.append(np.asarray(cuts, np.int64), V - 1)
        src = rng.integers(0, V, E).astype(np.int32)
        dst = rng.integers(0, V, E).astype(np.int32)
        sp, rp, ub = oracle.edge_parts(src, dst, distrib)
        assert np.array_equal(sp, np.searchsorted(distrib, src, side="left"))
        assert np.array_equal(rp, np.searchsorted(distrib, dst, side="left"))
        ref = np.zeros((p, p), np.int64)
        np.add.at(ref, (sp, rp), 1)
        np.fill_diagonal(ref, 0)
        assert np.array_equal(ub, ref)
        assert ub.sum() == int((sp != rp).sum())


def parse_rc(text):
    """The reference's chunked RECV file (p sp V E off gV / a u v w) -> in-CSR
    in file order, as scan_csr_idx + read_graph_csr (RECV) build it."""
    toks = text.split()
    i = toks.index("p")
    V, E, off = int(toks[i + 2]), int(toks[i + 3]), int(toks[i + 4])
    us, vs = [], []
    k = i + 6
    while len(us) < E:
        assert toks[k] == "a"
        us.append(int(toks[k + 1]) - 1)
        vs.append(int(toks[k + 2]) - off)
        k += 4
    rp = np.zeros(V + 1, np.int64)
    np.add.at(rp, np.array(vs) + 1, 1)
    rp = np.cumsum(rp)
    pos = rp[:-1].copy()
    col = np.empty(E, np.int32)
    for u, v in zip(us, vs):
        col[pos[v]] = u
        pos[v] += 1
    return V, rp.astype(np.int32), col


@pytest.mark.parametrize("case", load("pagerank.json"), ids=lambda c: f"V{c['V']}")
def test_pagerank_ref_vs_reference_pipeline(case):
    """Iterations equal; printed values (%lf, 6 decimals) within rounding."""
    V, irp, icol = parse_rc(case["rc_file"])
    recip = np.array([float(t) for t in case["recip_file"].split()[:V]])
    x, it = oracle.pagerank_ref(irp, icol, recip)
    assert it == case["iterations"]
    n = len(case["final"])
    assert np.max(np.abs(x[:n] - np.array(case["final"]))) <= 5.01e-7
    assert case["recip0_line"] == f"{recip[0]:1.10f}"


def test_pagerank_ref_trajectory_vs_reference():
    """The reference prints wb of the first vertices at the top of every
    iteration: wb_k = 0.3 + 0.7 * sum x_{k-1}/outdeg reproduced step by step."""
    case = load("pagerank.json")[0]
    V, irp, icol = parse_rc(case["rc_file"])
    recip = np.array([float(t) for t in case["recip_file"].split()[:V]])
    for k, wb in enumerate(case["iter_wb"], start=1):
        if k == 1:
            ours = np.full(V, 0.3)
        else:
            ours, _ = oracle.pagerank_ref(irp, icol, recip, max_iter=k - 1)
        assert np.max(np.abs(ours[:V] - np.array(wb[:V]))) <= 5.01e-7, k


def _nx_bc_directed_weighted(V, u, v, w):
    import networkx as nx
    G = nx.DiGraph()
    G.add_nodes_from(range(V))
    for a, b, c in zip(u, v, w):
        if G.has_edge(a, b):
            raise ValueError("parallel edges")
        G.add_edge(a, b, weight=c)
    r = nx.betweenness_centrality(G, normalized=False, weight="weight")
    return np.array([r[i] for i in range(V)])


@pytest.mark.parametrize("case", load("bc.json"), ids=lambda c: f"V{c['V']}E{len(c['src'])}")
def test_weighted_brandes_vs_reference_run_bc(case):
    """run_bc weights are v_to/V (graph/csr.c:42).  The oracle agrees with the
    reference where its predecessor duplication does not change sigma ratios
    (paths, stars, a diamond); everywhere it agrees with networkx."""
    V = case["V"]
    u, v = np.array(case["src"]), np.array(case["dst"])
    w = (v + 1) / V
    rp, col, ww = csr(V, u, v, w)
    rrp, rcol, rw = transpose(V, rp, col, ww)
    bc = oracle.brandes_weighted(rp, col, ww, rrp, rcol, rw, np.arange(V))
    ref = np.array(case["bc"])
    simple = len(set(zip(u.tolist(), v.tolist()))) == len(u)
    if simple:
        nxv = _nx_bc_directed_weighted(V, u.tolist(), v.tolist(), w.tolist())
        assert np.allclose(bc, nxv, rtol=1e-12, atol=1e-12)
    if V <= 6:
        assert np.allclose(bc, ref, atol=5.01e-7)


@pytest.mark.parametrize("case", load("pagerank_bin.json"), ids=lambda c: f"V{c['V']}")
def test_pagerank_ref_vs_reference_binary_pipeline(case):
    """bin = 1: read_csr_bin RECV keeps each target's in-edges in file order
    (graph/csr.c:271-291), read_recip_bin skips num_nodes distrib entries
    (partition/csr_partition_bin_mpi.c:312-332).  Exact recips, so the
    iteration counts and the printed trajectories match."""
    V = case["V"]
    p = np.array(case["rc_pairs"]).reshape(-1, 2) - 1
    irp, icol, _ = csr(V, p[:, 1], p[:, 0])
    recip = np.array(case["recip"])
    x, it = oracle.pagerank_ref(irp, icol, recip)
    assert it == case["iterations"]
    n = len(case["final"])
    assert np.max(np.abs(x[:n] - np.array(case["final"]))) <= 5.01e-7
    for k in (2, it // 2, it):
        ours, _ = oracle.pagerank_ref(irp, icol, recip, max_iter=k - 1)
        assert np.max(np.abs(ours[:n] - np.array(case["iter_wb"][k - 1][:n]))) <= 5.01e-7


def bin_bc_graph(case):
    """read_csr_bin SEND: rows by source in file order, and the weight of CSR
    slot i is 0.1 or 0.2 by the parity of i (graph/csr.c:293-298), whichever
    edge lands there."""
    V = case["V"]
    p = np.array(case["pairs"]).reshape(-1, 2) - 1
    rp, col, _ = csr(V, p[:, 0], p[:, 1])
    w = np.where(np.arange(len(col)) % 2 == 0, 0.1, 0.2)
    return V, rp, col, w


@pytest.mark.parametrize("case", load("bc_bin.json"), ids=lambda c: f"V{c['V']}E{len(c['pairs']) // 2}")
def test_weighted_brandes_vs_reference_binary_run_bc(case):
    """The oracle on the binary reader's graph: equal to the reference where
    no two shortest paths tie (path, star, diamond with unequal sides), equal
    to networkx on every simple graph.  On the 7-vertex double diamond both
    sides of each diamond weigh 0.1 + 0.2, and the reference's duplicated
    predecessors (sssp/csr_sspath.c:94-122) skew its sigma ratios."""
    V, rp, col, w = bin_bc_graph(case)
    rrp, rcol, rw = transpose(V, rp, col, w)
    bc = oracle.brandes_weighted(rp, col, w, rrp, rcol, rw, np.arange(V))
    src = np.repeat(np.arange(V), np.diff(rp))
    if len(set(zip(src.tolist(), col.tolist()))) == len(col):
        nxv = _nx_bc_directed_weighted(V, src.tolist(), col.tolist(), w.tolist())
        assert np.allclose(bc, nxv, rtol=1e-12, atol=1e-12)
    ref = np.array(case["bc"])
    if V <= 6:
        assert np.allclose(bc, ref, atol=5.01e-7)
    elif V == 7:
        assert not np.allclose(bc, ref, atol=1e-3)  # the tie case
        assert np.allclose(bc, [0, 2, 2, 9, 2.5, 1.5, 0], atol=1e-12)


# ------------------------------------------------------------------ mathematics

def rand_graph(seed, V, E, wlo=1, whi=255, allow_dup=True):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, V, E)
    d = rng.integers(0, V, E)
    m = s != d
    w = rng.integers(wlo, whi + 1, m.sum()).astype(np.uint32)
    return csr(V, s[m], d[m], w)


@pytest.mark.parametrize("seed", range(4))
def test_dijkstra_vs_scipy_and_bellman_ford(seed):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    V = 80
    rp, col, w = rand_graph(seed, V, 500)
    src = np.repeat(np.arange(V), np.diff(rp))
    best = {}
    for a, b, c in zip(src, col, w):
        best[(a, b)] = min(best.get((a, b), 1 << 40), int(c))
    k = np.array(list(best.keys()))
    A = sp.csr_matrix((np.array(list(best.values()), float), (k[:, 0], k[:, 1])), shape=(V, V))
    for s in (0, 17, 79):
        d = oracle.dijkstra(rp, col, w, s)
        ref = cg.dijkstra(A, indices=s)
        assert np.array_equal(np.where(d == oracle.UINF, np.inf, d.astype(float)), ref)
        # brute-force Bellman-Ford
        bf = np.full(V, np.inf)
        bf[s] = 0
        for _ in range(V):
            for a, b, c in zip(src, col, w):
                bf[b] = min(bf[b], bf[a] + c)
        assert np.array_equal(np.where(d == oracle.UINF, np.inf, d.astype(float)), bf)


@pytest.mark.parametrize("delta,light", [(1, 1), (7, 7), (16, 4), (40, 100), (1000, 1000)])
def test_delta_stepping_any_delta(delta, light):
    rp, col, w = rand_graph(5, 300, 3000)
    for s in (0, 99):
        a = oracle.dijkstra(rp, col, w, s)
        b, st = oracle.delta_stepping(rp, col, w, s, delta, light, stats=True)
        assert np.array_equal(a, b)
        assert st[2] >= st[3] > 0


def test_sssp_invariants():
    """Triangle inequality on every edge, and every reached vertex has a tight in-edge."""
    rp, col, w = rand_graph(8, 500, 4000)
    d = oracle.dijkstra(rp, col, w, 3).astype(np.int64)
    src = np.repeat(np.arange(500), np.diff(rp))
    fin = d != oracle.UINF
    m = fin[src]
    assert np.all(d[col[m]] <= d[src[m]] + w[m])
    tight = np.zeros(500, bool)
    tm = m & (d[col] == d[src] + w)
    tight[col[tm]] = True
    reached = np.where(fin)[0]
    assert all(tight[v] for v in reached if v != 3)


def test_bfs_levels_vs_networkx():
    import networkx as nx
    G = nx.gnm_random_graph(300, 900, seed=4)
    e = np.array(G.edges())
    rp, col, _ = csr(300, np.r_[e[:, 0], e[:, 1]], np.r_[e[:, 1], e[:, 0]])
    lv = oracle.bfs_levels(rp, col, 0)
    ref = nx.single_source_shortest_path_length(G, 0)
    assert all(lv[v] == ref.get(v, -1) for v in range(300))


def test_brandes_bfs_vs_networkx_undirected_and_directed():
    import networkx as nx
    G = nx.gnm_random_graph(150, 500, seed=6)
    e = np.array(G.edges())
    rp, col, _ = csr(150, np.r_[e[:, 0], e[:, 1]], np.r_[e[:, 1], e[:, 0]])
    bc = oracle.brandes_bfs(rp, col, np.arange(150))
    ref = nx.betweenness_centrality(G, normalized=False)
    # the reference does not halve undirected sums (csr_bc.c:57-61)
    assert np.allclose(bc, 2 * np.array([ref[i] for i in range(150)]), rtol=1e-12)
    D = nx.gnm_random_graph(120, 600, seed=7, directed=True)
    e = np.array(D.edges())
    rp, col, _ = csr(120, e[:, 0], e[:, 1])
    rrp, rcol, _ = transpose(120, rp, col)
    bc = oracle.brandes_bfs(rp, col, np.arange(120), rrp, rcol)
    ref = nx.betweenness_centrality(D, normalized=False)
    assert np.allclose(bc, [ref[i] for i in range(120)], rtol=1e-12)


def test_brandes_closed_forms():
    # path 0-1-...-(n-1): BC(i) = 2 i (n-1-i) unhalved
    n = 9
    a = np.arange(n - 1)
    rp, col, _ = csr(n, np.r_[a, a + 1], np.r_[a + 1, a])
    bc = oracle.brandes_bfs(rp, col, np.arange(n))
    assert np.array_equal(bc, 2.0 * np.arange(n) * (n - 1 - np.arange(n)))
    # star: centre = (n-1)(n-2)
    n = 7
    leaves = np.arange(1, n)
    rp, col, _ = csr(n, np.r_[np.zeros(n - 1, int), leaves], np.r_[leaves, np.zeros(n - 1, int)])
    bc = oracle.brandes_bfs(rp, col, np.arange(n))
    assert bc[0] == (n - 1) * (n - 2) and np.all(bc[1:] == 0)


def test_brandes_parallel_edges_count_paths():
    # 0 =2 edges=> 1 -> 2 and 0 -> 3 -> 2: sigma(2) = 2 + 1 paths
    rp, col, _ = csr(4, [0, 0, 0, 1, 3], [1, 1, 3, 2, 2])
    rrp, rcol, _ = transpose(4, rp, col)
    bc = oracle.brandes_bfs(rp, col, [0], rrp, rcol)
    assert np.allclose(bc, [0, 2 / 3, 0, 1 / 3])


def test_pagerank_cfg_vs_networkx_and_invariants():
    import networkx as nx
    D = nx.gnm_random_graph(200, 1200, seed=5, directed=True)
    e = np.array(D.edges())
    od = np.bincount(e[:, 0], minlength=200).astype(np.int32)
    irp, icol, _ = csr(200, e[:, 1], e[:, 0])
    x, it, r = oracle.pagerank_cfg(irp, icol, od, 0.85, 1e-13, 1000)
    ref = nx.pagerank(D, alpha=0.85, tol=1e-15, max_iter=1000)
    assert np.abs(x - np.array([ref[i] for i in range(200)])).sum() < 1e-10
    assert abs(x.sum() - 1) < 1e-12 and np.all(x >= 0.15 / 200 - 1e-15)
    assert r < 1e-13


def test_pagerank_cfg_closed_forms():
    # directed cycle: uniform, converged after one iteration
    n = 12
    a = np.arange(n)
    irp, icol, _ = csr(n, (a + 1) % n, a)
    x, it, r = oracle.pagerank_cfg(irp, icol, np.ones(n, np.int32), 0.85, 1e-12, 100)
    assert np.allclose(x, 1 / n) and it == 1
    # two nodes 0 -> 1, node 1 dangling: fixed point of x1 = .075 + .85(x0 + x1/2),
    # x0 = .075 + .85 x1/2
    irp, icol, _ = csr(2, [1], [0])
    x, it, r = oracle.pagerank_cfg(irp, icol, np.array([1, 0], np.int32), 0.85, 1e-15, 1000)
    x1 = (0.075 + 0.85 * 0.075) / (1 - 0.85 * 0.85 / 2 - 0.85 / 2)
    x0 = 0.075 + 0.85 * x1 / 2
    assert abs(x[1] - x1) < 1e-12 and abs(x[0] - x0) < 1e-12


def test_pagerank_ref_closed_form_and_fixed_point():
    # directed cycle with recip = 1: x = 1 for all, one iteration
    n = 5
    a = np.arange(n)
    irp, icol, _ = csr(n, (a + 1) % n, a)
    x, it = oracle.pagerank_ref(irp, icol, np.ones(n))
    assert np.array_equal(x, np.ones(n)) and it == 1
    # converged x satisfies x = 0.3 + 0.7 * sum x[u] recip[u] within THRESH scale
    rng = np.random.default_rng(1)
    s, d = rng.integers(0, 50, 300), rng.integers(0, 50, 300)
    m = s != d
    od = np.bincount(s[m], minlength=50)
    recip = np.where(od > 0, 1.0 / np.maximum(od, 1), np.inf)
    irp, icol, _ = csr(50, d[m], s[m])
    x, it = oracle.pagerank_ref(irp, icol, recip)
    fx = 0.3 + 0.7 * np.array([sum(x[u] * recip[u] for u in icol[irp[v]:irp[v + 1]])
                               for v in range(50)])
    assert np.max(np.abs(fx - x)) < 1e-4


def test_comp_distrib_invariants():
    rng = np.random.default_rng(3)
    for _ in range(20):
        V = int(rng.integers(1, 400))
        deg = rng.integers(0, 50, V)
        E = int(deg.sum())
        p = int(rng.integers(1, 12))
        d = oracle.comp_distrib(deg, E, p)
        assert d[-1] == V - 1
        cuts = [x for x in d[:-1] if x > 0] if p > 1 else []
        assert cuts == sorted(cuts)
        q = E // p
        start = 0
        for c in cuts:  # every cut closes a part whose sum just exceeded q
            assert deg[start:c + 1].sum() > q and deg[start:c].sum() <= q
            start = c + 1


def test_merge_path_split_invariants():
    rng = np.random.default_rng(2)
    deg = rng.integers(0, 40, 500) * (rng.random(500) < 0.7)
    rp = np.r_[0, np.cumsum(deg)].astype(np.int32)
    V, E = 500, int(rp[-1])
    for items in (1, 5, 64, 1000, V + E + 3):
        nt = (V + E + items - 1) // items
        tr, te = oracle.merge_path_split(rp, items, nt)
        assert tr[0] == 0 and te[0] == 0 and tr[-1] == V and te[-1] == E
        assert np.all(np.diff(tr) >= 0) and np.all(np.diff(te) >= 0)
        for t in range(nt + 1):
            i, j = tr[t], te[t]
            assert i + j == min(t * items, V + E)
            assert rp[i] <= j and (i == V or j <= rp[i + 1])


def test_mlp_gradient_vs_finite_differences():
    rng = np.random.default_rng(0)
    n, din, h, dout = 7, 5, 6, 2
    X = rng.standard_normal((n, din)).astype(np.float32)
    T = rng.standard_normal((n, dout)).astype(np.float32)
    P = [rng.standard_normal(s).astype(np.float32) * 0.5
         for s in [(h, din), (h,), (h, h), (h,), (dout, h), (dout,)]]
    loss, grads = oracle.mlp_loss_grad(X, T, *P)
    eps = 1e-3
    for k in range(6):
        flat = P[k].reshape(-1)
        for idx in range(0, flat.size, max(1, flat.size // 5)):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = oracle.mlp_loss_grad(X, T, *P)
            flat[idx] = old - eps
            lm, _ = oracle.mlp_loss_grad(X, T, *P)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[k].reshape(-1)[idx]) < 2e-3 * max(1, abs(fd)), (k, idx)
