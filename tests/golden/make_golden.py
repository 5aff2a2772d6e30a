"""Writes golden vectors from the reference build (oracle/_ref/ref_harness).

Run in the container that has the reference sources (build with
`make -C oracle/ref`).  Only data is written: inputs and the reference's
outputs, as JSON under tests/golden/.  Cases:
  sssp_list.json   run_sspath distances + raw predecessor lists (csr_sspath.c:38-122)
  distrib.json     comp_distrib_csr + bin_search (csr_partition.c:5-30, utils.c:222-239)
  separate.json    separate_csr (csr.c:75-130)
  pagerank.json    the reference CLI pipeline -k / -r / -i (pagerank_mpi.c:667-1044):
                   chunked rc and recip files, every ITER line, final values,
                   iteration count
  bc.json          the reference CLI -e (run_bc, csr_bc.c:16-68) on text graphs
  remove_duplicate.json
                   the reference CLI -y (remove_duplicate,
                   csr_partition_bin_mpi.c:31-106): binary edge files in, the
                   deduplicated file's bytes out
  chunk.json       the reference CLI -k (write_recip + chunk_files,
                   csr_partition.c:32-164) on text graphs: every recip, part
                   and send file it writes
  pagerank_bin.json
                   the reference CLI -l -i (bin = 1: read_csr_bin,
                   csr.c:211-299, and read_recip_bin,
                   csr_partition_bin_mpi.c:312-332, feeding
                   run_pagerank_csr_mpi): binary in-edge files whose pairs are
                   not in row order, exact 1/out-degree recips
  bc_bin.json      the reference CLI -l -e (read_csr_bin SEND feeding run_bc,
                   csr_bc.c:16-68): binary edge files in shuffled order, so
                   the per-slot 0.1 / 0.2 weights land on other edges than in
                   a sorted file
Arguments: fixture names to (re)write; none = all.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(args, stdin="", cwd=None, check=True):
    r = subprocess.run([HARNESS, *args], input=stdin, capture_output=True, text=True,
                       cwd=cwd, timeout=120)
    if check and r.returncode != 0:
        raise RuntimeError(f"{args}: rc={r.returncode}\n{r.stdout}\n{r.stderr}")
    return r


def edges_sorted(V, E, rng, wfun):
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    m = src != dst
    src, dst = src[m], dst[m]
    w = wfun(len(src))
    o = np.argsort(src, kind="stable")
    return src[o], dst[o], w[o]


def sssp_cases():
    cases = []
    rng = np.random.default_rng(2024)
    specs = [(8, 20, "frac", 0.1), (30, 120, "frac", 0.15), (60, 300, "int", 16.0),
             (200, 1200, "int", 16.0), (120, 600, "frac", 0.25)]
    for V, E, kind, delta in specs:
        if kind == "frac":
            wf = lambda n: np.round(rng.random(n) * 0.9 + 0.05, 3)
        else:
            wf = lambda n: rng.integers(1, 256, n).astype(np.float64)
        s, d, w = edges_sorted(V, E, rng, wf)
        for src in sorted(set([0, int(rng.integers(0, V)), V - 1])):
            txt = "".join(f"{int(a)} {int(b)} {float(c)!r}\n" for a, b, c in zip(s, d, w))
            with tempfile.TemporaryDirectory() as td:
                r = run(["sssp_list", str(V), str(len(s)), str(src), repr(delta)], txt, cwd=td)
            lines = r.stdout.strip().splitlines()
            dist = [float(x) for x in lines[0].split()[1:]]
            preds = [[int(x) for x in ln.split()[2:]] for ln in lines[1:]]
            cases.append(dict(V=V, src=[int(x) for x in s], dst=[int(x) for x in d],
                              w=[float(x) for x in w], source=src, delta=delta, dist=dist,
                              pred_raw=preds))
    return cases


def rmat_edges(scale, ef, seed):
    """Small R-MAT (a,b,c = .57,.19,.19) for fixtures; numpy RNG."""
    rng = np.random.default_rng(seed)
    n = (1 << scale) * ef
    u = np.zeros(n, np.int64)
    v = np.zeros(n, np.int64)
    for _ in range(scale):
        r = rng.random(n)
        bu = (r >= 0.76).astype(np.int64)
        bv = (((r >= 0.57) & (r < 0.76)) | (r >= 0.95)).astype(np.int64)
        u = (u << 1) | bu
        v = (v << 1) | bv
    m = u != v
    return u[m], v[m]


def sssp_rmat_case():
    u, v = rmat_edges(10, 8, 77)
    rng = np.random.default_rng(78)
    w = rng.integers(1, 256, len(u)).astype(np.float64)
    o = np.argsort(u, kind="stable")
    u, v, w = u[o], v[o], w[o]
    V = 1024
    txt = "".join(f"{int(a)} {int(b)} {float(c)!r}\n" for a, b, c in zip(u, v, w))
    with tempfile.TemporaryDirectory() as td:
        r = run(["sssp_list", str(V), str(len(u)), "0", "16.0"], txt, cwd=td)
    lines = r.stdout.strip().splitlines()
    dist = [float(x) for x in lines[0].split()[1:]]
    return dict(V=V, src=[int(x) for x in u], dst=[int(x) for x in v], w=[float(x) for x in w],
                source=0, delta=16.0, dist=dist, pred_raw=None)


def distrib_cases():
    out = []
    rng = np.random.default_rng(7)
    for V, p in [(6, 3), (50, 4), (200, 8), (200, 1), (1000, 16), (37, 5)]:
        sdeg = rng.integers(0, 30, V) * (rng.random(V) < 0.8)
        rdeg = rng.permutation(sdeg)
        for typ in (1, 0):
            txt = "".join(f"{a} {b}\n" for a, b in zip(sdeg, rdeg))
            r = run(["distrib", str(V), str(p), str(typ)], txt)
            rows = {ln.split()[0]: [int(x) for x in ln.split()[1:]]
                    for ln in r.stdout.strip().splitlines()}
            out.append(dict(V=V, p=p, type=typ, send_deg=[int(x) for x in sdeg],
                            recv_deg=[int(x) for x in rdeg], **rows))
    return out


def separate_cases():
    out = []
    rng = np.random.default_rng(9)
    for V, E, delta in [(5, 6, 0.1), (40, 200, 0.3), (100, 800, 0.55)]:
        s, d, w = edges_sorted(V, E, rng, lambda n: np.round(rng.random(n), 3))
        txt = "".join(f"{int(a)} {int(b)} {float(c)!r}\n" for a, b, c in zip(s, d, w))
        with tempfile.TemporaryDirectory() as td:
            r = run(["separate", str(V), str(len(s)), repr(delta)], txt, cwd=td)
        rows = {}
        for ln in r.stdout.strip().splitlines():
            k, *vals = ln.split()
            rows[k] = [float(x) if k.endswith("_w") else int(x) for x in vals]
        out.append(dict(V=V, src=[int(x) for x in s], dst=[int(x) for x in d],
                        w=[float(x) for x in w], delta=delta, **rows))
    return out


def text_graph(V, u, v):
    """The reference's text format: p sp V E off gV / a u v w (1-based)."""
    lines = [f"p sp {V} {len(u)} 1 {V}"]
    lines += [f"a {a + 1} {b + 1} 1" for a, b in zip(u, v)]
    return "\n".join(lines) + "\n"


def pagerank_cases():
    out = []
    rng = np.random.default_rng(11)
    graphs = [(6, [0, 0, 1, 2, 3, 3, 4, 5, 5], [1, 2, 2, 0, 2, 4, 5, 3, 0])]
    for V, E in [(8, 20), (10, 30), (10, 14), (9, 40)]:
        u = rng.integers(0, V, E)
        v = rng.integers(0, V, E)
        m = u != v
        o = np.argsort(u[m], kind="stable")
        graphs.append((V, u[m][o].tolist(), v[m][o].tolist()))
    # a larger one: only the first 10 ranks are printed by the reference
    ur, vr = rmat_edges(8, 8, 5)
    o = np.argsort(ur, kind="stable")
    graphs.append((256, ur[o].tolist(), vr[o].tolist()))
    for V, u, v in graphs:
        with tempfile.TemporaryDirectory() as td:
            g = os.path.join(td, "g") + "/"
            for sub in ("datex", "parex", "log"):
                os.makedirs(g + sub)
            with open(g + "g.txt", "w") as f:
                f.write(text_graph(V, u, v))
            base = ["cli", "-f", "g.txt", "-b", g, "-n", "1"]
            run([*base[:1], "-k", *base[1:]], cwd=td)
            run([*base[:1], "-r", *base[1:]], cwd=td)
            # print_result divides by recv_size == 0 at one rank (SIGFPE after
            # the results are printed): keep the printed part
            r = run([*base[:1], "-i", *base[1:]], cwd=td, check=False)
            rc_file = open(g + "datex/g.txt_rc_0_1").read()
            recip_file = open(g + "parex/g.txt_recip_0_1").read()
        text = r.stdout
        iters = [[float(t.split("|")[0]) for t in ln.split()[2:]]
                 for ln in text.splitlines() if ln.startswith("ITER")]
        fin = text.split("Final pagerank values:")[1].splitlines()[1].split()
        final = [float(t.split("|")[1]) for t in fin]
        it = int(re.search(r"iterations taken: (\d+)", text).group(1))
        out.append(dict(V=V, src=[int(x) for x in u], dst=[int(x) for x in v],
                        rc_file=rc_file, recip_file=recip_file, recip0_line=text.splitlines()[0],
                        iter_wb=iters, final=final[:min(10, V)], iterations=it))
    return out


def bc_cases():
    out = []
    rng = np.random.default_rng(13)
    graphs = [(5, [0, 1, 2, 3], [1, 2, 3, 4]),                       # path
              (6, [0, 0, 0, 0, 0], [1, 2, 3, 4, 5]),                 # star
              (6, [0, 0, 1, 2, 3, 4], [1, 2, 3, 3, 4, 5])]           # diamond + tail
    for V, E in [(12, 30), (20, 60)]:
        u = rng.integers(0, V, E)
        v = rng.integers(0, V, E)
        m = u != v
        o = np.argsort(u[m], kind="stable")
        graphs.append((V, u[m][o].tolist(), v[m][o].tolist()))
    for V, u, v in graphs:
        with tempfile.TemporaryDirectory() as td:
            g = os.path.join(td, "g") + "/"
            os.makedirs(g)
            with open(g + "g.txt", "w") as f:
                f.write(text_graph(V, u, v))
            r = run(["cli", "-e", "-f", "g.txt", "-b", g], cwd=td)
        vals = r.stdout.split("betweeness centrality is:")[1].split()
        bc = [float(t.split(":")[1]) for t in vals]
        out.append(dict(V=V, src=[int(x) for x in u], dst=[int(x) for x in v], bc=bc))
    return out


def bin_graph_bytes(V, pairs):
    """read_csr_bin's format (csr.c:221-229): int64 v_num, e_num, offset,
    g_v_size, then e_num (from, to) int64 pairs, both 1-based."""
    hdr = np.array([V, len(pairs), 1, V], np.int64)
    return hdr.tobytes() + np.asarray(pairs, np.int64).reshape(-1, 2).tobytes()


def pagerank_bin_cases():
    out = []
    rng = np.random.default_rng(17)
    graphs = [(6, [0, 0, 1, 2, 3, 3, 4, 5, 5], [1, 2, 2, 0, 2, 4, 5, 3, 0])]
    for V, E in [(10, 30), (12, 20)]:
        u = rng.integers(0, V, E)
        v = rng.integers(0, V, E)
        m = u != v
        o = np.argsort(u[m], kind="stable")
        graphs.append((V, u[m][o].tolist(), v[m][o].tolist()))
    ur, vr = rmat_edges(8, 8, 9)
    o = np.argsort(ur, kind="stable")
    graphs.append((256, ur[o].tolist(), vr[o].tolist()))
    for V, u, v in graphs:
        u = np.asarray(u, np.int64)
        v = np.asarray(v, np.int64)
        E = len(u)
        deg = np.bincount(u, minlength=V).astype(np.float64)
        with np.errstate(divide="ignore"):
            recip = 1.0 / deg
        perm = rng.permutation(E)  # in-edge file order: not grouped by target
        rc_pairs = np.stack([u[perm] + 1, v[perm] + 1], 1)
        with tempfile.TemporaryDirectory() as td:
            g = os.path.join(td, "g") + "/"
            for sub in ("datex", "parex", "log"):
                os.makedirs(g + sub)
            with open(g + "g.txt", "w") as f:
                f.write(text_graph(V, u.tolist(), v.tolist()))
            base = ["cli", "-f", "g.txt", "-b", g, "-n", "1"]
            # the text pipeline writes the files the binary run also opens
            # (partition index, send bounds); then the graph, recip and send
            # files are replaced by their binary forms
            run([*base[:1], "-k", *base[1:]], cwd=td)
            run([*base[:1], "-r", *base[1:]], cwd=td)
            distrib = [int(x) for x in open(g + "parex/g.txt_recip_0_1").read().split()[V:V + 1]]
            send = [int(x) for x in open(g + "parex/g.txt_send_1").read().split()[:1]]
            with open(g + "datex/g.txt_sd_0_1", "wb") as f:
                f.write(bin_graph_bytes(V, np.stack([u + 1, v + 1], 1)))
            with open(g + "datex/g.txt_rc_0_1", "wb") as f:
                f.write(bin_graph_bytes(V, rc_pairs))
            with open(g + "parex/g.txt_recip_0_1", "wb") as f:
                f.write(np.array(distrib, np.int64).tobytes() + recip.tobytes())
            with open(g + "parex/g.txt_send_1", "wb") as f:
                f.write(np.array(send, np.int64).tobytes())
            r = run(["cli", "-l", "-i", "-f", "g.txt", "-b", g, "-n", "1"], cwd=td, check=False)
        text = r.stdout
        iters = [[float(t.split("|")[0]) for t in ln.split()[2:]]
                 for ln in text.splitlines() if ln.startswith("ITER")]
        fin = text.split("Final pagerank values:")[1].splitlines()[1].split()
        final = [float(t.split("|")[1]) for t in fin]
        it = int(re.search(r"iterations taken: (\d+)", text).group(1))
        out.append(dict(V=V, rc_pairs=rc_pairs.ravel().tolist(), distrib=distrib,
                        recip=[float(x) for x in recip], iter_wb=iters,
                        final=final[:min(10, V)], iterations=it))
    return out


def bc_bin_cases():
    out = []
    rng = np.random.default_rng(19)
    graphs = [(5, [0, 1, 2, 3], [1, 2, 3, 4]),                       # path
              (6, [0, 0, 0, 0, 0], [1, 2, 3, 4, 5]),                 # star
              (6, [0, 0, 1, 2, 3, 4], [1, 2, 3, 3, 4, 5]),           # diamond + tail
              (7, [0, 0, 1, 2, 3, 3, 4, 5], [1, 2, 3, 3, 4, 5, 6, 6])]
    for V, E in [(12, 30), (20, 60)]:
        u = rng.integers(0, V, E)
        v = rng.integers(0, V, E)
        m = u != v
        graphs.append((V, u[m].tolist(), v[m].tolist()))
    for V, u, v in graphs:
        u = np.asarray(u, np.int64)
        v = np.asarray(v, np.int64)
        perm = rng.permutation(len(u))
        pairs = np.stack([u[perm] + 1, v[perm] + 1], 1)
        with tempfile.TemporaryDirectory() as td:
            g = os.path.join(td, "g") + "/"
            os.makedirs(g)
            with open(g + "g.bin", "wb") as f:
                f.write(bin_graph_bytes(V, pairs))
            r = run(["cli", "-l", "-e", "-f", "g.bin", "-b", g], cwd=td)
        vals = r.stdout.split("betweeness centrality is:")[1].split()
        bc = [float(t.split(":")[1]) for t in vals]
        out.append(dict(V=V, pairs=pairs.ravel().tolist(), bc=bc))
    return out


def chunk_cases():
    """The reference CLI -k (main.c:116-136: scan_csr_idx SEND/RECV,
    comp_distrib_csr, write_recip, chunk_files) on text graphs, num_nodes p:
    every file it writes, as text."""
    out = []
    rng = np.random.default_rng(41)
    graphs = []
    # hand-made: weights kept as written, a vertex without out-edges (inf recip)
    graphs.append(("hand", "p sp 6 8 1 6\na 1 2 5\na 1 3 1\na 2 3 2\na 3 1 4\n"
                           "a 4 5 1\na 5 6 3\na 6 1 1\na 6 4 2\n", (1, 2, 3)))
    # a comment line first, and edge lines past the header's count (not read)
    graphs.append(("extra", "c made for the chunking test 10 20\np sp 5 4 1 5\na 1 2 7\n"
                            "a 2 3 8\na 3 4 9\na 4 5 10\na 5 1 11\na 1 5 12\n", (1, 2)))
    for V, E in [(20, 60), (40, 200)]:
        u = rng.integers(0, V, E)
        v = rng.integers(0, V, E)
        w = rng.integers(1, 100, E)
        lines = [f"p sp {V} {E} 1 {V}"] + [f"a {a + 1} {b + 1} {c}" for a, b, c in zip(u, v, w)]
        graphs.append((f"rand{V}", "\n".join(lines) + "\n", (1, 2, 3, 4)))
    ur, vr = rmat_edges(8, 8, 3)
    lines = [f"p sp 256 {len(ur)} 1 256"] + [f"a {a + 1} {b + 1} 1" for a, b in zip(ur, vr)]
    graphs.append(("rmat8", "\n".join(lines) + "\n", (1, 4, 7)))
    for name, text, parts in graphs:
        for p in parts:
            with tempfile.TemporaryDirectory() as td:
                g = os.path.join(td, "g") + "/"
                for sub in ("datex", "parex", "log"):
                    os.makedirs(g + sub)
                with open(g + "g.txt", "w") as f:
                    f.write(text)
                run(["cli", "-k", "-f", "g.txt", "-b", g, "-n", str(p)], cwd=td)
                files = {}
                for sub in ("datex", "parex"):
                    for fn in sorted(os.listdir(g + sub)):
                        with open(g + sub + "/" + fn) as f:
                            files[sub + "/" + fn] = f.read()
            out.append(dict(name=name, num_nodes=p, text=text, files=files))
    return out


def remove_duplicate_cases():
    out = []
    rng = np.random.default_rng(31)
    inputs = [
        ("dups", rng.integers(0, 40, (600, 2))),            # many repeats
        ("few", rng.integers(0, 1 << 20, (500, 2))),        # almost none
        ("one_pair", np.tile([[7, 3]], (50, 1))),            # a single pair repeated
        ("reversed", np.array([[1, 2], [2, 1], [1, 2], [2, 1], [3, 3], [3, 3]])),
    ]
    for name, pairs in inputs:
        pairs = np.asarray(pairs, np.int64)
        for nodes, scale in ((1, 10), (2, 6)):
            with tempfile.TemporaryDirectory() as td:
                g = os.path.join(td, "g") + "/"
                os.makedirs(g + "datex")
                with open(g + "edges_0", "wb") as f:
                    f.write(np.int64(len(pairs)).tobytes())
                    f.write(pairs.tobytes())
                # with more than one node main() goes on to the partitioning
                # steps, which stop on files this case does not make; the
                # deduplicated file is complete by then
                run(["cli", "-y", "-f", "edges", "-b", g, "-n", str(nodes), "-s", str(scale)],
                    cwd=td, check=nodes == 1)
                raw = open(g + f"datex/edges_phy_0_{nodes}", "rb").read()
            out.append(dict(name=name, num_nodes=nodes, g_v_num=1 << scale,
                            pairs=pairs.ravel().tolist(),
                            output=np.frombuffer(raw, np.int64).tolist()))
    return out


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference first: make -C oracle/ref")
    builders = {
        "sssp_list.json": lambda: sssp_cases() + [sssp_rmat_case()],
        "distrib.json": distrib_cases,
        "separate.json": separate_cases,
        "pagerank.json": pagerank_cases,
        "bc.json": bc_cases,
        "remove_duplicate.json": remove_duplicate_cases,
        "chunk.json": chunk_cases,
        "pagerank_bin.json": pagerank_bin_cases,
        "bc_bin.json": bc_bin_cases,
    }
    wanted = sys.argv[1:] or list(builders)
    fixtures = {name: builders[name]() for name in wanted}
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump({"generator": "tests/golden/make_golden.py (reference build oracle/_ref)",
                       "cases": data}, f)
        print(name, len(data), "cases")


if __name__ == "__main__":
    main()
