"""The five measured workloads, defined once for the benchmark and the tests.

Each function generates its graph(s) on the host with the library's
counter-based generators (gb_gen_rmat / gb_gen_er), so a workload is fixed by
its parameters alone; ``csr_sha256`` fingerprints a generated CSR so that a
change to a generator shows up as a changed pin.

  C1  SSSP: directed R-MAT 16, edge factor 16, weights U[1,255], seed 42,
      8 sources (out-degree > 0), delta = max_w / avg_degree
  C2  PageRank: directed R-MAT 22, edge factor 16, unweighted, seed 7
  C3  Brandes: undirected R-MAT 20, edge factor 16, symmetrised, seed 11,
      256 sources (degree > 0)
  C4  512 uneven graphs, one source each: V = round(2^(10 + 10 U)), even
      indices R-MAT with edge factor U{8..32}, odd indices Erdos-Renyi with
      average degree U{8..32}, weights U[1,1023]
  C5  directed R-MAT 24, edge factor 16, seed 99, weights a 50/50 mix of
      U[1,255] and log-uniform [1,65535]; 32 SSSP sources plus PageRank
"""
from __future__ import annotations

import hashlib

import numpy as np

from . import HostCSR, gen_er, gen_rmat, pick_sources


def csr_sha256(h: HostCSR) -> str:
    """SHA-256 over V, E, row_ptr, col and (when present) w, little-endian."""
    m = hashlib.sha256()
    m.update(np.array([h.V, h.E], "<i8").tobytes())
    m.update(np.ascontiguousarray(h.row_ptr, "<i4").tobytes())
    m.update(np.ascontiguousarray(h.col, "<i4").tobytes())
    if h.w is not None:
        m.update(np.ascontiguousarray(h.w, "<u4").tobytes())
    return m.hexdigest()


def c1():
    h = gen_rmat(16, 16, seed=42, weights="uniform", wlo=1, whi=255)
    return h, pick_sources(h, 8, 42)


def c2():
    return gen_rmat(22, 16, seed=7, weights="none")


def c3():
    h = gen_rmat(20, 16, seed=11, weights="none", symmetric=True)
    return h, pick_sources(h, 256, 11)


def c4_specs(n: int = 512, seed: int = 2024):
    """(kind, V, degree, seed) of every graph of the 512-graph batch."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        V = int(round(2.0 ** (10 + 10 * rng.random())))
        deg = int(rng.integers(8, 33))
        out.append(("rmat" if i % 2 == 0 else "er", V, deg, 5000 + i))
    return out


def c4_graph(kind: str, V: int, deg: int, seed: int) -> HostCSR:
    if kind == "rmat":
        sc = max(1, (V - 1).bit_length())
        return gen_rmat(sc, deg, seed=seed, weights="uniform", wlo=1, whi=1023, nverts=V)
    return gen_er(V, deg, seed=seed, weights="uniform", wlo=1, whi=1023)


def c4():
    """The 512 graphs and one source per graph."""
    hs = [c4_graph(*s) for s in c4_specs()]
    srcs = [int(pick_sources(h, 1, 7 + i)[0]) for i, h in enumerate(hs)]
    return hs, srcs


def c5():
    h = gen_rmat(24, 16, seed=99, weights="mix", wlo=1, whi=255)
    return h, pick_sources(h, 32, 99)
