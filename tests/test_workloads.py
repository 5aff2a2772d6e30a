"""Pins of the synthetic inputs (graphbench_amd/workloads.py).

The reference ships no input graphs, so the generators define every measured
workload.  A SHA-256 of each generated CSR (V, E, row_ptr, col, w) at small
scales, of the C1 graph itself and of the 512-graph spec list, makes any
change to a generator, its RNG or its canonical row order a test failure.
The structural checks say what the pinned bytes are: R-MAT without
self-loops, weights in range, rows sorted by (destination, weight).
"""
import hashlib
import json

import numpy as np
import pytest

import graphbench_amd as gb
from graphbench_amd import workloads as w

PINS = {
    "rmat10_u1_255_seed42": (lambda: gb.gen_rmat(10, 16, seed=42, weights="uniform", wlo=1, whi=255),
                             1024, 16251,
                             "5fc1b31cfe58cf6cac15a4abeb9b4caa61ab6986723dd5e3dadcf2e6bb415c2c"),
    "rmat11_u1_255_seed42": (lambda: gb.gen_rmat(11, 16, seed=42, weights="uniform", wlo=1, whi=255),
                             2048, 32577,
                             "1340dcc6157d6f997782d57965ce38b95236b8c685b881d47eb07f77acdcbb90"),
    "rmat12_u1_255_seed42": (lambda: gb.gen_rmat(12, 16, seed=42, weights="uniform", wlo=1, whi=255),
                             4096, 65277,
                             "771306477a4cce1985e23162c985b83aefad93e38b5e547b35ebfe9f97793a6a"),
    "rmat12_none_seed7": (lambda: gb.gen_rmat(12, 16, seed=7, weights="none"),
                          4096, 65301,
                          "b9907f3370812e212d39f0b68a0b02bd6af45563f15ac9da15e81f4f49313385"),
    "rmat10_sym_seed11": (lambda: gb.gen_rmat(10, 16, seed=11, weights="none", symmetric=True),
                          1024, 32526,
                          "fdcb6815bc810395fa008b1f2828cd97f0a3fae86d827484528c879285d982ec"),
    "rmat12_sym_seed11": (lambda: gb.gen_rmat(12, 16, seed=11, weights="none", symmetric=True),
                          4096, 130660,
                          "c08a4b0e08d67fb460bb013c8ef1d2855ee2769c9fc948c73b74f4b503c6d3ee"),
    "rmat12_mix_seed99": (lambda: gb.gen_rmat(12, 16, seed=99, weights="mix", wlo=1, whi=255),
                          4096, 65319,
                          "c4f814f460357d785831fa27f0cf431f66cb7c4aac6ee5255154685df21bfc19"),
    "er3000_deg12_seed5001": (lambda: gb.gen_er(3000, 12, seed=5001, weights="uniform", wlo=1, whi=1023),
                              3000, 36000,
                              "7d5e657bc7867511d469c653b555752040688c1853280f94b0a50eac273fafe4"),
    "rmat_V3000_ef9_seed5000": (lambda: gb.gen_rmat(12, 9, seed=5000, weights="uniform", wlo=1,
                                                    whi=1023, nverts=3000),
                                3000, 26902,
                                "06e0887ae610181a0c666132f299a3a146f11a75573d1e765e2b7f782523af51"),
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_generated_csr_pinned(name):
    gen, V, E, digest = PINS[name]
    h = gen()
    assert (h.V, h.E) == (V, E)
    src = np.repeat(np.arange(h.V), np.diff(h.row_ptr))
    assert h.row_ptr[0] == 0 and np.all(np.diff(h.row_ptr) >= 0)
    assert np.all((h.col >= 0) & (h.col < h.V))
    if not name.startswith("er"):
        assert not np.any(src == h.col), "self-loop kept"
    # canonical row order: by destination, then weight
    key = h.col.astype(np.int64) * (1 << 32) + (0 if h.w is None else h.w.astype(np.int64))
    same_row = src[1:] == src[:-1]
    assert np.all(key[1:][same_row] >= key[:-1][same_row])
    if h.w is not None:
        hi = 65535 if "mix" in name else (255 if "255" in name else 1023)
        assert h.w.min() >= 1 and h.w.max() <= hi
    if "sym" in name:
        fwd = np.sort(src.astype(np.int64) * h.V + h.col)
        bwd = np.sort(h.col.astype(np.int64) * h.V + src)
        assert np.array_equal(fwd, bwd), "symmetrised graph is not symmetric"
    assert w.csr_sha256(h) == digest


def test_c1_graph_and_sources_pinned():
    h, srcs = w.c1()
    assert (h.V, h.E) == (65536, 1048039)
    assert w.csr_sha256(h) == "38a213bbae49a54ae8597658daf9a154c947546012fe470b9376c21a1a11ae22"
    assert srcs.tolist() == [20911, 46384, 43378, 15476, 5436, 10881, 10938, 22860]
    assert np.all(np.diff(h.row_ptr)[srcs] > 0)


def test_c4_specs_pinned():
    specs = w.c4_specs()
    assert len(specs) == 512
    assert sum(s[1] for s in specs) == 81763101
    assert all(1024 <= s[1] <= 1 << 20 and 8 <= s[2] <= 32 for s in specs)
    assert [s[0] for s in specs[:4]] == ["rmat", "er", "rmat", "er"]
    digest = hashlib.sha256(json.dumps(specs).encode()).hexdigest()
    assert digest == "da631367effb272114810726a8150189ce1d96ce3d07197ab7bd69109b947c14"
