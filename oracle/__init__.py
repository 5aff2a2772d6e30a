"""CPU oracle for the CSR analytics path -- TEST INFRASTRUCTURE ONLY.

This package is the test oracle: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.  The shipped library
(``graphbench_amd``) never imports, links or calls anything here.

Each wrapper forwards to ``oracle.c`` (a plain-C restatement of the reference
algorithm; every C function names the reference file:line it follows).
Graphs are passed as numpy arrays: ``rp`` int32[V+1], ``col`` int32[E],
``w`` uint32[E] or float64[E].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

UINF = 0xFFFFFFFF


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_f64 = C.c_double


def _declare(L):
    sig = {
        "orc_dijkstra_u32": (C.c_int, [_i32, _p, _p, _p, _i32, _p]),
        "orc_dijkstra_f64": (C.c_int, [_i32, _p, _p, _p, _i32, _p]),
        "orc_delta_stepping_u32": (C.c_int, [_i32, _p, _p, _p, _i32, _u32, _u32, _p, _p]),
        "orc_delta_stepping_f64": (C.c_int, [_i32, _p, _p, _p, _i32, _f64, _p]),
        "orc_bfs_levels": (None, [_i32, _p, _p, _i32, _p]),
        "orc_brandes_bfs": (None, [_i32, _p, _p, _p, _p, _p, _i32, _p]),
        "orc_brandes_weighted_f64": (None, [_i32, _p, _p, _p, _p, _p, _p, _p, _i32, _p]),
        "orc_pagerank_ref": (C.c_int, [_i32, _p, _p, _p, _i32, _p]),
        "orc_pagerank_cfg": (C.c_int, [_i32, _p, _p, _p, _f64, _f64, _i32, _p, _p]),
        "orc_comp_distrib": (None, [_i64, _p, _i64, _i32, _p]),
        "orc_bin_search": (_i64, [_p, _i64, _i64, _i64]),
        "orc_merge_path_split": (None, [_i32, _p, _i64, _i64, _p, _p]),
        "orc_separate_csr": (None, [_i32, _p, _p, _p, _u32, _p, _p, _p, _p, _p, _p]),
        "orc_separate_csr_f64": (None, [_i32, _p, _p, _p, _f64, _p, _p, _p, _p, _p, _p]),
        "orc_remove_duplicate": (_i64, [_i64, _p, _p]),
        "orc_edge_parts": (None, [_i64, _p, _p, _p, _i32, _p, _p, _p]),
        "orc_mlp_loss_grad": (_f64, [_i32, _i32, _i32, _i32] + [_p] * 14),
        "orc_set_threads": (None, [_i32]),
        "orc_get_threads": (_i32, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def dijkstra(rp, col, w, src):
    """Exact SSSP distances (uint32 weights -> uint32, float -> float64)."""
    rp, col = _c(rp, np.int32), _c(col, np.int32)
    V = len(rp) - 1
    if np.asarray(w).dtype.kind == "f":
        w = _c(w, np.float64)
        d = np.empty(V, np.float64)
        lib().orc_dijkstra_f64(V, _ptr(rp), _ptr(col), _ptr(w), int(src), _ptr(d))
        return d
    w = _c(w, np.uint32)
    d = np.empty(V, np.uint32)
    lib().orc_dijkstra_u32(V, _ptr(rp), _ptr(col), _ptr(w), int(src), _ptr(d))
    return d


def delta_stepping(rp, col, w, src, delta, light_max=None, stats=False):
    rp, col = _c(rp, np.int32), _c(col, np.int32)
    V = len(rp) - 1
    if np.asarray(w).dtype.kind == "f":
        w = _c(w, np.float64)
        d = np.empty(V, np.float64)
        lib().orc_delta_stepping_f64(V, _ptr(rp), _ptr(col), _ptr(w), int(src),
                                     float(delta), _ptr(d))
        return d
    w = _c(w, np.uint32)
    d = np.empty(V, np.uint32)
    st = np.zeros(4, np.int64)
    lm = int(delta) if light_max is None else int(light_max)
    lib().orc_delta_stepping_u32(V, _ptr(rp), _ptr(col), _ptr(w), int(src),
                                 int(delta), lm, _ptr(d), _ptr(st))
    return (d, st) if stats else d


def bfs_levels(rp, col, src):
    rp, col = _c(rp, np.int32), _c(col, np.int32)
    V = len(rp) - 1
    lv = np.empty(V, np.int32)
    lib().orc_bfs_levels(V, _ptr(rp), _ptr(col), int(src), _ptr(lv))
    return lv


def brandes_bfs(rp, col, sources, rrp=None, rcol=None):
    rp, col = _c(rp, np.int32), _c(col, np.int32)
    rrp = rp if rrp is None else _c(rrp, np.int32)
    rcol = col if rcol is None else _c(rcol, np.int32)
    V = len(rp) - 1
    s = _c(sources, np.int32)
    bc = np.empty(V, np.float64)
    lib().orc_brandes_bfs(V, _ptr(rp), _ptr(col), _ptr(rrp), _ptr(rcol), _ptr(s),
                          len(s), _ptr(bc))
    return bc


def brandes_weighted(rp, col, w, rrp, rcol, rw, sources):
    rp, col, w = _c(rp, np.int32), _c(col, np.int32), _c(w, np.float64)
    rrp, rcol, rw = _c(rrp, np.int32), _c(rcol, np.int32), _c(rw, np.float64)
    V = len(rp) - 1
    s = _c(sources, np.int32)
    bc = np.empty(V, np.float64)
    lib().orc_brandes_weighted_f64(V, _ptr(rp), _ptr(col), _ptr(w), _ptr(rrp),
                                   _ptr(rcol), _ptr(rw), _ptr(s), len(s), _ptr(bc))
    return bc


def pagerank_ref(irp, icol, recip, max_iter=100000):
    irp, icol, recip = _c(irp, np.int32), _c(icol, np.int32), _c(recip, np.float64)
    V = len(irp) - 1
    x = np.empty(V, np.float64)
    it = lib().orc_pagerank_ref(V, _ptr(irp), _ptr(icol), _ptr(recip), int(max_iter), _ptr(x))
    return x, int(it)


def pagerank_cfg(irp, icol, outdeg, d=0.85, tol=1e-7, max_iter=100):
    irp, icol, od = _c(irp, np.int32), _c(icol, np.int32), _c(outdeg, np.int32)
    V = len(irp) - 1
    x = np.empty(V, np.float64)
    r = C.c_double(0.0)
    it = lib().orc_pagerank_cfg(V, _ptr(irp), _ptr(icol), _ptr(od), float(d), float(tol),
                                int(max_iter), _ptr(x), C.byref(r))
    return x, int(it), float(r.value)


def comp_distrib(deg, E, p):
    deg = _c(deg, np.int64)
    out = np.empty(p, np.int64)
    lib().orc_comp_distrib(len(deg), _ptr(deg), int(E), int(p), _ptr(out))
    return out


def bin_search(distrib, val):
    distrib = _c(distrib, np.int64)
    return int(lib().orc_bin_search(_ptr(distrib), int(val), 0, len(distrib) - 1))


def merge_path_split(rp, items, ntiles):
    rp = _c(rp, np.int32)
    V = len(rp) - 1
    tr = np.empty(ntiles + 1, np.int32)
    te = np.empty(ntiles + 1, np.int32)
    lib().orc_merge_path_split(V, _ptr(rp), int(items), int(ntiles), _ptr(tr), _ptr(te))
    return tr, te


def edge_parts(src, dst, distrib):
    """Per-edge (send part, recv part) and the p x p cross-part counts of
    chunk_files: (sp int32[E], rp int32[E], ub int64[p, p])."""
    src = _c(src, np.int32)
    dst = _c(dst, np.int32)
    distrib = _c(distrib, np.int64)
    p = len(distrib)
    sp = np.empty(len(src), np.int32)
    rp = np.empty(len(src), np.int32)
    ub = np.empty(p * p, np.int64)
    lib().orc_edge_parts(len(src), _ptr(src), _ptr(dst), _ptr(distrib), p, _ptr(sp), _ptr(rp),
                         _ptr(ub))
    return sp, rp, ub.reshape(p, p)


def chunk_files(text, file, p):
    """The serial chunking step of main.c:116-136 on a text graph held in
    `text` (the file <dir><file>), with num_nodes = p: scan_csr_idx SEND and
    RECV (graph/csr.c:133-202), comp_distrib_csr SEND
    (partition/csr_partition.c:5-30), write_recip (133-164) and chunk_files
    (32-125).  Returns {path relative to dir: file text}.  Pure Python over
    the tokens, for small graphs."""
    tok = text.split()

    def scan(kind):  # degrees and header, as scan_csr_idx reads them
        deg, hdr, i, n = None, None, 0, 0
        while i < len(tok):
            t = tok[i]
            i += 1
            if t == "p":
                v_num, e_num, off = int(tok[i + 1]), int(tok[i + 2]), int(tok[i + 3])
                i += 5
                hdr = (v_num, e_num, off)
                deg = [0] * v_num
            elif t == "a":
                u, v = int(tok[i]), int(tok[i + 1])
                i += 3
                deg[u - hdr[2] if kind == "send" else v - hdr[2]] += 1
                n += 1
                if n == hdr[1]:
                    break
        return deg, hdr

    sdeg, (V, E, _) = scan("send")
    rdeg, _ = scan("recv")
    # comp_distrib_csr (SEND): cut on the send degrees at E // p
    distrib, ss, rs, j = [0] * p, [0] * p, [0] * p, 0
    for i in range(V):
        ss[j] += sdeg[i]
        rs[j] += rdeg[i]
        if ss[j] > E // p:
            distrib[j] = i
            j += 1
    distrib[p - 1] = V - 1
    assert distrib == comp_distrib(sdeg, E, p).tolist()
    ext = str(p)
    out = {}
    # write_recip
    rec = [""] * p
    j = 0
    for i in range(V):
        rec[j] += "%f " % (1.0 / sdeg[i] if sdeg[i] else float("inf"))
        if i == distrib[j]:
            j += 1
    tail = "\n" + "".join("%d " % d for d in distrib) + "\n\n"
    for k in range(p):
        out[f"parex/{file}_recip_{k}_{ext}"] = rec[k] + tail
    # chunk_files
    sd, rc = [""] * p, [""] * p
    ub = np.zeros((p, p), np.int64)
    i, n, off, v_num, e_num = 0, 0, 0, 0, 0
    while i < len(tok):
        t = tok[i]
        i += 1
        if t == "p":
            v_num, e_num, off = int(tok[i + 1]), int(tok[i + 2]), int(tok[i + 3])
            i += 4
            for k in range(p):
                size = distrib[0] + 1 if k == 0 else distrib[k] - distrib[k - 1]
                o = (0 if k == 0 else distrib[k - 1] + 1) + 1
                sd[k] += f"p sp {size} {ss[k]} {o} {v_num}\n"
                rc[k] += f"p sp {size} {rs[k]} {o} {v_num}\n"
        elif t == "a":
            u, v, w = int(tok[i]), int(tok[i + 1]), int(tok[i + 2])
            i += 3
            s_, r_, c = edge_parts([u - 1], [v - off], distrib)
            ub += c
            sd[int(s_[0])] += f"a {u} {v} {w}\n"
            rc[int(r_[0])] += f"a {u} {v} {w}\n"
            n += 1
            if n == e_num:
                break
    for k in range(p):
        out[f"datex/{file}_sd_{k}_{ext}"] = sd[k]
        out[f"datex/{file}_rc_{k}_{ext}"] = rc[k]
    out[f"parex/{file}_send_{ext}"] = "".join(
        "".join("%d " % x for x in row) + "\n" for row in ub.tolist())
    return out


def remove_duplicate(pairs):
    """First occurrences of (from, to) pairs, in order: (k, 2) int64."""
    pairs = _c(np.asarray(pairs).reshape(-1, 2), np.int64)
    out = np.empty_like(pairs)
    k = lib().orc_remove_duplicate(len(pairs), _ptr(pairs), _ptr(out))
    return out[:k]


def separate_csr(rp, col, w, light_max):
    if np.asarray(w).dtype.kind == "f":
        return _separate_f64(rp, col, w, light_max)
    rp, col, w = _c(rp, np.int32), _c(col, np.int32), _c(w, np.uint32)
    V = len(rp) - 1
    E = len(col)
    lrp = np.empty(V + 1, np.int32)
    hrp = np.empty(V + 1, np.int32)
    lcol = np.empty(E, np.int32)
    hcol = np.empty(E, np.int32)
    lw = np.empty(E, np.uint32)
    hw = np.empty(E, np.uint32)
    lib().orc_separate_csr(V, _ptr(rp), _ptr(col), _ptr(w), int(light_max), _ptr(lrp),
                           _ptr(lcol), _ptr(lw), _ptr(hrp), _ptr(hcol), _ptr(hw))
    nl, nh = int(lrp[-1]), int(hrp[-1])
    return (lrp, lcol[:nl], lw[:nl]), (hrp, hcol[:nh], hw[:nh])


def _separate_f64(rp, col, w, delta):
    rp, col, w = _c(rp, np.int32), _c(col, np.int32), _c(w, np.float64)
    V = len(rp) - 1
    E = len(col)
    lrp = np.empty(V + 1, np.int32)
    hrp = np.empty(V + 1, np.int32)
    lcol = np.empty(max(E, 1), np.int32)
    hcol = np.empty(max(E, 1), np.int32)
    lw = np.empty(max(E, 1), np.float64)
    hw = np.empty(max(E, 1), np.float64)
    lib().orc_separate_csr_f64(V, _ptr(rp), _ptr(col), _ptr(w), float(delta), _ptr(lrp),
                               _ptr(lcol), _ptr(lw), _ptr(hrp), _ptr(hcol), _ptr(hw))
    nl, nh = int(lrp[-1]), int(hrp[-1])
    return (lrp, lcol[:nl], lw[:nl]), (hrp, hcol[:nh], hw[:nh])


def mlp_loss_grad(X, T, W1, b1, W2, b2, W3, b3):
    X, T = _c(X, np.float32), _c(T, np.float32)
    ps = [_c(a, np.float32) for a in (W1, b1, W2, b2, W3, b3)]
    n, din = X.shape
    h = ps[0].shape[0]
    dout = ps[4].shape[0]
    gs = [np.empty(a.shape, np.float64) for a in ps]
    loss = lib().orc_mlp_loss_grad(n, din, h, dout, _ptr(X), _ptr(T),
                                   *[_ptr(a) for a in ps], *[_ptr(g) for g in gs])
    return float(loss), gs
