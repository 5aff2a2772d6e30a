"""graphbench_amd -- MI355X-native CSR analytics (SSSP, PageRank, Brandes BC).

Python host side over the C-ABI of ``libgraphbench.so`` (declared in
``include/graphbench.h`` and ``include/gb_legacy.h``).  Every compute call goes
to hand-written HIP kernels for gfx950; there is no CPU fallback: if the
library is missing, ``lib()`` raises.  PyTorch is used only to hold device
memory and to name the current stream.

Reference counterparts (stplaydog/GRAPHBENCH):
  sssp()          run_sspath_arr  sssp/csr_sspath_array.c:149-363
  pagerank()      run_pagerank_csr_mpi  pagerank/pagerank_mpi.c:667-1044
  pagerank_ref()  the same loop with the reference constants (FP64)
  bfs(), bc()     run_bc  btwcentral/csr_bc.c:16-130
  merge_path_split(), edge_balanced_parts()
                  comp_distrib_csr  partition/csr_partition.c:5-30
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraphbench.so")
_lib = None

UINF = 0xFFFFFFFF

GB_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "out of device memory", 3: "HIP runtime error",
          4: "uint32 distance overflow", 5: "iteration cap reached", 6: "no HIP device",
          7: "unsupported"}


class GBError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_u64 = C.c_uint64
_f64 = C.c_double


class _HostCSR(C.Structure):
    _fields_ = [("V", C.c_int32), ("E", C.c_int64), ("row_ptr", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("w", C.POINTER(C.c_uint32))]


class _SsspParams(C.Structure):
    _fields_ = [("delta", C.c_uint32), ("light_max", C.c_uint32), ("max_phases", C.c_int32),
                ("reserved", C.c_int32)]


class _SsspStats(C.Structure):
    _fields_ = [("buckets", C.c_int64), ("phases", C.c_int64), ("relax_edges", C.c_int64),
                ("reached", C.c_int64), ("traversed", C.c_int64), ("ms", C.c_double),
                ("improved_near", C.c_int64), ("improved_far", C.c_int64),
                ("expanded_edges", C.c_int64)]


class _PrParams(C.Structure):
    _fields_ = [("damping", C.c_double), ("tol", C.c_double), ("max_iter", C.c_int32),
                ("items_per_thread", C.c_int32), ("use_graph", C.c_int32),
                ("time_kernels", C.c_int32), ("d_x0", C.c_void_p), ("relabel", C.c_int32),
                ("row_block", C.c_int32), ("binned", C.c_int32)]


class _PrResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("residual", C.c_double),
                ("dangling", C.c_double), ("gather_ms", C.c_double),
                ("gather_launches", C.c_int32), ("tiles", C.c_int32),
                ("items_per_thread", C.c_int32), ("hot", C.c_int32), ("relabeled", C.c_int32),
                ("row_block", C.c_int32), ("binned", C.c_int32)]


class _BcStats(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("batches", C.c_int32), ("ms", C.c_double)]


_SIGS = {
    "gb_status_string": (C.c_char_p, [_i32]),
    "gb_last_error": (C.c_char_p, []),
    "gb_init": (_i32, [_i32]),
    "gb_set_mem_cap": (None, [_u64]),
    "gb_release_workspaces": (None, []),
    "gb_version": (C.c_char_p, []),
    "gb_graph_wrap": (_i32, [C.POINTER(_p), _i32, _i64, _p, _p, _p, _p, _p, _p]),
    "gb_graph_upload": (_i32, [C.POINTER(_p), _i32, _i64, _p, _p, _p, _p, _p, _p]),
    "gb_graph_set_f64_weights": (_i32, [_p, _p, _p]),
    "gb_graph_destroy": (_i32, [_p]),
    "gb_merge_path_split": (_i32, [_i32, _p, _i64, _i64, _p, _p, _p]),
    "gb_edge_balanced_parts": (_i32, [_i32, _p, _i32, _p, _p]),
    "gb_separate": (_i32, [_p, _u32, _p, _p, _p, _p]),
    "gb_sssp": (_i32, [_p, C.POINTER(_SsspParams), _p, _i32, _p, C.POINTER(_SsspStats), _p]),
    "gb_sssp_f64": (_i32, [_p, _f64, _p, _i32, _p, C.POINTER(_SsspStats), _p]),
    "gb_sssp_multi": (_i32, [_i32, _p, _p, _p, _p, _p, C.POINTER(_SsspStats), _p]),
    "gb_pagerank": (_i32, [_p, C.POINTER(_PrParams), _p, C.POINTER(_PrResult), _p]),
    "gb_pagerank_ref": (_i32, [_p, _p, _i32, _p, C.POINTER(_i32), _p]),
    "gb_bfs": (_i32, [_p, _p, _i32, _p, _p]),
    "gb_bc": (_i32, [_p, _p, _i32, _i32, _p, C.POINTER(_BcStats), _p]),
    "gb_bc_weighted": (_i32, [_p, _p, _i32, _i32, _p, C.POINTER(_BcStats), _p]),
    "gb_gen_rmat": (_i32, [C.POINTER(_HostCSR), _i32, _i64, _f64, _f64, _f64, _f64, _u64, _i32,
                           _u32, _u32, _i32]),
    "gb_gen_er": (_i32, [C.POINTER(_HostCSR), _i64, _f64, _u64, _i32, _u32, _u32]),
    "gb_transpose": (_i32, [C.POINTER(_HostCSR), C.POINTER(_HostCSR)]),
    "gb_pick_sources": (_i32, [C.POINTER(_HostCSR), _i32, _u64, _p]),
    "gb_host_csr_free": (None, [C.POINTER(_HostCSR)]),
    "gb_splitmix64": (_u64, [_u64]),
    "gb_mlp_param_count": (_i32, [_i32, _i32]),
    "gb_mlp_forward": (_i32, [_p, _i32, _i32, _p, _i64, _p, _p]),
    "gb_mlp_forward_bf16": (_i32, [_p, _i32, _i32, _p, _i64, _p, _p]),
    "gb_mlp_forward_mfma": (_i32, [_p, _i32, _i32, _p, _i64, _p, _p]),
    "gb_mlp_loss_grad": (_i32, [_p, _i32, _i32, _p, _p, _i64, _p, C.POINTER(_f64), _p]),
    "gb_mlp_adam": (_i32, [_p, _p, _p, _p, _i32, C.c_float, C.c_float, C.c_float, C.c_float,
                           _i32, _p]),
    "gb_csr_from_edges": (_i32, [_i32, _i64, _p, _p, _p, _i32, _p, _p, _p, _p]),
    "gb_csr_transpose": (_i32, [_i32, _i64, _p, _p, _p, _i32, _p, _p, _p, _p]),
    "gb_edges_unique": (_i32, [_i64, _p, _p, _p, _p, C.POINTER(_i64), _p]),
    "gb_edge_parts": (_i32, [_i64, _p, _p, _p, _i32, _p, _p, _p, _p]),
    "gb_graph_features": (_i32, [_p, _i32, _p, _p]),
    "gb_graph_features_multi": (_i32, [_i32, _p, _i32, _p, _p]),
    "gb_pr_warm_start": (_i32, [_p, _p, _p, _p, _p]),
}


def lib():
    """The loaded libgraphbench.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m graphbench_amd.build`")
        # torch ships its own libamdhip64 (same soname).  Loading torch first
        # makes the library bind to that runtime, so the process holds one HIP
        # runtime and torch's device pointers and streams are valid here.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != GB_OK:
        raise GBError(status, lib().gb_last_error().decode(errors="replace"))


# --------------------------------------------------------------------------
# host-side CSR (numpy)

@dataclass
class HostCSR:
    row_ptr: np.ndarray  # int32[V+1]
    col: np.ndarray      # int32[E]
    w: Optional[np.ndarray] = None  # uint32[E]

    @property
    def V(self) -> int:
        return len(self.row_ptr) - 1

    @property
    def E(self) -> int:
        return len(self.col)

    def outdeg(self) -> np.ndarray:
        return np.diff(self.row_ptr).astype(np.int32)

    def _c(self) -> _HostCSR:
        s = _HostCSR()
        s.V = self.V
        s.E = self.E
        s.row_ptr = self.row_ptr.ctypes.data_as(C.POINTER(C.c_int32))
        s.col = self.col.ctypes.data_as(C.POINTER(C.c_int32))
        s.w = self.w.ctypes.data_as(C.POINTER(C.c_uint32)) if self.w is not None else None
        return s

    def transpose(self) -> "HostCSR":
        out = _HostCSR()
        _check(lib().gb_transpose(C.byref(self._c()), C.byref(out)))
        return _take(out)

    @staticmethod
    def from_edges(V: int, src, dst, w=None) -> "HostCSR":
        """CSR with rows sorted by (dst, w), the generators' canonical order."""
        src = np.asarray(src, np.int64)
        dst = np.asarray(dst, np.int64)
        if w is None:
            order = np.lexsort((dst, src))
        else:
            w = np.asarray(w, np.uint32)
            order = np.lexsort((w, dst, src))
        rp = np.zeros(V + 1, np.int64)
        np.add.at(rp, src + 1, 1)
        rp = np.cumsum(rp).astype(np.int32)
        return HostCSR(rp, dst[order].astype(np.int32),
                       None if w is None else w[order].astype(np.uint32))


def _take(s: _HostCSR) -> HostCSR:
    V, E = s.V, s.E
    rp = np.ctypeslib.as_array(s.row_ptr, shape=(V + 1,)).copy()
    col = np.ctypeslib.as_array(s.col, shape=(max(E, 1),))[:E].copy() if E else np.zeros(0, np.int32)
    w = None
    if s.w:
        w = np.ctypeslib.as_array(s.w, shape=(max(E, 1),))[:E].copy() if E else np.zeros(0, np.uint32)
    lib().gb_host_csr_free(C.byref(s))
    return HostCSR(rp, col, w)


def gen_rmat(scale: int, edge_factor: float = 16, a=0.57, b=0.19, c=0.19, seed: int = 42,
             weights: str = "uniform", wlo: int = 1, whi: int = 255, symmetric=False,
             nverts: int = 0) -> HostCSR:
    """Directed R-MAT (self-loops dropped, duplicates kept), rows sorted by (dst, w).

    weights: 'none', 'uniform' ([wlo, whi]) or 'mix' (50/50 uniform / log-uniform [1,65535]).
    """
    mode = {"none": 0, "uniform": 1, "mix": 2}[weights]
    out = _HostCSR()
    _check(lib().gb_gen_rmat(C.byref(out), scale, nverts, edge_factor, a, b, c, seed, mode, wlo, whi,
                             1 if symmetric else 0))
    return _take(out)


def gen_er(n: int, avg_degree: float, seed: int = 1, weights: str = "uniform", wlo: int = 1,
           whi: int = 255) -> HostCSR:
    mode = {"none": 0, "uniform": 1, "mix": 2}[weights]
    out = _HostCSR()
    _check(lib().gb_gen_er(C.byref(out), n, avg_degree, seed, mode, wlo, whi))
    return _take(out)


def pick_sources(g: HostCSR, k: int, seed: int) -> np.ndarray:
    out = np.empty(k, np.int32)
    _check(lib().gb_pick_sources(C.byref(g._c()), k, seed, out.ctypes.data))
    return out


def splitmix64(x: int) -> int:
    return int(lib().gb_splitmix64(x))


# --------------------------------------------------------------------------
# device side

def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def init(device: int = 0) -> None:
    _check(lib().gb_init(device))


def to_device(a: np.ndarray, device="cuda"):
    torch = _torch()
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


class Graph:
    """Device graph handle over torch tensors (int32 CSR, uint32 weights as int32 storage)."""

    def __init__(self, V: int, E: int, row_ptr, col, w=None, rrow_ptr=None, rcol=None, rw=None,
                 w64=None, rw64=None):
        self.V, self.E = V, E
        self.row_ptr, self.col, self.w = row_ptr, col, w
        self.rrow_ptr, self.rcol, self.rw = rrow_ptr, rcol, rw
        self.w64, self.rw64 = w64, rw64
        h = C.c_void_p()
        _check(lib().gb_graph_wrap(C.byref(h), V, E, _ptr(row_ptr), _ptr(col), _ptr(w),
                                   _ptr(rrow_ptr), _ptr(rcol), _ptr(rw)))
        self.handle = h
        if w64 is not None:
            _check(lib().gb_graph_set_f64_weights(h, _ptr(w64), _ptr(rw64)))

    @staticmethod
    def from_host(g: HostCSR, reverse: Optional[HostCSR] = None, need_reverse: bool = True,
                  w64: Optional[np.ndarray] = None, rw64: Optional[np.ndarray] = None,
                  device="cuda") -> "Graph":
        if need_reverse and reverse is None:
            reverse = g.transpose()
        d = lambda a: None if a is None else to_device(a, device)
        return Graph(g.V, g.E, d(g.row_ptr), d(g.col), d(g.w),
                     d(reverse.row_ptr) if reverse is not None else None,
                     d(reverse.col) if reverse is not None else None,
                     d(reverse.w) if reverse is not None else None,
                     d(w64), d(rw64))

    def close(self):
        if self.handle:
            lib().gb_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sssp(g: Graph, sources: Sequence[int], delta: int = 0, light_max: int = 0,
         out=None, stats: bool = False, max_phases: int = 0):
    """Delta-stepping distances, uint32 (returned as an int32-storage tensor [nsrc, V])."""
    torch = _torch()
    src = np.ascontiguousarray(np.asarray(sources, np.int32))
    if out is None:
        out = torch.empty((len(src), g.V), dtype=torch.int32, device="cuda")
    p = _SsspParams(delta, light_max, max_phases, 0)
    st = _SsspStats()
    _check(lib().gb_sssp(g.handle, C.byref(p), src.ctypes.data, len(src), _ptr(out), C.byref(st),
                         _stream()))
    if stats:
        return out, {k: getattr(st, k) for k, _ in _SsspStats._fields_}
    return out


def sssp_f64(g: Graph, sources: Sequence[int], delta: float, stats: bool = False):
    torch = _torch()
    src = np.ascontiguousarray(np.asarray(sources, np.int32))
    out = torch.empty((len(src), g.V), dtype=torch.float64, device="cuda")
    st = _SsspStats()
    _check(lib().gb_sssp_f64(g.handle, float(delta), src.ctypes.data, len(src), _ptr(out),
                             C.byref(st), _stream()))
    if stats:
        return out, {k: getattr(st, k) for k, _ in _SsspStats._fields_}
    return out


def sssp_multi(graphs: Sequence[Graph], sources: Sequence[int], deltas=None, light_max=None,
               outs=None, stats: bool = False):
    torch = _torch()
    n = len(graphs)
    hs = (C.c_void_p * n)(*[g.handle.value for g in graphs])
    src = np.ascontiguousarray(np.asarray(sources, np.int32))
    dl = None if deltas is None else np.ascontiguousarray(np.asarray(deltas, np.uint32))
    lm = None if light_max is None else np.ascontiguousarray(np.asarray(light_max, np.uint32))
    if outs is None:
        outs = [torch.empty(g.V, dtype=torch.int32, device="cuda") for g in graphs]
    ptrs = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
    st = _SsspStats()
    _check(lib().gb_sssp_multi(n, hs, src.ctypes.data, None if dl is None else dl.ctypes.data,
                               None if lm is None else lm.ctypes.data, ptrs, C.byref(st),
                               _stream()))
    if stats:
        return outs, {k: getattr(st, k) for k, _ in _SsspStats._fields_}
    return outs


def pagerank(g: Graph, damping: float = 0.85, tol: float = 1e-7, max_iter: int = 100,
             items_per_thread: int = 0, use_graph: bool = True, time_kernels: bool = False,
             x0=None, out=None, relabel: int = 0, row_block: int = 0, binned: int = 0):
    """Configured PageRank (FP32 ranks, compensated or fixed-point row sums,
    FP64 residual).

    relabel: 0 auto (on for E >= 2^20), 1 on, -1 off -- gather in out-degree
    order (built once per graph; results are in the caller's vertex ids).
    row_block: 0 autotuned form, 1024/2048/4096 row blocks of that many
    edges, -1 merge path only.  binned: 0 left to the autotuner, -1 never,
    1..6 the binned (propagation-blocking) form with that (hub, chunk, bin) shape."""
    torch = _torch()
    if out is None:
        out = torch.empty(g.V, dtype=torch.float32, device="cuda")
    p = _PrParams(damping, tol, max_iter, items_per_thread, 1 if use_graph else 0,
                  1 if time_kernels else 0, x0.data_ptr() if x0 is not None else None,
                  relabel, row_block, binned)
    r = _PrResult()
    _check(lib().gb_pagerank(g.handle, C.byref(p), _ptr(out), C.byref(r), _stream()))
    return out, {k: getattr(r, k) for k, _ in _PrResult._fields_}


def pagerank_ref(g: Graph, recip, max_iter: int = 100000):
    """Reference-semantics PageRank (damping 0.7 + teleport 0.3, FP64, delta form)."""
    torch = _torch()
    out = torch.empty(g.V, dtype=torch.float64, device="cuda")
    it = C.c_int32(0)
    st = lib().gb_pagerank_ref(g.handle, _ptr(recip), max_iter, _ptr(out), C.byref(it), _stream())
    if st not in (GB_OK, 5):
        _check(st)
    return out, int(it.value)


def bfs(g: Graph, sources: Sequence[int]):
    torch = _torch()
    src = np.ascontiguousarray(np.asarray(sources, np.int32))
    out = torch.empty((len(src), g.V), dtype=torch.int32, device="cuda")
    _check(lib().gb_bfs(g.handle, src.ctypes.data, len(src), _ptr(out), _stream()))
    return out


def bc(g: Graph, sources: Sequence[int], batch: int = 0, weighted: bool = False,
       stats: bool = False):
    torch = _torch()
    src = np.ascontiguousarray(np.asarray(sources, np.int32))
    out = torch.empty(g.V, dtype=torch.float64, device="cuda")
    st = _BcStats()
    fn = lib().gb_bc_weighted if weighted else lib().gb_bc
    _check(fn(g.handle, src.ctypes.data, len(src), batch, _ptr(out), C.byref(st), _stream()))
    if stats:
        return out, {k: getattr(st, k) for k, _ in _BcStats._fields_}
    return out


def merge_path_split(row_ptr, items: int, ntiles: int):
    torch = _torch()
    tr = torch.empty(ntiles + 1, dtype=torch.int32, device="cuda")
    te = torch.empty(ntiles + 1, dtype=torch.int32, device="cuda")
    V = row_ptr.numel() - 1
    _check(lib().gb_merge_path_split(V, _ptr(row_ptr), items, ntiles, _ptr(tr), _ptr(te), _stream()))
    return tr, te


def edge_balanced_parts(row_ptr, p: int):
    torch = _torch()
    out = torch.empty(p, dtype=torch.int64, device="cuda")
    V = row_ptr.numel() - 1
    _check(lib().gb_edge_balanced_parts(V, _ptr(row_ptr), p, _ptr(out), _stream()))
    return out


def separate(g: Graph, light_max: int):
    torch = _torch()
    col = torch.empty(max(g.E, 1), dtype=torch.int32, device="cuda")
    w = torch.empty(max(g.E, 1), dtype=torch.int32, device="cuda")
    le = torch.empty(max(g.V, 1), dtype=torch.int32, device="cuda")
    _check(lib().gb_separate(g.handle, light_max, _ptr(col), _ptr(w), _ptr(le), _stream()))
    return col[:g.E], w[:g.E], le[:g.V]


def csr_from_edges(V: int, src, dst, w=None):
    """Device CSR (row_ptr, col, w) of device edge arrays; rows in source order,
    a row's edges in input order.  w: None, uint32 (int32 storage) or float64."""
    torch = _torch()
    E = int(src.numel())
    rp = torch.empty(V + 1, dtype=torch.int32, device="cuda")
    col = torch.empty(max(E, 1), dtype=torch.int32, device="cuda")
    wb = 0 if w is None else w.element_size()
    wo = None if w is None else torch.empty(max(E, 1), dtype=w.dtype, device="cuda")
    _check(lib().gb_csr_from_edges(V, E, _ptr(src), _ptr(dst), _ptr(w), wb, _ptr(rp), _ptr(col),
                                   _ptr(wo), _stream()))
    return rp, col[:E], (None if wo is None else wo[:E])


def csr_transpose(row_ptr, col, w=None):
    """Device reverse CSR of a device CSR (rows keep the order of their sources)."""
    torch = _torch()
    V = row_ptr.numel() - 1
    E = int(col.numel())
    rp = torch.empty(V + 1, dtype=torch.int32, device="cuda")
    rc = torch.empty(max(E, 1), dtype=torch.int32, device="cuda")
    wb = 0 if w is None else w.element_size()
    wo = None if w is None else torch.empty(max(E, 1), dtype=w.dtype, device="cuda")
    _check(lib().gb_csr_transpose(V, E, _ptr(row_ptr), _ptr(col), _ptr(w), wb, _ptr(rp), _ptr(rc),
                                  _ptr(wo), _stream()))
    return rp, rc[:E], (None if wo is None else wo[:E])


def edges_unique(src, dst):
    """First occurrences of the device edge pairs (src[e], dst[e]), in input
    order (the dedup of remove_duplicate): (src_out, dst_out) device tensors."""
    torch = _torch()
    E = int(src.numel())
    so = torch.empty(max(E, 1), dtype=torch.int32, device="cuda")
    do = torch.empty(max(E, 1), dtype=torch.int32, device="cuda")
    k = _i64(0)
    _check(lib().gb_edges_unique(E, _ptr(src), _ptr(dst), _ptr(so), _ptr(do), C.byref(k),
                                 _stream()))
    return so[:k.value], do[:k.value]


def edge_parts(src, dst, distrib):
    """Owners of both endpoints of every device edge under the 1D partition
    distrib (each part's last vertex; part = the reference's bin_search), and
    the p x p cross-part edge counts, as chunk_files computes them:
    (send_part, recv_part, send_ub) device tensors (int32, int32, int64)."""
    torch = _torch()
    E = int(src.numel())
    distrib = torch.as_tensor(distrib, dtype=torch.int64).to("cuda")
    p = int(distrib.numel())
    sp = torch.empty(max(E, 1), dtype=torch.int32, device="cuda")
    rp = torch.empty(max(E, 1), dtype=torch.int32, device="cuda")
    ub = torch.empty(max(p * p, 1), dtype=torch.int64, device="cuda")
    _check(lib().gb_edge_parts(E, _ptr(src), _ptr(dst), _ptr(distrib), p, _ptr(sp), _ptr(rp),
                               _ptr(ub), _stream()))
    return sp[:E], rp[:E], ub[:p * p].view(p, p)


def u32(t) -> np.ndarray:
    """Device int32-storage tensor -> numpy uint32 view."""
    return t.cpu().numpy().view(np.uint32)
