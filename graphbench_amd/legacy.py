"""ctypes mirror of the reference's t_csr / p_list interface (include/gb_legacy.h).

Lets Python callers (and the tests) drive the drop-in entry points
run_sspath_arr, run_sspath, run_pagerank_csr_mpi, run_bc and the readers
exactly as C code linked against the reference would.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _lib

lint = C.c_longlong
sint = C.c_int
RECV, SEND = 0, 1


class TVi(C.Structure):  # graph.h:95-102
    _fields_ = [("weight", C.c_double), ("wb", C.c_double), ("recip", C.c_double),
                ("vet_deg", lint)]


class TE(C.Structure):  # graph.h:105-110
    _fields_ = [("node_id", lint), ("edge_weight", C.c_double)]


class TCsr(C.Structure):  # graph.h:113-125
    _fields_ = [("vet_idx", C.POINTER(lint)), ("edge_idx", C.POINTER(lint)),
                ("vet_info", C.POINTER(TVi)), ("edge_info", C.POINTER(TE)),
                ("v_size", lint), ("e_size", lint), ("offset", lint), ("csr_type", sint),
                ("global_v_size", lint)]


class TList(C.Structure):  # list.h:3-9
    pass


TList._fields_ = [("int_val", lint), ("int_val1", lint), ("double_val", C.c_double),
                  ("next", C.POINTER(TList))]

_declared = False


def lib():
    global _declared
    L = _lib()
    if not _declared:
        P = C.POINTER
        for name, args in {
            "run_sspath_arr": [P(TCsr), P(P(TList)), lint, C.c_double],
            "run_sspath": [P(TCsr), P(P(TList)), lint, C.c_double],
            "run_pagerank_csr_mpi": [C.c_char_p, C.c_char_p],
            "run_bc": [C.c_char_p, C.c_char_p],
            "scan_csr_idx": [P(TCsr), C.c_char_p, C.c_char_p, sint],
            "read_graph_csr": [P(TCsr), C.c_char_p, C.c_char_p, sint],
            "read_csr_bin": [P(TCsr), C.c_char_p, C.c_char_p, sint, C.c_int],
            "separate_csr": [P(TCsr), P(TCsr), P(TCsr), C.c_double],
            "comp_distrib_csr": [P(TCsr), P(TCsr), P(lint), P(lint), P(lint), sint],
            "init_node_id": [P(TCsr), P(lint), sint],
            "remove_duplicate": [C.c_char_p, C.c_char_p, sint],
            "write_recip": [C.c_char_p, C.c_char_p, P(lint), P(TCsr)],
            "chunk_files": [C.c_char_p, C.c_char_p, P(lint), P(lint), P(lint)],
        }.items():
            f = getattr(L, name)
            f.restype = None
            f.argtypes = args
        L.bin_search.restype = lint
        L.bin_search.argtypes = [P(lint), lint, lint, lint]
        _declared = True
    return L


class LegacyGraph:
    """A t_csr built in memory from a CSR (rp int[V+1], col, f64 weights)."""

    def __init__(self, rp, col, w):
        rp = np.asarray(rp, np.int64)
        self.V = len(rp) - 1
        self.E = int(rp[-1])
        self._vet_idx = np.ascontiguousarray(rp[1:], dtype=np.int64)
        self._edge_idx = np.ascontiguousarray(col, dtype=np.int64)
        self._vet_info = (TVi * max(self.V, 1))()
        self._edge_info = (TE * max(self.E, 1))()
        deg = np.diff(rp)
        for i in range(self.V):
            self._vet_info[i].vet_deg = int(deg[i])
        for e in range(self.E):
            self._edge_info[e].edge_weight = float(w[e])
        self.csr = TCsr()
        self.csr.vet_idx = self._vet_idx.ctypes.data_as(C.POINTER(lint))
        self.csr.edge_idx = self._edge_idx.ctypes.data_as(C.POINTER(lint))
        self.csr.vet_info = C.cast(self._vet_info, C.POINTER(TVi))
        self.csr.edge_info = C.cast(self._edge_info, C.POINTER(TE))
        self.csr.v_size = self.V
        self.csr.e_size = self.E
        self.csr.offset = 1
        self.csr.csr_type = SEND
        self.csr.global_v_size = self.V

    def weights_out(self) -> np.ndarray:
        return np.array([self._vet_info[i].weight for i in range(self.V)], np.float64)


def new_pred_lists(V):
    """V list heads initialised like creat_list (list.c:4-9)."""
    heads = (TList * V)()
    for h in heads:
        h.int_val, h.int_val1, h.double_val = -1, -1, -1.0
        h.next = None
    arr = (C.POINTER(TList) * V)(*[C.pointer(h) for h in heads])
    return heads, arr


def list_values(head: TList):
    out = []
    p = head.next
    while p:
        out.append(p.contents.int_val)
        p = p.contents.next
    return out


def csr_arrays(g: TCsr):
    """t_csr -> (row_ptr int64[V+1], col int64[E], weights f64[E])."""
    V, E = g.v_size, g.e_size
    rp = np.zeros(V + 1, np.int64)
    rp[1:] = np.ctypeslib.as_array(g.vet_idx, shape=(V,))
    col = np.ctypeslib.as_array(g.edge_idx, shape=(max(E, 1),))[:E].copy()
    w = np.array([g.edge_info[e].edge_weight for e in range(E)], np.float64)
    return rp, col, w
