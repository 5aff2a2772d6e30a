"""gb_pagerank_ref_noalloc (no device allocation) against gb_pagerank_ref.

Both entry points are called through the C-ABI on the same inputs; the
no-allocation one must give bit-identical ranks and the same iteration count,
on two calls in a row (the second on a smaller graph, so stale scratch from
the first would show), and must keep working with the library's allocation
cap at one byte, where every allocating entry point fails.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph(gb, seed, V, E):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    m = src != dst
    h = gb.HostCSR.from_edges(V, src[m], dst[m])
    r = h.transpose()
    od = h.outdeg().astype(np.float64)
    with np.errstate(divide="ignore"):
        recip = np.where(od > 0, 1.0 / od, np.inf)  # the reference's recip (csr.c:194-200)
    return h, r, recip


def _noalloc(gb, torch, r, recip_d, max_iter=100000):
    L = gb.lib()
    f = L.gb_pagerank_ref_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                  C.POINTER(C.c_int32), C.c_void_p]
    irp = torch.from_numpy(r.row_ptr).cuda()
    icol = torch.from_numpy(r.col).cuda()
    x = torch.full((r.V,), -1.0, dtype=torch.float64, device="cuda")
    it = C.c_int32(0)
    st = f(r.V, irp.data_ptr(), icol.data_ptr(), recip_d.data_ptr(), max_iter, x.data_ptr(),
           C.byref(it), gb._stream())
    torch.cuda.synchronize()
    return st, x, int(it.value)


def test_noalloc_matches_pagerank_ref_twice(gpu):
    import torch
    gb = gpu
    for seed, V, E in [(1, 3000, 24000), (2, 700, 4000)]:
        print(f"[noalloc] seed {seed} V={V} E={E}", flush=True)
        h, r, recip = _graph(gb, seed, V, E)
        g = gb.Graph.from_host(h, reverse=r)
        recip_d = torch.from_numpy(recip).cuda()
        x_ref, it_ref = gb.pagerank_ref(g, recip_d)
        st, x, it = _noalloc(gb, torch, r, recip_d)
        assert st == gb.GB_OK, gb.lib().gb_last_error()
        assert it == it_ref > 0
        assert torch.equal(x, x_ref)


def test_noalloc_needs_no_allocation_and_rejects_oversize(gpu):
    import torch
    gb = gpu
    h, r, recip = _graph(gb, 3, 500, 3000)
    g = gb.Graph.from_host(h, reverse=r)
    recip_d = torch.from_numpy(recip).cuda()
    x_ref, it_ref = gb.pagerank_ref(g, recip_d)
    L = gb.lib()
    L.gb_set_mem_cap.argtypes = [C.c_uint64]
    L.gb_set_mem_cap(1)
    try:
        st, x, it = _noalloc(gb, torch, r, recip_d)
        assert st == gb.GB_OK, L.gb_last_error()
        assert it == it_ref and torch.equal(x, x_ref)
        with pytest.raises(gb.GBError):  # the allocating form hits the cap
            gb.pagerank_ref(g, recip_d)
    finally:
        L.gb_set_mem_cap(0)
    # V beyond the scratch: refused before any launch
    f = L.gb_pagerank_ref_noalloc
    dummy = torch.zeros(4, dtype=torch.int64, device="cuda")
    p = dummy.data_ptr()
    it = C.c_int32(0)
    st = f((1 << 22) + 1, p, p, p, 10, p, C.byref(it), gb._stream())
    assert st == 7  # GB_ERR_UNSUPPORTED
    assert b"exceeds the fixed scratch" in L.gb_last_error()
