"""Size-independent checks of SSSP distances, run on the GPU with torch.

For uint32 distances d from source s on a CSR with non-negative weights, the
exact shortest-path distances are the unique vector with
  * d[s] = 0;
  * no edge can still relax: d[v] <= d[u] + w(u, v) for every reached u;
  * every reached v != s has a tight in-edge: d[u] + w(u, v) == d[v];
  * an unreached v has no reached in-neighbour.
(The first three give d <= true distances by induction along tight edges and
d >= true distances by the no-relax condition along any path.)
"""
from __future__ import annotations

import numpy as np

UINF = 0xFFFFFFFF


class DeviceCSR:
    """int64 copies of a host CSR on the GPU, with the source of every edge."""

    def __init__(self, torch, h):
        self.V = h.V
        rp = torch.from_numpy(h.row_ptr.astype(np.int64)).cuda()
        self.src = torch.repeat_interleave(torch.arange(h.V, device="cuda"), rp[1:] - rp[:-1])
        self.col = torch.from_numpy(h.col.astype(np.int64)).cuda()
        self.w = torch.from_numpy(h.w.astype(np.int64)).cuda()


def check_distances(torch, g: DeviceCSR, d_row, source: int, tag=""):
    """Asserts the four properties for one row of distances (int32 storage of
    uint32, on the GPU)."""
    d = d_row.to(torch.int64) & 0xFFFFFFFF
    assert int(d[source]) == 0, f"{tag}: d[source] != 0"
    du = d[g.src]
    dv = d[g.col]
    reached_u = du != UINF
    assert bool(((dv <= du + g.w) | ~reached_u).all()), f"{tag}: an edge can still relax"
    assert bool((~reached_u | (dv != UINF)).all()), f"{tag}: unreached target of a reached source"
    tight = reached_u & (du + g.w == dv)
    has_tight = torch.zeros(g.V, dtype=torch.bool, device="cuda")
    has_tight[g.col[tight]] = True
    reached = d != UINF
    reached[source] = False
    assert bool((has_tight | ~reached).all()), f"{tag}: a reached vertex has no tight in-edge"
