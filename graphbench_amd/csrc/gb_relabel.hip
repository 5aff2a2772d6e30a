// gb_relabel.hip -- out-degree relabelling of a graph for the PageRank gather.
//
// The pull gather reads contrib[u] for every in-edge u->v.  How often u is
// read is its out-degree, and on power-law graphs a small set of vertices
// takes most reads.  Renumbering vertices by out-degree (descending, ties by
// id) packs those hot contributions into the first cache lines of the array,
// so the L1/L2 working set of the gather shrinks to a dense prefix instead of
// one 128-byte line per hub.  The rows of the renumbered in-CSR are sorted by
// (new) source id, so lanes that walk a row read ascending addresses.
//
// Built once per graph handle and cached; PageRank solves in the new
// numbering and maps x0 in and the ranks out through the permutation, so
// callers see the original ids.  Row sums run over the same in-edges in a
// different order, which changes FP32 rounding only (compensated sums).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>
#include <hipcub/device/device_segmented_radix_sort.hpp>

#include "gb_internal.hpp"

namespace {

__global__ void deg_from_rp_kernel(int32_t V, const int32_t *__restrict__ rp,
                                   int32_t *__restrict__ deg) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) deg[v] = rp[v + 1] - rp[v];
}

__global__ void deg_count_kernel(int64_t E, const int32_t *__restrict__ icol,
                                 int32_t *__restrict__ deg) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) atomicAdd(&deg[icol[e]], 1);
}

// key = ~outdeg so that an ascending (stable) radix sort orders by degree
// descending and then by id
__global__ void sort_keys_kernel(int32_t V, const int32_t *__restrict__ deg,
                                 uint32_t *__restrict__ key, int32_t *__restrict__ id) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) {
    key[v] = ~(uint32_t)deg[v];
    id[v] = (int32_t)v;
  }
}

__global__ void rank_kernel(int32_t V, const int32_t *__restrict__ perm,
                            const int32_t *__restrict__ rrp, const float *__restrict__ inv,
                            int32_t *__restrict__ rank, int32_t *__restrict__ indeg_new,
                            float *__restrict__ inv_new) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const int32_t old = perm[r];
  rank[old] = (int32_t)r;
  indeg_new[r] = rrp[old + 1] - rrp[old];
  inv_new[r] = inv[old];
}

// one wave per new row: copy the old row's sources, renumbered
__global__ void fill_rows_kernel(int32_t V, const int32_t *__restrict__ perm,
                                 const int32_t *__restrict__ rank, const int32_t *__restrict__ rrp,
                                 const int32_t *__restrict__ rcol,
                                 const int32_t *__restrict__ nrp, int32_t *__restrict__ ncol) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = wave; r < V; r += nwaves) {
    const int32_t old = perm[r];
    const int32_t b = rrp[old], n = rrp[old + 1] - b, o = nrp[r];
    for (int32_t k = lane; k < n; k += 64) ncol[o + k] = rank[rcol[b + k]];
  }
}

}  // namespace

gb_status gb_graph_outdeg(gb_graph g, int32_t *d_deg, hipStream_t st) {
  const int32_t V = g->V;
  const int blocks = (int)((V + 255) / 256);
  if (g->rp) {
    hipLaunchKernelGGL(deg_from_rp_kernel, dim3(blocks), dim3(256), 0, st, V, g->rp, d_deg);
  } else {
    GB_HIP(hipMemsetAsync(d_deg, 0, sizeof(int32_t) * V, st));
    if (g->E > 0)
      hipLaunchKernelGGL(deg_count_kernel, dim3((unsigned)((g->E + 255) / 256)), dim3(256), 0, st,
                         g->E, g->rcol, d_deg);
  }
  GB_HIP(hipGetLastError());
  return GB_OK;
}

gb_status gb_graph_pr_relabel(gb_graph g, hipStream_t st) {
  if (g->rl_valid) return GB_OK;
  if (!g->rrp || (!g->rcol && g->E > 0) || !g->inv_outdeg_valid)
    return gb_fail(GB_ERR_INVALID, "relabel: needs the reverse CSR and 1/outdeg");
  const int32_t V = g->V;
  const int64_t E = g->E;
  const int blocks = (int)((V + 255) / 256);
  DevBuf deg, keys, keys_out, ids, indeg, tmp;
  GB_TRY(deg.ensure(sizeof(int32_t) * V));
  GB_TRY(keys.ensure(sizeof(uint32_t) * V));
  GB_TRY(keys_out.ensure(sizeof(uint32_t) * V));
  GB_TRY(ids.ensure(sizeof(int32_t) * V));
  GB_TRY(indeg.ensure(sizeof(int32_t) * V));
  GB_TRY(g->rl_perm.ensure(sizeof(int32_t) * V));
  GB_TRY(g->rl_rank.ensure(sizeof(int32_t) * V));
  GB_TRY(g->rl_rrp.ensure(sizeof(int32_t) * (V + 1)));
  GB_TRY(g->rl_rcol.ensure(sizeof(int32_t) * (E + 1)));
  GB_TRY(g->rl_inv.ensure(sizeof(float) * V));
  int32_t *perm = g->rl_perm.as<int32_t>();
  GB_TRY(gb_graph_outdeg(g, deg.as<int32_t>(), st));
  hipLaunchKernelGGL(sort_keys_kernel, dim3(blocks), dim3(256), 0, st, V, deg.as<int32_t>(),
                     keys.as<uint32_t>(), ids.as<int32_t>());
  GB_HIP(hipGetLastError());
  // stable sort by ~degree: perm[new] = old
  size_t tb = 0;
  GB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint32_t>(),
                                            keys_out.as<uint32_t>(), ids.as<int32_t>(), perm, V,
                                            0, 32, st));
  size_t tb_scan = 0;
  GB_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb_scan, indeg.as<int32_t>(),
                                          g->rl_rrp.as<int32_t>() + 1, V, st));
  size_t tb_seg = 0;
  GB_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(
      nullptr, tb_seg, g->rl_rcol.as<int32_t>(), g->rl_rcol.as<int32_t>(), (int)E, V,
      g->rl_rrp.as<int32_t>(), g->rl_rrp.as<int32_t>() + 1, 0, 32, st));
  GB_TRY(tmp.ensure(std::max(tb, std::max(tb_scan, tb_seg)) + 16));
  GB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.as<uint32_t>(),
                                            keys_out.as<uint32_t>(), ids.as<int32_t>(), perm, V,
                                            0, 32, st));
  hipLaunchKernelGGL(rank_kernel, dim3(blocks), dim3(256), 0, st, V, perm, g->rrp,
                     g->inv_outdeg.as<float>(), g->rl_rank.as<int32_t>(), indeg.as<int32_t>(),
                     g->rl_inv.as<float>());
  GB_HIP(hipGetLastError());
  GB_HIP(hipMemsetAsync(g->rl_rrp.p, 0, sizeof(int32_t), st));
  GB_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb_scan, indeg.as<int32_t>(),
                                          g->rl_rrp.as<int32_t>() + 1, V, st));
  if (E > 0) {
    // renumbered rows into a scratch array, then each row sorted into rl_rcol
    DevBuf unsorted;
    GB_TRY(unsorted.ensure(sizeof(int32_t) * E));
    const int fb = (int)std::min<int64_t>(((int64_t)V + 3) / 4, (int64_t)gb_num_cus() * 64);
    hipLaunchKernelGGL(fill_rows_kernel, dim3(std::max(fb, 1)), dim3(256), 0, st, V, perm,
                       g->rl_rank.as<int32_t>(), g->rrp, g->rcol, g->rl_rrp.as<int32_t>(),
                       unsorted.as<int32_t>());
    GB_HIP(hipGetLastError());
    GB_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(
        tmp.p, tb_seg, unsorted.as<int32_t>(), g->rl_rcol.as<int32_t>(), (int)E, V,
        g->rl_rrp.as<int32_t>(), g->rl_rrp.as<int32_t>() + 1, 0, 32, st));
    GB_HIP(hipStreamSynchronize(st));  // scratch buffers are freed on return
  } else {
    GB_HIP(hipStreamSynchronize(st));
  }
  g->rl_valid = true;
  return GB_OK;
}
