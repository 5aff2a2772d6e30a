// gb_csr_build.hip -- CSR construction on the device from an edge list.
//
// The reference builds its CSRs on the host from the edge files: a counting
// pass (scan_csr_idx, graph/csr.c:133-202) and a scatter in file order
// (read_graph_csr, graph/csr.c:4-66; read_csr_bin, graph/csr.c:211-299).  Here
// the edge list is on the device and the CSR is one stable radix sort of the
// row keys (edges of a row keep their input order, as the reference's scatter
// does), a histogram and a prefix sum.  The transpose of a CSR is the same
// construction with the endpoints swapped.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include "graphbench.h"
#include "gb_internal.hpp"

namespace {

__global__ void iota_kernel(int64_t n, int32_t *__restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (int32_t)i;
}

__global__ void count_rows_kernel(int64_t E, const int32_t *__restrict__ key,
                                  int32_t *__restrict__ deg) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&deg[key[e]], 1);
}

template <int WB>
__global__ void gather_cols_kernel(int64_t E, const int32_t *__restrict__ perm,
                                   const int32_t *__restrict__ dst, const void *__restrict__ w,
                                   int32_t *__restrict__ col, void *__restrict__ w_out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = perm[i];
    col[i] = dst[p];
    if (WB == 4) ((uint32_t *)w_out)[i] = ((const uint32_t *)w)[p];
    if (WB == 8) ((double *)w_out)[i] = ((const double *)w)[p];
  }
}

// out-of-range endpoints are reported, not written
__global__ void check_ids_kernel(int64_t E, int32_t V, const int32_t *__restrict__ a,
                                 const int32_t *__restrict__ b, int *__restrict__ bad) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    if (a[e] < 0 || a[e] >= V || b[e] < 0 || b[e] >= V) *bad = 1;
}

// row id of every edge of a CSR
__global__ void expand_rows_kernel(int32_t V, const int32_t *__restrict__ rp,
                                   int32_t *__restrict__ row) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wave; v < V; v += nwaves)
    for (int32_t e = rp[v] + lane; e < rp[v + 1]; e += 64) row[e] = (int32_t)v;
}

int grid_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)gb_num_cus() * 16));
}

// ---- duplicate edges ------------------------------------------------------
__global__ void pair_keys_kernel(int64_t E, const int32_t *__restrict__ src,
                                 const int32_t *__restrict__ dst, uint64_t *__restrict__ key,
                                 int32_t *__restrict__ ids, int *__restrict__ bad) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = src[e], d = dst[e];
    if (s < 0 || d < 0) *bad = 1;
    key[e] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)d;
    ids[e] = (int32_t)e;
  }
}

// keep[e] = 1 for the first index of each run of equal keys (the sort is
// stable, so that is the pair's first occurrence in input order)
__global__ void first_of_run_kernel(int64_t E, const uint64_t *__restrict__ skey,
                                    const int32_t *__restrict__ sid, int32_t *__restrict__ keep) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x)
    keep[sid[i]] = (i == 0 || skey[i] != skey[i - 1]) ? 1 : 0;
}

__global__ void compact_pairs_kernel(int64_t E, const int32_t *__restrict__ keep,
                                     const int32_t *__restrict__ pos,
                                     const int32_t *__restrict__ src,
                                     const int32_t *__restrict__ dst, int32_t *__restrict__ src_out,
                                     int32_t *__restrict__ dst_out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    if (keep[e]) {
      src_out[pos[e]] = src[e];
      dst_out[pos[e]] = dst[e];
    }
}

}  // namespace

// Keeps the first occurrence of every (src, dst) pair, in input order, as
// remove_duplicate (partition/csr_partition_bin_mpi.c:31-106) does with a hash
// set over the file order: a stable radix sort of the 64-bit pair keys with
// their indices marks the first index of each run; a scan compacts.
extern "C" gb_status gb_edges_unique(int64_t E, const int32_t *d_src, const int32_t *d_dst,
                                     int32_t *d_src_out, int32_t *d_dst_out, int64_t *h_kept,
                                     void *stream) {
  if (E < 0 || E >= ((int64_t)1 << 31) || !h_kept ||
      (E > 0 && (!d_src || !d_dst || !d_src_out || !d_dst_out)))
    return gb_fail(GB_ERR_INVALID, "gb_edges_unique: bad arguments");
  *h_kept = 0;
  if (E == 0) return GB_OK;
  hipStream_t st = gb_stream(stream);
  DevBuf key, skey, ids, sids, keep, pos, flag, tmp;
  GB_TRY(key.ensure(sizeof(uint64_t) * E));
  GB_TRY(skey.ensure(sizeof(uint64_t) * E));
  GB_TRY(ids.ensure(sizeof(int32_t) * E));
  GB_TRY(sids.ensure(sizeof(int32_t) * E));
  GB_TRY(keep.ensure(sizeof(int32_t) * E));
  GB_TRY(pos.ensure(sizeof(int32_t) * (E + 1)));
  GB_TRY(flag.ensure(sizeof(int)));
  GB_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(pair_keys_kernel, dim3(grid_for(E)), dim3(256), 0, st, E, d_src, d_dst,
                     key.as<uint64_t>(), ids.as<int32_t>(), flag.as<int>());
  GB_HIP(hipGetLastError());
  int bad = 0;
  GB_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  if (bad) return gb_fail(GB_ERR_INVALID, "gb_edges_unique: negative vertex id");
  size_t tb_sort = 0, tb_scan = 0;
  GB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb_sort, key.as<uint64_t>(),
                                            skey.as<uint64_t>(), ids.as<int32_t>(),
                                            sids.as<int32_t>(), (int)E, 0, 64, st));
  GB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, keep.as<int32_t>(),
                                          pos.as<int32_t>(), (int)E, st));
  GB_TRY(tmp.ensure(std::max(tb_sort, tb_scan) + 16));
  GB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb_sort, key.as<uint64_t>(),
                                            skey.as<uint64_t>(), ids.as<int32_t>(),
                                            sids.as<int32_t>(), (int)E, 0, 64, st));
  hipLaunchKernelGGL(first_of_run_kernel, dim3(grid_for(E)), dim3(256), 0, st, E,
                     skey.as<uint64_t>(), sids.as<int32_t>(), keep.as<int32_t>());
  GB_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb_scan, keep.as<int32_t>(), pos.as<int32_t>(),
                                          (int)E, st));
  hipLaunchKernelGGL(compact_pairs_kernel, dim3(grid_for(E)), dim3(256), 0, st, E,
                     keep.as<int32_t>(), pos.as<int32_t>(), d_src, d_dst, d_src_out, d_dst_out);
  GB_HIP(hipGetLastError());
  int32_t last_pos = 0, last_keep = 0;
  GB_HIP(hipMemcpyAsync(&last_pos, pos.as<int32_t>() + (E - 1), sizeof(int32_t),
                        hipMemcpyDeviceToHost, st));
  GB_HIP(hipMemcpyAsync(&last_keep, keep.as<int32_t>() + (E - 1), sizeof(int32_t),
                        hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));  // scratch buffers are freed on return
  *h_kept = (int64_t)last_pos + last_keep;
  return GB_OK;
}

extern "C" gb_status gb_csr_from_edges(int32_t V, int64_t E, const int32_t *d_src,
                                       const int32_t *d_dst, const void *d_w, int32_t w_bytes,
                                       int32_t *d_row_ptr, int32_t *d_col, void *d_w_out,
                                       void *stream) {
  if (V < 0 || E < 0 || E >= ((int64_t)1 << 31) || !d_row_ptr || (E > 0 && (!d_src || !d_dst || !d_col)))
    return gb_fail(GB_ERR_INVALID, "gb_csr_from_edges: bad arguments");
  if (w_bytes != 0 && w_bytes != 4 && w_bytes != 8)
    return gb_fail(GB_ERR_INVALID, "gb_csr_from_edges: w_bytes must be 0, 4 or 8");
  if (w_bytes && E > 0 && (!d_w || !d_w_out))
    return gb_fail(GB_ERR_INVALID, "gb_csr_from_edges: weights without a weight buffer");
  hipStream_t st = gb_stream(stream);
  DevBuf deg, keys_out, ids, perm, flag, tmp;
  GB_TRY(deg.ensure(sizeof(int32_t) * ((size_t)V + 1)));
  GB_TRY(flag.ensure(sizeof(int)));
  GB_HIP(hipMemsetAsync(deg.p, 0, sizeof(int32_t) * ((size_t)V + 1), st));
  GB_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
  if (E > 0) {
    hipLaunchKernelGGL(check_ids_kernel, dim3(grid_for(E)), dim3(256), 0, st, E, V, d_src, d_dst,
                       flag.as<int>());
    int bad = 0;
    GB_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    if (bad) return gb_fail(GB_ERR_INVALID, "gb_csr_from_edges: endpoint outside [0, V)");
    GB_TRY(keys_out.ensure(sizeof(int32_t) * E));
    GB_TRY(ids.ensure(sizeof(int32_t) * E));
    GB_TRY(perm.ensure(sizeof(int32_t) * E));
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(E)), dim3(256), 0, st, E, ids.as<int32_t>());
    hipLaunchKernelGGL(count_rows_kernel, dim3(grid_for(E)), dim3(256), 0, st, E, d_src,
                       deg.as<int32_t>() + 1);
    GB_HIP(hipGetLastError());
    // stable sort of the edge indices by row (LSD radix sort keeps ties in input order)
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < (int64_t)V) bits++;
    size_t tb_sort = 0, tb_scan = 0;
    const uint32_t *ukeys = (const uint32_t *)d_src;  // checked non-negative above
    GB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb_sort, ukeys, keys_out.as<uint32_t>(),
                                              ids.as<int32_t>(), perm.as<int32_t>(), (int)E, 0,
                                              bits, st));
    GB_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb_scan, deg.as<int32_t>(), d_row_ptr,
                                            V + 1, st));
    GB_TRY(tmp.ensure(std::max(tb_sort, tb_scan) + 16));
    GB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb_sort, ukeys, keys_out.as<uint32_t>(),
                                              ids.as<int32_t>(), perm.as<int32_t>(), (int)E, 0,
                                              bits, st));
    GB_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb_scan, deg.as<int32_t>(), d_row_ptr, V + 1,
                                            st));
    if (w_bytes == 8)
      hipLaunchKernelGGL(gather_cols_kernel<8>, dim3(grid_for(E)), dim3(256), 0, st, E,
                         perm.as<int32_t>(), d_dst, d_w, d_col, d_w_out);
    else if (w_bytes == 4)
      hipLaunchKernelGGL(gather_cols_kernel<4>, dim3(grid_for(E)), dim3(256), 0, st, E,
                         perm.as<int32_t>(), d_dst, d_w, d_col, d_w_out);
    else
      hipLaunchKernelGGL(gather_cols_kernel<0>, dim3(grid_for(E)), dim3(256), 0, st, E,
                         perm.as<int32_t>(), d_dst, d_w, d_col, d_w_out);
    GB_HIP(hipGetLastError());
  } else {
    GB_HIP(hipMemsetAsync(d_row_ptr, 0, sizeof(int32_t) * ((size_t)V + 1), st));
  }
  GB_HIP(hipStreamSynchronize(st));  // scratch buffers are freed on return
  return GB_OK;
}

extern "C" gb_status gb_csr_transpose(int32_t V, int64_t E, const int32_t *d_row_ptr,
                                      const int32_t *d_col, const void *d_w, int32_t w_bytes,
                                      int32_t *d_trow_ptr, int32_t *d_tcol, void *d_tw,
                                      void *stream) {
  if (V < 0 || E < 0 || !d_row_ptr || !d_trow_ptr || (E > 0 && (!d_col || !d_tcol)))
    return gb_fail(GB_ERR_INVALID, "gb_csr_transpose: bad arguments");
  hipStream_t st = gb_stream(stream);
  DevBuf rows;
  GB_TRY(rows.ensure(sizeof(int32_t) * ((size_t)E + 1)));
  if (E > 0 && V > 0) {
    hipLaunchKernelGGL(expand_rows_kernel, dim3(grid_for((int64_t)V * 64)), dim3(256), 0, st, V,
                       d_row_ptr, rows.as<int32_t>());
    GB_HIP(hipGetLastError());
  }
  // transposed rows are keyed by the column; edges of a transposed row keep
  // the order of their source rows
  GB_TRY(gb_csr_from_edges(V, E, d_col, rows.as<int32_t>(), d_w, w_bytes, d_trow_ptr, d_tcol,
                           d_tw, stream));
  return GB_OK;
}
