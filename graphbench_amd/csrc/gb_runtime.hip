// gb_runtime.hip -- status/errors, device memory, graph handles and the
// graph-structure kernels (merge-path split, edge-balanced parts, light/heavy
// separation, max weight).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {
thread_local char g_err[1024] = "";
std::mutex g_mem_mu;
uint64_t g_mem_cap = 0;
uint64_t g_mem_used = 0;
int g_device = -1;
int g_cus = 256;
std::unordered_map<void *, size_t> g_sizes;  // live allocations, for the cap
}  // namespace

gb_status gb_fail(gb_status s, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return s;
}

hipStream_t gb_stream(void *s) { return (hipStream_t)s; }
int gb_num_cus() { return g_cus; }

gb_status gb_dmalloc(void **p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    if (g_mem_cap && g_mem_used + bytes > g_mem_cap)
      return gb_fail(GB_ERR_NOMEM, "device allocation of %zu bytes exceeds the cap (%llu used of %llu)",
                     bytes, (unsigned long long)g_mem_used, (unsigned long long)g_mem_cap);
  }
  void *raw = nullptr;
  hipError_t e = hipMalloc(&raw, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return gb_fail(GB_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_mem_used += bytes;
    g_sizes[raw] = bytes;
  }
  *p = raw;
  return GB_OK;
}

void gb_dfree(void *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_sizes.find(p);
    if (it != g_sizes.end()) {
      g_mem_used -= it->second;
      g_sizes.erase(it);
    }
  }
  (void)hipFree(p);
}

extern "C" {

const char *gb_version(void) { return "graphbench-mi355x 0.1 (gfx950)"; }

const char *gb_status_string(gb_status s) {
  switch (s) {
    case GB_OK: return "ok";
    case GB_ERR_INVALID: return "invalid argument";
    case GB_ERR_NOMEM: return "out of device memory";
    case GB_ERR_HIP: return "HIP runtime error";
    case GB_ERR_OVERFLOW: return "uint32 distance overflow";
    case GB_ERR_NOT_CONVERGED: return "iteration cap reached";
    case GB_ERR_NO_DEVICE: return "no HIP device";
    case GB_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown";
}

const char *gb_last_error(void) { return g_err; }

static void split_release();  // the light/heavy split's scratch (below)

void gb_release_workspaces(void) {
  gb_pr_release();
  gb_sssp_release();
  gb_msssp_release();
  gb_bc_release();
  gb_msbc_release();
  gb_mlp_release();
  gb_learned_release();
  split_release();
}

void gb_set_mem_cap(uint64_t bytes) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_mem_cap = bytes;
}

gb_status gb_init(int32_t device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return gb_fail(GB_ERR_NO_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= n) return gb_fail(GB_ERR_INVALID, "device %d out of range", device);
  GB_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  GB_HIP(hipGetDeviceProperties(&prop, device));
  g_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  GB_TRY(gb_noalloc_setup(device));
  g_device = device;
  return GB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// graph handles

extern "C" gb_status gb_graph_wrap(gb_graph *out, int32_t V, int64_t E,
                                   const int32_t *d_row_ptr, const int32_t *d_col,
                                   const uint32_t *d_w, const int32_t *d_rrow_ptr,
                                   const int32_t *d_rcol, const uint32_t *d_rw) {
  // the forward CSR may be absent when only pull kernels (reverse CSR) are used
  const bool fwd = d_row_ptr && (d_col || E == 0);
  const bool rev = d_rrow_ptr && (d_rcol || E == 0);
  if (!out || V < 0 || E < 0 || E >= (1ll << 31) || (V > 0 && !fwd && !rev) ||
      (d_row_ptr && !d_col && E > 0) || (d_rrow_ptr && !d_rcol && E > 0))
    return gb_fail(GB_ERR_INVALID, "gb_graph_wrap: bad graph (V=%d E=%lld)", V, (long long)E);
  if (g_device < 0) GB_TRY(gb_init(0));
  gb_graph g = new gb_graph_s();
  g->V = V;
  g->E = E;
  g->rp = d_row_ptr;
  g->col = d_col;
  g->w = d_w;
  g->rrp = d_rrow_ptr;
  g->rcol = d_rcol;
  g->rw = d_rw;
  *out = g;
  return GB_OK;
}

uint64_t gb_next_graph_uid() {
  static std::atomic<uint64_t> next{1};
  return next.fetch_add(1);
}

extern "C" gb_status gb_graph_upload(gb_graph *out, int32_t V, int64_t E,
                                     const int32_t *row_ptr, const int32_t *col,
                                     const uint32_t *w, const int32_t *rrow_ptr,
                                     const int32_t *rcol, const uint32_t *rw) {
  if (!out || V < 0 || E < 0 || E >= (1ll << 31) || !row_ptr || (!col && E > 0))
    return gb_fail(GB_ERR_INVALID, "gb_graph_upload: bad graph");
  if (row_ptr[0] != 0 || row_ptr[V] != E)
    return gb_fail(GB_ERR_INVALID, "gb_graph_upload: row_ptr[0]=%d row_ptr[V]=%d, E=%lld",
                   row_ptr[0], row_ptr[V], (long long)E);
  if (g_device < 0) GB_TRY(gb_init(0));
  gb_graph g = new gb_graph_s();
  g->V = V;
  g->E = E;
  g->owns = true;
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> gb_status {
    if (!src) return GB_OK;
    gb_status s = b.ensure(bytes);
    if (s != GB_OK) return s;
    if (bytes) GB_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return GB_OK;
  };
  gb_status s = GB_OK;
  if ((s = up(g->own_rp, row_ptr, sizeof(int32_t) * (V + 1))) != GB_OK ||
      (s = up(g->own_col, col, sizeof(int32_t) * E)) != GB_OK ||
      (s = up(g->own_w, w, sizeof(uint32_t) * E)) != GB_OK ||
      (s = up(g->own_rrp, rrow_ptr, sizeof(int32_t) * (V + 1))) != GB_OK ||
      (s = up(g->own_rcol, rcol, sizeof(int32_t) * E)) != GB_OK ||
      (s = up(g->own_rw, rw, sizeof(uint32_t) * E)) != GB_OK) {
    delete g;
    return s;
  }
  g->rp = g->own_rp.as<int32_t>();
  g->col = g->own_col.as<int32_t>();
  g->w = w ? g->own_w.as<uint32_t>() : nullptr;
  g->rrp = rrow_ptr ? g->own_rrp.as<int32_t>() : nullptr;
  g->rcol = rcol ? g->own_rcol.as<int32_t>() : nullptr;
  g->rw = rw ? g->own_rw.as<uint32_t>() : nullptr;
  *out = g;
  return GB_OK;
}

extern "C" gb_status gb_graph_set_f64_weights(gb_graph g, const double *d_w64,
                                              const double *d_rw64) {
  if (!g || (!d_w64 && g->E > 0)) return gb_fail(GB_ERR_INVALID, "gb_graph_set_f64_weights");
  g->w64 = d_w64;
  g->rw64 = d_rw64;
  g->split64_valid = false;
  return GB_OK;
}

extern "C" gb_status gb_graph_destroy(gb_graph g) {
  if (!g) return GB_OK;
  delete g;
  return GB_OK;
}

// ---------------------------------------------------------------------------
// merge-path split: one thread per tile boundary.
// Row end i (= row_ptr[i+1]) is consumed before edge j iff row_ptr[i+1] <= j.

__global__ void merge_path_split_kernel(int32_t V, const int32_t *__restrict__ rp,
                                        int64_t items, int64_t ntiles,
                                        int32_t *__restrict__ tile_row,
                                        int32_t *__restrict__ tile_edge) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const int64_t E = rp[V];
  int64_t d = t * items;
  const int64_t total = (int64_t)V + E;
  if (d > total) d = total;
  int64_t lo = d - E > 0 ? d - E : 0;
  int64_t hi = d < V ? d : V;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)rp[mid + 1] <= d - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  tile_row[t] = (int32_t)lo;
  tile_edge[t] = (int32_t)(d - lo);
}

extern "C" gb_status gb_merge_path_split(int32_t V, const int32_t *d_row_ptr, int64_t items,
                                         int64_t ntiles, int32_t *d_tile_row,
                                         int32_t *d_tile_edge, void *stream) {
  if (V < 0 || items <= 0 || ntiles < 0 || !d_row_ptr || !d_tile_row || !d_tile_edge)
    return gb_fail(GB_ERR_INVALID, "gb_merge_path_split: bad arguments");
  int64_t n = ntiles + 1;
  int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(merge_path_split_kernel, dim3(blocks), dim3(256), 0, gb_stream(stream), V,
                     d_row_ptr, items, ntiles, d_tile_row, d_tile_edge);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

// comp_distrib_csr rule (partition/csr_partition.c:5-30): part j grows from
// the vertex after the previous cut; it is cut at the first vertex i where the
// part's degree sum exceeds q = floor(E/p).  With prefix sums this is one
// binary search per part; the cut chain is sequential, so one thread walks it.
// Vertices after the last cut form part ncuts; later parts stay empty, and
// distrib[p-1] is V-1 (csr_partition.c:29).
__global__ void edge_balanced_parts_kernel(int32_t V, const int32_t *__restrict__ rp,
                                           const int32_t *__restrict__ srp,
                                           const int32_t *__restrict__ rrp, int32_t p,
                                           int64_t *__restrict__ distrib,
                                           int64_t *__restrict__ ssize,
                                           int64_t *__restrict__ rsize) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int64_t E = rp[V];
  const int64_t q = E / p;
  for (int32_t k = 0; k < p; k++) {
    distrib[k] = 0;
    if (ssize) ssize[k] = 0;
    if (rsize) rsize[k] = 0;
  }
  int64_t start = 0;  // first vertex of the current part
  int32_t j = 0;
  while (j < p && start < V) {
    // smallest i >= start with rp[i+1] - rp[start] > q
    const int64_t base = rp[start];
    int64_t lo = start, hi = V;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if ((int64_t)rp[mid + 1] - base > q)
        hi = mid;
      else
        lo = mid + 1;
    }
    const int64_t end = lo < V ? lo : V - 1;  // last vertex of part j
    if (ssize) ssize[j] = (int64_t)srp[end + 1] - srp[start];
    if (rsize) rsize[j] = (int64_t)rrp[end + 1] - rrp[start];
    if (lo >= V) break;  // no cut: part j runs to the last vertex
    distrib[j] = lo;
    start = lo + 1;
    j++;
  }
  distrib[p - 1] = (int64_t)V - 1;
}

gb_status gb_edge_balanced_parts_ex(int32_t V, const int32_t *d_cut_rp, const int32_t *d_send_rp,
                                    const int32_t *d_recv_rp, int32_t p, int64_t *d_distrib,
                                    int64_t *d_send_size, int64_t *d_recv_size, hipStream_t st) {
  if (V <= 0 || p <= 0 || !d_cut_rp || !d_distrib)
    return gb_fail(GB_ERR_INVALID, "gb_edge_balanced_parts: bad arguments");
  hipLaunchKernelGGL(edge_balanced_parts_kernel, dim3(1), dim3(64), 0, st, V, d_cut_rp,
                     d_send_rp ? d_send_rp : d_cut_rp, d_recv_rp ? d_recv_rp : d_cut_rp, p,
                     d_distrib, d_send_size, d_recv_size);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

extern "C" gb_status gb_edge_balanced_parts(int32_t V, const int32_t *d_row_ptr, int32_t p,
                                            int64_t *d_distrib, void *stream) {
  return gb_edge_balanced_parts_ex(V, d_row_ptr, nullptr, nullptr, p, d_distrib, nullptr, nullptr,
                                   gb_stream(stream));
}

// bin_search (utils/utils.c:222-239), same branch structure, iterative.
__device__ __host__ inline int64_t gb_bin_search(const int64_t *distrib, int64_t val,
                                                 int64_t start, int64_t end) {
  for (;;) {
    const int64_t mid = start + (end - start) / 2;
    if (start == end) return start;
    if (start == end - 1) return val <= distrib[start] ? start : end;
    if (val <= distrib[mid] && val > distrib[mid - 1]) return mid;
    if (val > distrib[mid]) {
      start = mid;
      continue;
    }
    if (val <= distrib[mid - 1]) {
      end = mid;
      continue;
    }
    return -1;
  }
}

__global__ void node_ids_kernel(int64_t E, const int32_t *__restrict__ col,
                                const int64_t *__restrict__ distrib, int32_t p,
                                int64_t *__restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = gb_bin_search(distrib, col[e], 0, p - 1);
}

gb_status gb_node_ids(int64_t E, const int32_t *d_col, const int64_t *d_distrib, int32_t p,
                      int64_t *d_out, hipStream_t st) {
  if (E == 0) return GB_OK;
  int64_t blocks = (E + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(node_ids_kernel, dim3((unsigned)blocks), dim3(256), 0, st, E, d_col,
                     d_distrib, p, d_out);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

// ---------------------------------------------------------------------------
// per-wave statistics folded into counters (one block, fixed order)

__global__ __launch_bounds__(256) void wstat_fold_kernel(int64_t nwaves,
                                                         const unsigned long long *__restrict__ s,
                                                         unsigned long long *__restrict__ cnt,
                                                         int i0, int i1, int i2, int i3) {
  __shared__ unsigned long long s_sum[GB_WSTAT][4];
  unsigned long long a[GB_WSTAT] = {};
  for (int64_t w = threadIdx.x; w < nwaves; w += blockDim.x)
    for (int k = 0; k < GB_WSTAT; k++) a[k] += s[w * GB_WSTAT + k];
  for (int k = 0; k < GB_WSTAT; k++) {
    unsigned long long v = a[k];
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (gb_lane() == 0) s_sum[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < GB_WSTAT) {
    const int k = threadIdx.x;
    const int idx = k == 0 ? i0 : k == 1 ? i1 : k == 2 ? i2 : i3;
    const unsigned long long t = s_sum[k][0] + s_sum[k][1] + s_sum[k][2] + s_sum[k][3];
    if (idx >= 0) cnt[idx] += t;
  }
}

gb_status gb_wstat_fold(int64_t nwaves, const unsigned long long *slots, unsigned long long *cnt,
                        int idx0, int idx1, int idx2, hipStream_t st, int idx3) {
  hipLaunchKernelGGL(wstat_fold_kernel, dim3(1), dim3(256), 0, st, nwaves, slots, cnt, idx0, idx1,
                     idx2, idx3);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

// ---------------------------------------------------------------------------
// edge owners and cross-part counts (the per-edge work of chunk_files,
// partition/csr_partition.c:99-108)

constexpr int EP_LDS_CELLS = 8192;  // p * p counters held in LDS per block at most

__global__ void fill_u64_kernel(unsigned long long *a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0ull;
}

// One thread per edge (grid-stride).  Parts come from the reference's
// bin_search, which stays in [0, p-1] for any distrib.  With
// p * p <= EP_LDS_CELLS the block counts in LDS and adds its counts to the
// global ones once.
__global__ __launch_bounds__(256) void edge_parts_kernel(
    int64_t E, const int32_t *__restrict__ src, const int32_t *__restrict__ dst,
    const int64_t *__restrict__ distrib, int32_t p, int32_t *__restrict__ send_part,
    int32_t *__restrict__ recv_part, unsigned long long *__restrict__ ub) {
  __shared__ unsigned int s_cnt[EP_LDS_CELLS];
  const int64_t cells = (int64_t)p * p;
  const bool in_lds = ub && cells <= EP_LDS_CELLS;
  if (in_lds) {
    for (int64_t k = threadIdx.x; k < cells; k += blockDim.x) s_cnt[k] = 0u;
    __syncthreads();
  }
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = gb_bin_search(distrib, src[e], 0, p - 1);
    const int64_t r = gb_bin_search(distrib, dst[e], 0, p - 1);
    if (send_part) send_part[e] = (int32_t)s;
    if (recv_part) recv_part[e] = (int32_t)r;
    if (ub && s != r) {
      if (in_lds)
        atomicAdd(&s_cnt[s * p + r], 1u);
      else
        atomicAdd(&ub[s * p + r], 1ull);
    }
  }
  if (in_lds) {
    __syncthreads();
    for (int64_t k = threadIdx.x; k < cells; k += blockDim.x)
      if (s_cnt[k]) atomicAdd(&ub[k], (unsigned long long)s_cnt[k]);
  }
}

extern "C" gb_status gb_edge_parts(int64_t E, const int32_t *d_src, const int32_t *d_dst,
                                   const int64_t *d_distrib, int32_t p, int32_t *d_send_part,
                                   int32_t *d_recv_part, unsigned long long *d_send_ub,
                                   void *stream) {
  if (E < 0 || p < 1 || p > GB_MAX_PARTS)
    return gb_fail(GB_ERR_INVALID, "gb_edge_parts: need E >= 0 and 1 <= p <= %d", GB_MAX_PARTS);
  if (E > 0 && (!d_src || !d_dst || !d_distrib))
    return gb_fail(GB_ERR_INVALID, "gb_edge_parts: null edge or distrib array");
  hipStream_t st = gb_stream(stream);
  const int64_t cells = (int64_t)p * p;
  if (d_send_ub) {
    const int64_t fb = std::min<int64_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)fb), dim3(256), 0, st, d_send_ub, cells);
    GB_HIP(hipGetLastError());
  }
  if (E == 0) return GB_OK;
  // enough blocks to fill the chip; with LDS counts, few enough that each
  // block's flush of p * p counters stays small against its edges
  int64_t blocks = (E + 255) / 256;
  const int64_t cap = cells <= EP_LDS_CELLS ? (int64_t)gb_num_cus() * 8 : 65536;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(edge_parts_kernel, dim3((unsigned)blocks), dim3(256), 0, st, E, d_src, d_dst,
                     d_distrib, p, d_send_part, d_recv_part, d_send_ub);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

// ---------------------------------------------------------------------------
// max weight

// max over a[0..n): wave, then block reduction, one atomicMax per block
// (blockDim 256)
__global__ void max_u32_kernel(const uint32_t *__restrict__ a, int64_t n, unsigned int *out) {
  __shared__ uint32_t s_max[4];
  uint32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = a[i] > m ? a[i] : m;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    uint32_t o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if (gb_lane() == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
    if (m) atomicMax(out, m);
  }
}

// Scratch of the light/heavy split and of the cached graph statistics below,
// grown on demand and reused: a batch of graphs would otherwise allocate and
// free device buffers per graph.  Every user synchronises before returning.
struct SplitWork {
  DevBuf flags, lrank, tmp;
};
static SplitWork &g_splitw() {  // heap singleton, never destroyed
  static SplitWork *w = new SplitWork();
  return *w;
}
static void split_release() {
  g_splitw().flags.release();
  g_splitw().lrank.release();
  g_splitw().tmp.release();
}

gb_status gb_graph_max_outdeg(gb_graph g, int32_t *out, hipStream_t st) {
  if (g->max_outdeg >= 0) {
    *out = g->max_outdeg;
    return GB_OK;
  }
  if (g->E == 0 || g->V == 0) {
    *out = g->max_outdeg = 0;
    return GB_OK;
  }
  DevBuf &deg = g_splitw().flags, &tmp = g_splitw().tmp;  // scratch, reused (synced below)
  GB_TRY(deg.ensure(sizeof(int32_t) * g->V));
  GB_TRY(tmp.ensure(sizeof(unsigned int)));
  GB_TRY(gb_graph_outdeg(g, deg.as<int32_t>(), st));
  GB_HIP(hipMemsetAsync(tmp.p, 0, sizeof(unsigned int), st));
  int blocks = (int)std::min<int64_t>((g->V + 255) / 256, (int64_t)gb_num_cus() * 4);
  hipLaunchKernelGGL(max_u32_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st,
                     (const uint32_t *)deg.as<int32_t>(), (int64_t)g->V, tmp.as<unsigned int>());
  GB_HIP(hipGetLastError());
  unsigned int h = 0;
  GB_HIP(hipMemcpyAsync(&h, tmp.p, sizeof(h), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  g->max_outdeg = (int32_t)h;
  *out = g->max_outdeg;
  return GB_OK;
}

gb_status gb_graph_max_w(gb_graph g, uint32_t *out, hipStream_t st) {
  if (g->max_w_valid) {
    *out = g->max_w;
    return GB_OK;
  }
  if (g->E == 0) {
    *out = 0;
    return GB_OK;
  }
  if (!g->w) return gb_fail(GB_ERR_INVALID, "graph has no uint32 weights");
  DevBuf &tmp = g_splitw().tmp;  // scratch, reused (synced below)
  GB_TRY(tmp.ensure(sizeof(unsigned int)));
  GB_HIP(hipMemsetAsync(tmp.p, 0, sizeof(unsigned int), st));
  int blocks = (int)std::min<int64_t>((g->E + 255) / 256, (int64_t)gb_num_cus() * 4);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(max_u32_kernel, dim3(blocks), dim3(256), 0, st, g->w, g->E,
                     tmp.as<unsigned int>());
  GB_HIP(hipGetLastError());
  unsigned int h = 0;
  GB_HIP(hipMemcpyAsync(&h, tmp.p, sizeof(h), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  g->max_w = h;
  g->max_w_valid = true;
  *out = h;
  return GB_OK;
}

// ---------------------------------------------------------------------------
// light/heavy separation (separate_csr, graph/csr.c:75-130) as a stable
// in-row partition: light edges first, each part in row order.  It is
// edge-parallel, so a hub row is spread over the whole grid:
//   * lrank[i + 1] = light(w[i]), and an inclusive scan makes lrank[i] the
//     number of light edges before edge i;
//   * edge i of row v goes to rp[v] + (lrank[i] - lrank[rp[v]]) when light,
//     and after the row's lrank[rp[v+1]] - lrank[rp[v]] light edges, in order,
//     when heavy;
//   * le[v] = rp[v] + the row's light count.

template <class W, class Pred>
__global__ void light_flags_kernel(int64_t E, const W *__restrict__ w, Pred light,
                                   int32_t *__restrict__ f) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x)
    f[i] = light(w[i]) ? 1 : 0;
}

__global__ void light_end_kernel(int32_t V, const int32_t *__restrict__ rp,
                                 const int32_t *__restrict__ lrank, int32_t *__restrict__ le) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x)
    le[v] = rp[v] + (lrank[rp[v + 1]] - lrank[rp[v]]);
}

constexpr int SPLIT_IPT = 8;  // edges per lane and wave tile

// Last row r in [lo, hi] with rp[r] <= i, given rp[lo] <= i (rp ascending):
// the whole wave searches together, 64 probes per step (lane j reads
// rp[lo + (j + 1) * span / 64]), so a range of V rows takes about log64 V
// dependent reads.  lo, hi and i are the same in every lane.
__device__ __forceinline__ int32_t wave_row_of_edge(const int32_t *__restrict__ rp, int32_t lo,
                                                    int32_t hi, int64_t i) {
  const int lane = gb_lane();
  while (lo < hi) {
    const int64_t span = (int64_t)hi - lo;
    const int32_t p = lo + (int32_t)((span * (lane + 1)) / GB_WAVE);  // p_63 == hi
    const bool ok = rp[p] <= i;  // true on a prefix of the lanes
    const int c = __popcll(__ballot(ok));
    const int32_t pl = __shfl(p, c > 0 ? c - 1 : 0, GB_WAVE);
    const int32_t ph = __shfl(p, c < GB_WAVE ? c : GB_WAVE - 1, GB_WAVE);
    const int32_t nlo = c > 0 ? pl : lo;
    hi = c < GB_WAVE ? ph - 1 : hi;
    lo = nlo;
  }
  return lo;
}

// A wave takes tiles of 64 * SPLIT_IPT consecutive edges (lane-consecutive
// within each of SPLIT_IPT steps).  The tile's first and last rows come from
// two wave-wide searches; each lane then finds the rows of its SPLIT_IPT edges
// with fixed-step binary searches run side by side (branch-free, on clamped
// edge indices, so each step's reads are one batch), reads the edges and the
// rows' bounds in two batches, and writes every edge to its slot: light edges
// first in each row, both parts in row order.
template <class W, class Pred>
__global__ void split_scatter_kernel(int32_t V, int64_t E, const int32_t *__restrict__ rp,
                                     const int32_t *__restrict__ col, const W *__restrict__ w,
                                     Pred light, const int32_t *__restrict__ lrank,
                                     int32_t *__restrict__ col_out, W *__restrict__ w_out) {
  constexpr int64_t TILE = (int64_t)GB_WAVE * SPLIT_IPT;
  const int lane = gb_lane();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t t0 = wave * TILE; t0 < E; t0 += nwaves * TILE) {
    const int64_t t1 = t0 + TILE < E ? t0 + TILE : E;
    const int32_t r_lo = wave_row_of_edge(rp, 0, V - 1, t0);
    const int32_t r_hi = wave_row_of_edge(rp, r_lo, V - 1, t1 - 1);
    int64_t ie[SPLIT_IPT];
    int32_t lo[SPLIT_IPT], hi[SPLIT_IPT];
#pragma unroll
    for (int k = 0; k < SPLIT_IPT; k++) {
      const int64_t i = t0 + (int64_t)k * GB_WAVE + lane;
      ie[k] = i < t1 ? i : t1 - 1;
      lo[k] = r_lo;
      hi[k] = r_hi;
    }
    // steps: enough for a range of r_hi - r_lo + 1 rows
    const uint32_t span = (uint32_t)(r_hi - r_lo);
    const int steps = span ? 33 - __clz(span) : 0;
    for (int s = 0; s < steps; s++) {
      int32_t x[SPLIT_IPT], mid[SPLIT_IPT];
#pragma unroll
      for (int k = 0; k < SPLIT_IPT; k++) {
        mid[k] = (int32_t)(((int64_t)lo[k] + hi[k] + 1) >> 1);  // lo when lo == hi
        x[k] = rp[mid[k]];
      }
#pragma unroll
      for (int k = 0; k < SPLIT_IPT; k++) {
        const bool le = x[k] <= ie[k];
        lo[k] = le ? mid[k] : lo[k];
        hi[k] = le ? hi[k] : mid[k] - 1;
      }
    }
    int32_t rs[SPLIT_IPT], re[SPLIT_IPT], lr[SPLIT_IPT], ci[SPLIT_IPT];
    W wi[SPLIT_IPT];
#pragma unroll
    for (int k = 0; k < SPLIT_IPT; k++) {
      rs[k] = rp[lo[k]];
      re[k] = rp[lo[k] + 1];
      lr[k] = lrank[ie[k]];
      wi[k] = w[ie[k]];
      ci[k] = col[ie[k]];
    }
    int32_t lrs[SPLIT_IPT], lre[SPLIT_IPT];
#pragma unroll
    for (int k = 0; k < SPLIT_IPT; k++) {
      lrs[k] = lrank[rs[k]];
      lre[k] = lrank[re[k]];
    }
#pragma unroll
    for (int k = 0; k < SPLIT_IPT; k++) {
      const int64_t i = t0 + (int64_t)k * GB_WAVE + lane;
      if (i < t1) {
        const int32_t nl = lre[k] - lrs[k], l = lr[k] - lrs[k];
        const int64_t dst = light(wi[k]) ? rs[k] + l : rs[k] + nl + (i - rs[k]) - l;
        col_out[dst] = ci[k];
        w_out[dst] = wi[k];
      }
    }
  }
}

struct LightU32 {
  uint32_t t;
  __device__ bool operator()(uint32_t x) const { return x <= t; }
};
struct LightF64 {
  double t;
  __device__ bool operator()(double x) const { return x <= t; }
};

static int split_grid(int64_t n, int64_t per_thread) {
  const int64_t blocks = (n + 256 * per_thread - 1) / (256 * per_thread);
  const int64_t cap = (int64_t)gb_num_cus() * 16;
  return (int)std::max<int64_t>(1, std::min(blocks, cap));
}

// light/heavy split of a CSR into col_out / w_out / le (see above).  The
// scratch is shared by all splits: with sync it returns after the work has
// run; without, the work stays queued on st, and a later split may queue
// behind it on the same stream only (a scratch buffer that must grow waits
// for the stream before it is freed).
template <class W, class Pred>
static gb_status split_rows(int32_t V, int64_t E, const int32_t *rp, const int32_t *col,
                            const W *w, Pred light, int32_t *col_out, W *w_out, int32_t *le,
                            hipStream_t st, bool sync) {
  if (V <= 0) return GB_OK;
  DevBuf &flags = g_splitw().flags, &lrank = g_splitw().lrank, &tmp = g_splitw().tmp;
  auto reserve = [st](DevBuf &b, size_t n) -> gb_status {
    if (n > b.bytes) GB_HIP(hipStreamSynchronize(st));
    return b.ensure(n);
  };
  GB_TRY(reserve(lrank, sizeof(int32_t) * ((size_t)E + 1)));
  GB_HIP(hipMemsetAsync(lrank.p, 0, sizeof(int32_t), st));
  if (E > 0) {
    GB_TRY(reserve(flags, sizeof(int32_t) * (size_t)E));
    hipLaunchKernelGGL((light_flags_kernel<W, Pred>), dim3(split_grid(E, 4)), dim3(256), 0, st, E,
                       w, light, flags.as<int32_t>());
    GB_HIP(hipGetLastError());
    size_t tb = 0;
    GB_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, flags.as<int32_t>(),
                                            lrank.as<int32_t>() + 1, (int)E, st));
    GB_TRY(reserve(tmp, tb + 16));
    GB_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, flags.as<int32_t>(),
                                            lrank.as<int32_t>() + 1, (int)E, st));
    hipLaunchKernelGGL((split_scatter_kernel<W, Pred>), dim3(split_grid(E, 4 * SPLIT_IPT)),
                       dim3(256), 0, st, V, E, rp, col, w, light, lrank.as<int32_t>(), col_out,
                       w_out);
    GB_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(light_end_kernel, dim3(split_grid(V, 4)), dim3(256), 0, st, V, rp,
                     lrank.as<int32_t>(), le);
  GB_HIP(hipGetLastError());
  if (sync) GB_HIP(hipStreamSynchronize(st));
  return GB_OK;
}

gb_status gb_graph_split(gb_graph g, uint32_t light_max, hipStream_t st, bool sync) {
  if (g->split_valid && g->split_light_max == light_max) return GB_OK;
  if (!g->w && g->E > 0) return gb_fail(GB_ERR_INVALID, "graph has no uint32 weights");
  GB_TRY(g->split_col.ensure(sizeof(int32_t) * (g->E + 1)));
  GB_TRY(g->split_w.ensure(sizeof(uint32_t) * (g->E + 1)));
  GB_TRY(g->split_le.ensure(sizeof(int32_t) * (g->V + 1)));
  GB_TRY(split_rows<uint32_t>(g->V, g->E, g->rp, g->col, g->w, LightU32{light_max},
                              g->split_col.as<int32_t>(), g->split_w.as<uint32_t>(),
                              g->split_le.as<int32_t>(), st, sync));
  g->split_valid = true;
  g->split_light_max = light_max;
  return GB_OK;
}

gb_status gb_graph_split64(gb_graph g, double light_max, hipStream_t st) {
  if (g->split64_valid && g->split64_delta == light_max) return GB_OK;
  if (!g->w64 && g->E > 0) return gb_fail(GB_ERR_INVALID, "graph has no f64 weights");
  GB_TRY(g->split64_col.ensure(sizeof(int32_t) * (g->E + 1)));
  GB_TRY(g->split64_w.ensure(sizeof(double) * (g->E + 1)));
  GB_TRY(g->split64_le.ensure(sizeof(int32_t) * (g->V + 1)));
  GB_TRY(split_rows<double>(g->V, g->E, g->rp, g->col, g->w64, LightF64{light_max},
                            g->split64_col.as<int32_t>(), g->split64_w.as<double>(),
                            g->split64_le.as<int32_t>(), st, true));
  g->split64_valid = true;
  g->split64_delta = light_max;
  return GB_OK;
}

extern "C" gb_status gb_separate(gb_graph g, uint32_t light_max, int32_t *d_col_out,
                                 uint32_t *d_w_out, int32_t *d_light_end, void *stream) {
  if (!g || !g->w || !d_col_out || !d_w_out || !d_light_end)
    return gb_fail(GB_ERR_INVALID, "gb_separate: bad arguments");
  hipStream_t st = gb_stream(stream);
  return split_rows<uint32_t>(g->V, g->E, g->rp, g->col, g->w, LightU32{light_max}, d_col_out,
                              d_w_out, d_light_end, st, true);
}
