// gb_bc.hip -- weighted Brandes betweenness (f64 weights) for gfx950.
//
// Replaces the weighted form of run_bc (btwcentral/csr_bc.c:16-130): the
// per-source run_sspath with predecessor lists, the distance sort
// (q_sort_double) and the comp_num_spath / comp_delta loops become kernels over
// a batch of sources on the tight-edge DAG of the exact distances.  An edge
// u->x is tight when fl(d[u] + w) == d[x] (the '<=' predecessor rule of
// csr_sspath.c:101-118 with each edge recorded once).  Every (source slot,
// vertex) pair is an entry of one `order` array, grouped by DAG depth, which
// is also Brandes' stack.
//   indeg     count of tight in-edges per (slot, vertex)
//   expand    one wave per depth-d entry, lanes over its out-edges: the last
//             tight in-edge to be processed assigns depth d+1 (Kahn order).
//   sigma     one wave per depth-(d+1) entry, pull over its tight in-edges
//             (FP64; integer values are exact below 2^53).
//   delta     backwards, one wave per depth-d entry, pull over tight out-edges:
//             delta[v] = sum ((1+delta[w])*sigma[v])/sigma[w] (the operand
//             order of csr_bc.c:116), fixed-order wave sum.
//   accum     bc[v] += delta[slot][v] for v != source, slots in order
//             (csr_bc.c:57-61, sums not halved).
// Unweighted BC and BFS levels are the batched kernels of gb_msbc.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {

constexpr int BC_BLOCK = 256;

__global__ void bc_fill_kernel(int64_t n, int32_t *__restrict__ level, double *__restrict__ sigma,
                               double *__restrict__ delta) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    level[i] = -1;
    if (sigma) sigma[i] = 0.0;
    if (delta) delta[i] = 0.0;
  }
}

__global__ void bc_seed_kernel(int32_t V, int32_t nslots, const int32_t *__restrict__ src,
                               int32_t *__restrict__ level, double *__restrict__ sigma,
                               uint64_t *__restrict__ order) {
  int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int64_t base = (int64_t)s * V;
  level[base + src[s]] = 0;
  if (sigma) sigma[base + src[s]] = 1.0;
  order[s] = ((uint64_t)s << 32) | (uint32_t)src[s];
}

// Tight-DAG discovery (weighted): decrement the count of unprocessed tight
// in-edges of each tight successor; the last one assigns depth d+1.
__global__ __launch_bounds__(BC_BLOCK) void bcw_expand_kernel(
    int32_t V, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
    const double *__restrict__ w, const double *__restrict__ dist, int32_t *__restrict__ indeg,
    uint64_t *__restrict__ order, uint64_t lo, uint64_t hi, int32_t d, int32_t *__restrict__ level,
    unsigned long long *n_order) {
  const int lane = gb_lane();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = lo + wave; i < hi; i += nwaves) {
    const uint64_t ent = order[i];
    const uint32_t s = (uint32_t)(ent >> 32);
    const int32_t u = (int32_t)(uint32_t)ent;
    const int64_t base = (int64_t)s * V;
    const double du = dist[base + u];
    const int32_t beg = rp[u], end = rp[u + 1];
    for (int32_t b0 = beg; b0 < end; b0 += GB_WAVE) {
      const int32_t e = b0 + lane;
      bool push = false;
      int32_t x = 0;
      if (e < end) {
        x = col[e];
        if (__dadd_rn(du, w[e]) == dist[base + x] && level[base + x] != 0) {
          if (atomicSub(&indeg[base + x], 1) == 1) {
            level[base + x] = d + 1;
            push = true;
          }
        }
      }
      const uint64_t pb = gb_wave_reserve(n_order, push ? 1u : 0u);
      if (push) order[pb] = ((uint64_t)s << 32) | (uint32_t)x;
    }
  }
}

// Count tight in-edges per (slot, vertex) for the weighted form.
__global__ __launch_bounds__(BC_BLOCK) void bcw_indeg_kernel(
    int32_t V, int32_t nslots, const int32_t *__restrict__ rrp, const int32_t *__restrict__ rcol,
    const double *__restrict__ rw, const double *__restrict__ dist, int32_t *__restrict__ indeg) {
  const int64_t n = (int64_t)nslots * V;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i / V;
    const int32_t v = (int32_t)(i - s * V);
    const double dv = dist[i];
    int32_t c = 0;
    if (dv != __longlong_as_double(0x7FF0000000000000ll)) {
      const double *ds = dist + s * V;
      for (int32_t e = rrp[v]; e < rrp[v + 1]; e++) {
        const double du = ds[rcol[e]];
        if (__dadd_rn(du, rw[e]) == dv) c++;
      }
    }
    indeg[i] = c;
  }
}

// Reached (slot, vertex) pairs: finite distances, one atomic per block.
__global__ void bcw_reached_kernel(int64_t n, const double *__restrict__ dist,
                                   unsigned long long *__restrict__ count) {
  unsigned long long c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c += dist[i] != __longlong_as_double(0x7FF0000000000000ll) ? 1ull : 0ull;
  __shared__ unsigned long long s_c[256 / GB_WAVE];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, GB_WAVE);
  if (gb_lane() == 0) s_c[threadIdx.x / GB_WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < (int)(blockDim.x / GB_WAVE); k++) t += s_c[k];
    atomicAdd(count, t);
  }
}

struct PredTight {
  const double *dist;
  const double *w;   // forward weights
  const double *rw;  // reverse weights
  __device__ __forceinline__ bool in_edge(int64_t base, int32_t u, int32_t v, int32_t e,
                                          int32_t) const {
    return __dadd_rn(dist[base + u], rw[e]) == dist[base + v];
  }
  __device__ __forceinline__ bool out_edge(int64_t base, int32_t v, int32_t x, int32_t e,
                                           int32_t) const {
    return __dadd_rn(dist[base + v], w[e]) == dist[base + x];
  }
};

// sigma pull for the entries [lo, hi) at level d (d >= 1)
template <class P>
__global__ __launch_bounds__(BC_BLOCK) void bc_sigma_kernel(
    int32_t V, const int32_t *__restrict__ rrp, const int32_t *__restrict__ rcol,
    const uint64_t *__restrict__ order, uint64_t lo, uint64_t hi, int32_t d, P pred,
    double *__restrict__ sigma) {
  const int lane = gb_lane();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = lo + wave; i < hi; i += nwaves) {
    const uint64_t ent = order[i];
    const int64_t base = (int64_t)(ent >> 32) * V;
    const int32_t v = (int32_t)(uint32_t)ent;
    double acc = 0.0;
    for (int32_t e = rrp[v] + lane; e < rrp[v + 1]; e += GB_WAVE) {
      const int32_t u = rcol[e];
      if (pred.in_edge(base, u, v, e, d)) acc += sigma[base + u];
    }
    acc = gb_wave_sum_f64(acc);
    if (lane == 0) sigma[base + v] = acc;
  }
}

// delta pull for the entries [lo, hi) at level d
template <class P>
__global__ __launch_bounds__(BC_BLOCK) void bc_delta_kernel(
    int32_t V, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
    const uint64_t *__restrict__ order, uint64_t lo, uint64_t hi, int32_t d, P pred,
    const double *__restrict__ sigma, double *__restrict__ delta) {
  const int lane = gb_lane();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = lo + wave; i < hi; i += nwaves) {
    const uint64_t ent = order[i];
    const int64_t base = (int64_t)(ent >> 32) * V;
    const int32_t v = (int32_t)(uint32_t)ent;
    const double sv = sigma[base + v];
    double acc = 0.0;
    for (int32_t e = rp[v] + lane; e < rp[v + 1]; e += GB_WAVE) {
      const int32_t x = col[e];
      if (pred.out_edge(base, v, x, e, d))
        acc += __ddiv_rn(__dmul_rn(__dadd_rn(1.0, delta[base + x]), sv), sigma[base + x]);
    }
    acc = gb_wave_sum_f64(acc);
    if (lane == 0) delta[base + v] = acc;
  }
}

__global__ void bc_accum_kernel(int32_t V, int32_t nslots, const int32_t *__restrict__ src,
                                const double *__restrict__ delta, double *__restrict__ bc) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  double acc = bc[v];
  for (int32_t s = 0; s < nslots; s++)
    if (src[s] != (int32_t)v) acc += delta[(int64_t)s * V + v];
  bc[v] = acc;
}

struct BcWork {
  DevBuf level, sigma, delta, order, src, cnt, indeg, dist;
  void release() {
    for (DevBuf *b : {&level, &sigma, &delta, &order, &src, &cnt, &indeg, &dist}) b->release();
  }
};
// Heap singleton, never destroyed: device memory must not be freed from a
// static destructor, which can run after the HIP runtime has shut down.
BcWork &g_bcw_ref() {
  static BcWork *w = new BcWork();
  return *w;
}
#define g_bcw (g_bcw_ref())

int wave_grid(uint64_t n) {
  uint64_t blocks = (n + 3) / 4;
  uint64_t cap = (uint64_t)gb_num_cus() * 16;
  if (blocks > cap) blocks = cap;
  return blocks < 1 ? 1 : (int)blocks;
}

int fill_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  int64_t cap = (int64_t)gb_num_cus() * 32;
  if (b > cap) b = cap;
  return b < 1 ? 1 : (int)b;
}

// Forward levels for nslots sources; returns level offsets (size depth+2).
gb_status forward_levels(gb_graph g, int32_t nslots, hipStream_t st,
                         std::vector<uint64_t> &off, bool with_sigma, int32_t *level) {
  BcWork &w = g_bcw;
  const int32_t V = g->V;
  unsigned long long *cnt = w.cnt.as<unsigned long long>();
  uint64_t *order = w.order.as<uint64_t>();
  double *sigma = with_sigma ? w.sigma.as<double>() : nullptr;
  off.assign(1, 0);
  off.push_back((uint64_t)nslots);
  unsigned long long h = nslots;
  GB_HIP(hipMemcpyAsync(cnt, &h, sizeof(h), hipMemcpyHostToDevice, st));
  int32_t d = 0;
  PredTight pt{w.dist.as<double>(), g->w64, g->rw64};
  for (;;) {
    const uint64_t lo = off[d], hi = off[d + 1];
    if (hi == lo) break;
    hipLaunchKernelGGL(bcw_expand_kernel, dim3(wave_grid(hi - lo)), dim3(BC_BLOCK), 0, st, V,
                         g->rp, g->col, g->w64, w.dist.as<double>(), w.indeg.as<int32_t>(), order,
                         lo, hi, d, level, cnt);
    GB_HIP(hipGetLastError());
    GB_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    off.push_back(h);
    d++;
    if (with_sigma && h > hi) {
      const int32_t *rrp = g->rrp ? g->rrp : g->rp;
      const int32_t *rcol = g->rcol ? g->rcol : g->col;
      hipLaunchKernelGGL(bc_sigma_kernel<PredTight>, dim3(wave_grid(h - hi)), dim3(BC_BLOCK), 0,
                         st, V, rrp, rcol, order, hi, (uint64_t)h, d, pt, sigma);
      GB_HIP(hipGetLastError());
    }
  }
  return GB_OK;
}

gb_status bc_run(gb_graph g, const int32_t *h_sources, int32_t nsrc, int32_t batch, double *d_bc,
                 gb_bc_stats *stats, void *stream) {
  if (!g || (!h_sources && nsrc > 0) || nsrc < 0 || !d_bc)
    return gb_fail(GB_ERR_INVALID, "gb_bc: null argument");
  const int32_t V = g->V;
  for (int32_t i = 0; i < nsrc; i++)
    if (h_sources[i] < 0 || h_sources[i] >= V)
      return gb_fail(GB_ERR_INVALID, "gb_bc: source %d out of range", h_sources[i]);
  if (!g->w64 || (g->rrp && !g->rw64))
    return gb_fail(GB_ERR_INVALID, "gb_bc_weighted: graph needs f64 weights (forward and reverse)");
  hipStream_t st = gb_stream(stream);
  if (batch <= 0 || batch > nsrc) batch = nsrc;
  if (batch > 65535) batch = 65535;
  BcWork &w = g_bcw;
  const int64_t n = (int64_t)batch * V;
  GB_TRY(w.level.ensure(sizeof(int32_t) * (n + 1)));
  GB_TRY(w.sigma.ensure(sizeof(double) * (n + 1)));
  GB_TRY(w.delta.ensure(sizeof(double) * (n + 1)));
  GB_TRY(w.order.ensure(sizeof(uint64_t) * (n + 1)));
  GB_TRY(w.src.ensure(sizeof(int32_t) * batch));
  GB_TRY(w.cnt.ensure(2 * sizeof(unsigned long long)));
  GB_TRY(w.indeg.ensure(sizeof(int32_t) * (n + 1)));
  GB_TRY(w.dist.ensure(sizeof(double) * (n + 1)));
  hipEvent_t e0, e1;
  GB_HIP(hipEventCreate(&e0));
  GB_HIP(hipEventCreate(&e1));
  GB_HIP(hipEventRecord(e0, st));
  GB_HIP(hipMemsetAsync(d_bc, 0, sizeof(double) * V, st));
  int32_t max_depth = 0, batches = 0;
  std::vector<uint64_t> off;
  for (int32_t b0 = 0; b0 < nsrc; b0 += batch) {
    const int32_t ns = std::min(batch, nsrc - b0);
    const int64_t nn = (int64_t)ns * V;
    GB_HIP(hipMemcpyAsync(w.src.p, h_sources + b0, sizeof(int32_t) * ns, hipMemcpyHostToDevice, st));
    // exact distances of the batch, then tight in-degree counts; bucket
    // width: the last f64 split of this graph, else the reference's BC
    // constant (run_sspath(gs, pred, source, 0.1), csr_bc.c:43)
    GB_TRY(gb_sssp_f64(g, g->split64_valid ? g->split64_delta : 0.1, h_sources + b0, ns,
                       w.dist.as<double>(), nullptr, stream));
    const int32_t *rrp = g->rrp ? g->rrp : g->rp;
    const int32_t *rcol = g->rcol ? g->rcol : g->col;
    const double *rw = g->rrp ? g->rw64 : g->w64;
    hipLaunchKernelGGL(bcw_indeg_kernel, dim3(fill_grid(nn)), dim3(256), 0, st, V, ns, rrp, rcol,
                       rw, w.dist.as<double>(), w.indeg.as<int32_t>());
    GB_HIP(hipGetLastError());
    hipLaunchKernelGGL(bc_fill_kernel, dim3(fill_grid(nn)), dim3(256), 0, st, nn,
                       w.level.as<int32_t>(), w.sigma.as<double>(), w.delta.as<double>());
    hipLaunchKernelGGL(bc_seed_kernel, dim3((ns + 255) / 256), dim3(256), 0, st, V, ns,
                       w.src.as<int32_t>(), w.level.as<int32_t>(), w.sigma.as<double>(),
                       w.order.as<uint64_t>());
    GB_HIP(hipGetLastError());
    GB_TRY(forward_levels(g, ns, st, off, true, w.level.as<int32_t>()));
    // Every reached pair must have been given a depth.  A cycle of tight edges
    // (a zero weight, or a weight below half an ulp of the distance it is
    // added to) never lets Kahn's count reach zero: its vertices would keep
    // sigma 0 and divide by it below.
    {
      unsigned long long *rc = w.cnt.as<unsigned long long>() + 1;
      GB_HIP(hipMemsetAsync(rc, 0, sizeof(unsigned long long), st));
      hipLaunchKernelGGL(bcw_reached_kernel, dim3(fill_grid(nn)), dim3(256), 0, st, nn,
                         (const double *)w.dist.as<double>(), rc);
      GB_HIP(hipGetLastError());
      unsigned long long reached = 0;
      GB_HIP(hipMemcpyAsync(&reached, rc, sizeof(reached), hipMemcpyDeviceToHost, st));
      GB_HIP(hipStreamSynchronize(st));
      if (reached != off.back())
        return gb_fail(GB_ERR_INVALID,
                       "gb_bc_weighted: %llu reached vertices lie on a cycle of tight edges "
                       "(weights must be positive and not vanish against the distances)",
                       (unsigned long long)(reached - off.back()));
    }
    const int32_t depth = (int32_t)off.size() - 2;  // number of non-empty levels
    max_depth = std::max(max_depth, depth);
    PredTight pt{w.dist.as<double>(), g->w64, g->rw64};
    for (int32_t d = depth - 1; d >= 0; d--) {
      const uint64_t lo = off[d], hi = off[d + 1];
      if (hi == lo) continue;
      hipLaunchKernelGGL(bc_delta_kernel<PredTight>, dim3(wave_grid(hi - lo)), dim3(BC_BLOCK), 0,
                         st, V, g->rp, g->col, w.order.as<uint64_t>(), lo, hi, d, pt,
                         w.sigma.as<double>(), w.delta.as<double>());
      GB_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(bc_accum_kernel, dim3((V + 255) / 256), dim3(256), 0, st, V, ns,
                       w.src.as<int32_t>(), w.delta.as<double>(), d_bc);
    GB_HIP(hipGetLastError());
    batches++;
  }
  GB_HIP(hipEventRecord(e1, st));
  GB_HIP(hipEventSynchronize(e1));
  float ms = 0;
  GB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (stats) {
    stats->max_depth = max_depth;
    stats->batches = batches;
    stats->ms = ms;
  }
  return GB_OK;
}

}  // namespace

void gb_bc_release() { g_bcw.release(); }

extern "C" gb_status gb_bc_weighted(gb_graph g, const int32_t *h_sources, int32_t nsrc,
                                    int32_t batch, double *d_bc, gb_bc_stats *st, void *stream) {
  return bc_run(g, h_sources, nsrc, batch, d_bc, st, stream);
}
