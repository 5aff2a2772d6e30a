// gb_learned.hip -- inputs and outputs of the learned delta and warm start.
//
// gb_graph_features: the delta model's inputs, reduced on the device
//   [log2 V, log2 avg degree, log2 max weight, log2 mean weight,
//    log2 (max out-degree / avg degree), log2 sources per batch]
// gb_pr_warm_start: per-vertex features [log2(1+indeg), log2(1+outdeg),
//   dangling], standardised, through the MLP (predicting the standardised
//   log2(N * rank)), back to ranks, normalised to sum 1 -- the x0 of
//   gb_pagerank.  The plain start (1/N) stays the default.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {

// Standardised features of a vertex with these degrees (one MLP input row).
struct WsNorm {
  float m0, s0, m1, s1, m2, s2;
};
__device__ __forceinline__ void ws_features(int32_t indeg, int32_t outdeg, const WsNorm &n,
                                            float *__restrict__ x) {
  x[0] = (log2f(1.0f + (float)indeg) - n.m0) / n.s0;
  x[1] = (log2f(1.0f + (float)outdeg) - n.m1) / n.s1;
  x[2] = ((outdeg == 0 ? 1.0f : 0.0f) - n.m2) / n.s2;
}

// The features depend on (indeg, outdeg) only, and the MLP maps each row on
// its own, so vertices with both degrees below WS_T share the prediction of
// one row of a WS_T x WS_T table (row indeg * WS_T + outdeg); the others get
// rows of their own after the table.  The predictions are those of one row
// per vertex, bit for bit, for a fraction of the MLP work.
constexpr int32_t WS_T = 512;

__global__ void ws_table_kernel(WsNorm n, float *__restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (int64_t)WS_T * WS_T) ws_features((int32_t)(i / WS_T), (int32_t)(i % WS_T), n, X + i * 3);
}

__global__ void ws_rows_kernel(int32_t V, const int32_t *__restrict__ rp,
                               const int32_t *__restrict__ rrp, WsNorm n, float *__restrict__ X,
                               int32_t *__restrict__ row, unsigned int *__restrict__ n_own) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int32_t indeg = rrp[v + 1] - rrp[v];
  const int32_t outdeg = rp[v + 1] - rp[v];
  if (indeg < WS_T && outdeg < WS_T) {
    row[v] = indeg * WS_T + outdeg;
  } else {
    const int64_t r = (int64_t)WS_T * WS_T + atomicAdd(n_own, 1u);
    ws_features(indeg, outdeg, n, X + r * 3);
    row[v] = (int32_t)r;
  }
}

__global__ void ranks_from_pred_kernel(int32_t V, const float *__restrict__ Y,
                                       const int32_t *__restrict__ row, float ym, float ys,
                                       float *__restrict__ x, double *__restrict__ part) {
  __shared__ double s_red[256 / 64];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (v < V) {
    const float r = exp2f(Y[row[v]] * ys + ym);  // N * rank
    x[v] = r;
    s = (double)r;
  }
  s = gb_wave_sum_f64(s);
  if (gb_lane() == 0) s_red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// fixed-order FP64 sum of the block partials (one block)
__global__ void sum_parts_kernel(const double *__restrict__ part, int64_t nparts,
                                 double *__restrict__ total) {
  __shared__ double s_red[1024 / 64];
  double t = 0.0;
  for (int64_t i = threadIdx.x; i < nparts; i += blockDim.x) t += part[i];
  t = gb_wave_sum_f64(t);
  if (gb_lane() == 0) s_red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) a += s_red[k];
    *total = a;
  }
}

__global__ void scale_kernel(int32_t V, float *__restrict__ x, const double *__restrict__ total) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) x[v] = (float)((double)x[v] / *total);
}

// Per-graph inputs of the statistics kernel.
struct GStat {
  const uint32_t *w;
  const int32_t *rp;
  int64_t E;
  int32_t V;
};

// Weight sum, largest weight and largest out-degree of n graphs in one
// launch: blockIdx.y is the graph, the x blocks of a row stride over its
// edges and vertices; out[3 * g + k] (zeroed first) takes the block results.
__global__ __launch_bounds__(256) void graph_stats_multi_kernel(const GStat *__restrict__ gs,
                                                                unsigned long long *__restrict__ out) {
  const GStat G = gs[blockIdx.y];
  unsigned long long s = 0, mw = 0, md = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < G.E;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = G.w[i];
    s += x;
    mw = x > mw ? x : mw;
  }
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < G.V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long d = (unsigned long long)(G.rp[v + 1] - G.rp[v]);
    md = d > md ? d : md;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    s += __shfl_xor(s, d, 64);
    const unsigned long long a = __shfl_xor(mw, d, 64), b = __shfl_xor(md, d, 64);
    mw = a > mw ? a : mw;
    md = b > md ? b : md;
  }
  if (gb_lane() == 0) {
    unsigned long long *o = out + 3 * (int64_t)blockIdx.y;
    if (s) atomicAdd(&o[0], s);
    if (mw) atomicMax(&o[1], mw);
    if (md) atomicMax(&o[2], md);
  }
}

__global__ void fill_zero_u64_kernel(unsigned long long *a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0ull;
}

struct FeatWork {
  DevBuf in, out;
  DevBuf wx, wy, wpart, wrow;  // gb_pr_warm_start: rows, predictions, sums, row of each vertex
};
FeatWork &g_featw() {  // heap singleton, never destroyed (as the other workspaces)
  static FeatWork *w = new FeatWork();
  return *w;
}

// Features from the reduced statistics (the layout of gb_graph_features).
void features_from(const gb_graph g, int32_t nsrc, double sum_w, double max_w, double max_deg,
                   float *f) {
  const double V = g->V, E = (double)g->E, avg = E / V;
  f[0] = (float)std::log2(V);
  f[1] = (float)std::log2(avg);
  f[2] = (float)std::log2(std::max(1.0, max_w));
  f[3] = (float)std::log2(std::max(1.0, sum_w / E));
  f[4] = (float)std::log2(std::max(1.0, max_deg / avg));
  f[5] = (float)std::log2((double)nsrc);
}

}  // namespace

void gb_learned_release() {
  g_featw().in.release();
  g_featw().out.release();
  g_featw().wx.release();
  g_featw().wy.release();
  g_featw().wpart.release();
  g_featw().wrow.release();
}

extern "C" gb_status gb_graph_features_multi(int32_t n, const gb_graph *graphs, int32_t nsrc,
                                             float *h_feat, void *stream) {
  if (n < 0 || (n > 0 && (!graphs || !h_feat)) || nsrc < 1 || n > 65535)
    return gb_fail(GB_ERR_INVALID, "gb_graph_features_multi: bad arguments");
  if (n == 0) return GB_OK;
  std::vector<GStat> hs(n);
  int64_t maxE = 1;
  for (int32_t i = 0; i < n; i++) {
    const gb_graph g = graphs[i];
    if (!g || !g->rp || g->V <= 0 || g->E <= 0)
      return gb_fail(GB_ERR_INVALID, "gb_graph_features_multi: graph %d: bad graph", i);
    if (!g->w)
      return gb_fail(GB_ERR_INVALID, "gb_graph_features_multi: graph %d has no uint32 weights", i);
    hs[i] = GStat{g->w, g->rp, g->E, g->V};
    maxE = std::max<int64_t>(maxE, std::max<int64_t>(g->E, g->V));
  }
  hipStream_t st = gb_stream(stream);
  FeatWork &w = g_featw();
  GB_TRY(w.in.ensure(sizeof(GStat) * n));
  GB_TRY(w.out.ensure(sizeof(unsigned long long) * 3 * n));
  GB_HIP(hipMemcpyAsync(w.in.p, hs.data(), sizeof(GStat) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(fill_zero_u64_kernel, dim3((unsigned)std::min<int64_t>((3 * n + 255) / 256, 1024)),
                     dim3(256), 0, st, w.out.as<unsigned long long>(), (int64_t)3 * n);
  // x blocks per graph, from the largest graph: about 256 items per thread,
  // at most 256 blocks (the blocks of smaller graphs finish at once)
  const int64_t bx = std::max<int64_t>(1, std::min<int64_t>((maxE + 65535) / 65536, 256));
  hipLaunchKernelGGL(graph_stats_multi_kernel, dim3((unsigned)bx, (unsigned)n), dim3(256), 0, st,
                     w.in.as<GStat>(), w.out.as<unsigned long long>());
  GB_HIP(hipGetLastError());
  std::vector<unsigned long long> h((size_t)3 * n);
  GB_HIP(hipMemcpyAsync(h.data(), w.out.p, sizeof(unsigned long long) * 3 * n,
                        hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  for (int32_t i = 0; i < n; i++) {
    const gb_graph g = graphs[i];
    if (!g->max_w_valid) {  // the graph's cached largest weight (gb_graph_max_w)
      g->max_w = (uint32_t)h[3 * i + 1];
      g->max_w_valid = true;
    }
    features_from(g, nsrc, (double)h[3 * i], (double)h[3 * i + 1], (double)h[3 * i + 2],
                  h_feat + 6 * (int64_t)i);
  }
  return GB_OK;
}

extern "C" gb_status gb_graph_features(gb_graph g, int32_t nsrc, float *h_feat, void *stream) {
  if (!g || !h_feat || !g->rp || g->V <= 0 || g->E <= 0 || nsrc < 1)
    return gb_fail(GB_ERR_INVALID, "gb_graph_features: bad arguments");
  return gb_graph_features_multi(1, &g, nsrc, h_feat, stream);
}

extern "C" gb_status gb_pr_warm_start(gb_graph g, const float *d_params, const float *h_norm,
                                      float *d_x0, void *stream) {
  if (!g || !d_params || !h_norm || !d_x0 || !g->rp || !g->rrp || g->V <= 0)
    return gb_fail(GB_ERR_INVALID, "gb_pr_warm_start: bad arguments (needs both CSRs)");
  hipStream_t st = gb_stream(stream);
  const int32_t V = g->V;
  // reused across calls (allocating and freeing them cost more than the
  // kernels at 16 M vertices); the call ends synchronised, so they are free
  // again when it returns
  DevBuf &X = g_featw().wx, &Y = g_featw().wy, &part = g_featw().wpart;
  DevBuf &row = g_featw().wrow;
  const int64_t nb = (V + 255) / 256;
  const int64_t tab = (int64_t)WS_T * WS_T;
  GB_TRY(X.ensure(sizeof(float) * 3 * (size_t)(tab + V)));
  GB_TRY(Y.ensure(sizeof(float) * (size_t)(tab + V)));
  GB_TRY(row.ensure(sizeof(int32_t) * (size_t)V));
  GB_TRY(part.ensure(sizeof(double) * (nb + 2)));
  // h_norm: x mean[3], x std[3], y mean, y std
  const WsNorm n{h_norm[0], h_norm[3], h_norm[1], h_norm[4], h_norm[2], h_norm[5]};
  unsigned int *n_own = reinterpret_cast<unsigned int *>(part.as<double>() + nb + 1);
  GB_HIP(hipMemsetAsync(n_own, 0, sizeof(unsigned int), st));
  hipLaunchKernelGGL(ws_table_kernel, dim3((unsigned)((tab + 255) / 256)), dim3(256), 0, st, n,
                     X.as<float>());
  hipLaunchKernelGGL(ws_rows_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, g->rp, g->rrp, n,
                     X.as<float>(), row.as<int32_t>(), n_own);
  GB_HIP(hipGetLastError());
  unsigned int h_own = 0;
  GB_HIP(hipMemcpyAsync(&h_own, n_own, sizeof(h_own), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  // the table and the vertices' own rows: the FP32 inference, hidden layer on MFMA
  GB_TRY(gb_mlp_forward_mfma(d_params, 3, 1, X.as<float>(), tab + (int64_t)h_own, Y.as<float>(),
                             stream));
  hipLaunchKernelGGL(ranks_from_pred_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, Y.as<float>(),
                     (const int32_t *)row.as<int32_t>(), h_norm[6], h_norm[7], d_x0,
                     part.as<double>());
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(1024), 0, st, part.as<double>(), nb,
                     part.as<double>() + nb);
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, d_x0,
                     part.as<double>() + nb);
  GB_HIP(hipGetLastError());
  GB_HIP(hipStreamSynchronize(st));
  return GB_OK;
}
