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

__global__ void weight_stats_kernel(const uint32_t *__restrict__ w, int64_t E,
                                    unsigned long long *__restrict__ sum) {
  unsigned long long s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x)
    s += w[i];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  if (gb_lane() == 0) atomicAdd(sum, s);
}

__global__ void max_degree_kernel(int32_t V, const int32_t *__restrict__ rp,
                                  unsigned int *__restrict__ out) {
  unsigned int m = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const unsigned int d = (unsigned int)(rp[v + 1] - rp[v]);
    m = d > m ? d : m;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned int o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if (gb_lane() == 0) atomicMax(out, m);
}

__global__ void vertex_features_kernel(int32_t V, const int32_t *__restrict__ rp,
                                       const int32_t *__restrict__ rrp, float m0, float s0,
                                       float m1, float s1, float m2, float s2,
                                       float *__restrict__ X) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const float indeg = (float)(rrp[v + 1] - rrp[v]);
  const float outdeg = (float)(rp[v + 1] - rp[v]);
  X[v * 3 + 0] = (log2f(1.0f + indeg) - m0) / s0;
  X[v * 3 + 1] = (log2f(1.0f + outdeg) - m1) / s1;
  X[v * 3 + 2] = ((outdeg == 0.0f ? 1.0f : 0.0f) - m2) / s2;
}

__global__ void ranks_from_pred_kernel(int32_t V, const float *__restrict__ Y, float ym, float ys,
                                       float *__restrict__ x, double *__restrict__ part) {
  __shared__ double s_red[256 / 64];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (v < V) {
    const float r = exp2f(Y[v] * ys + ym);  // N * rank
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

}  // namespace

extern "C" gb_status gb_graph_features(gb_graph g, int32_t nsrc, float *h_feat, void *stream) {
  if (!g || !h_feat || !g->rp || g->V <= 0 || g->E <= 0 || nsrc < 1)
    return gb_fail(GB_ERR_INVALID, "gb_graph_features: bad arguments");
  if (!g->w) return gb_fail(GB_ERR_INVALID, "gb_graph_features: graph has no uint32 weights");
  hipStream_t st = gb_stream(stream);
  uint32_t mw = 0;
  GB_TRY(gb_graph_max_w(g, &mw, st));
  DevBuf buf;
  GB_TRY(buf.ensure(16));
  GB_HIP(hipMemsetAsync(buf.p, 0, 16, st));
  unsigned long long *sum = buf.as<unsigned long long>();
  unsigned int *maxd = (unsigned int *)(sum + 1);
  const int cap = gb_num_cus() * 8;
  hipLaunchKernelGGL(weight_stats_kernel, dim3(std::max(1, (int)std::min<int64_t>((g->E + 255) / 256, cap))),
                     dim3(256), 0, st, g->w, g->E, sum);
  hipLaunchKernelGGL(max_degree_kernel, dim3(std::max(1, (int)std::min<int64_t>((g->V + 255) / 256, cap))),
                     dim3(256), 0, st, g->V, g->rp, maxd);
  GB_HIP(hipGetLastError());
  unsigned long long h[2] = {0, 0};
  GB_HIP(hipMemcpyAsync(h, buf.p, 16, hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  const double V = g->V, E = (double)g->E, avg = E / V;
  const double mean_w = (double)h[0] / E;
  const double maxdeg = (double)(uint32_t)h[1];
  h_feat[0] = (float)std::log2(V);
  h_feat[1] = (float)std::log2(avg);
  h_feat[2] = (float)std::log2(std::max(1.0, (double)mw));
  h_feat[3] = (float)std::log2(std::max(1.0, mean_w));
  h_feat[4] = (float)std::log2(std::max(1.0, maxdeg / avg));
  h_feat[5] = (float)std::log2((double)nsrc);
  return GB_OK;
}

extern "C" gb_status gb_pr_warm_start(gb_graph g, const float *d_params, const float *h_norm,
                                      float *d_x0, void *stream) {
  if (!g || !d_params || !h_norm || !d_x0 || !g->rp || !g->rrp || g->V <= 0)
    return gb_fail(GB_ERR_INVALID, "gb_pr_warm_start: bad arguments (needs both CSRs)");
  hipStream_t st = gb_stream(stream);
  const int32_t V = g->V;
  DevBuf X, Y, part;
  const int64_t nb = (V + 255) / 256;
  GB_TRY(X.ensure(sizeof(float) * 3 * (size_t)V));
  GB_TRY(Y.ensure(sizeof(float) * (size_t)V));
  GB_TRY(part.ensure(sizeof(double) * (nb + 1)));
  // h_norm: x mean[3], x std[3], y mean, y std
  hipLaunchKernelGGL(vertex_features_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, g->rp,
                     g->rrp, h_norm[0], h_norm[3], h_norm[1], h_norm[4], h_norm[2], h_norm[5],
                     X.as<float>());
  GB_HIP(hipGetLastError());
  // one row per vertex: the MFMA inference (a warm start needs no more than
  // bf16 hidden activations; the solve is judged by its own residual)
  GB_TRY(gb_mlp_forward_bf16(d_params, 3, 1, X.as<float>(), V, Y.as<float>(), stream));
  hipLaunchKernelGGL(ranks_from_pred_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, Y.as<float>(),
                     h_norm[6], h_norm[7], d_x0, part.as<double>());
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(1024), 0, st, part.as<double>(), nb,
                     part.as<double>() + nb);
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, d_x0,
                     part.as<double>() + nb);
  GB_HIP(hipGetLastError());
  GB_HIP(hipStreamSynchronize(st));
  return GB_OK;
}
