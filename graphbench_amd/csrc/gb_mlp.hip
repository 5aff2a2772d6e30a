// gb_mlp.hip -- the small MLP of the learned delta / PageRank warm start.
//
// y = W3 relu(W2 relu(W1 x + b1) + b2) + b3, hidden width 64, FP32, trained
// with mean squared error and Adam.  Training batches are small: one thread
// per row, the parameters staged in LDS and read by all lanes at once (LDS
// broadcast).  Inference over one row per vertex (the PageRank warm start,
// millions of rows) runs the 64 x 64 layer on bf16 MFMA instead
// (mlp_infer_mfma_kernel below).
// Training step kernels:
//   mlp_forward      per row: pre-activations A1, A2 (kept) and output Y
//   mlp_backward     per row: output error G3, GA2 = (W3^T G3)*[A2>0],
//                    GA1 = (W2^T GA2)*[A1>0], squared-error partials
//   mlp_grad_reduce  every parameter's gradient summed over a chunk of rows
//   mlp_grad_finish  chunk partials summed in chunk order (deterministic)
//   mlp_adam         Adam update
// Parameter layout (floats): W1[64][din] b1[64] W2[64][64] b2[64] W3[dout][64] b3[dout].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {

constexpr int H = 64;
constexpr int MAXIN = 16;
constexpr int MAXOUT = 4;
constexpr int RB = 256;          // rows per block
constexpr int64_t CHUNK = 4096;  // rows per gradient partial

struct Lay {
  int din, dout, oW1, ob1, oW2, ob2, oW3, ob3, total;
};
__host__ __device__ inline Lay layout(int din, int dout) {
  Lay l;
  l.din = din;
  l.dout = dout;
  l.oW1 = 0;
  l.ob1 = l.oW1 + H * din;
  l.oW2 = l.ob1 + H;
  l.ob2 = l.oW2 + H * H;
  l.oW3 = l.ob2 + H;
  l.ob3 = l.oW3 + dout * H;
  l.total = l.ob3 + dout;
  return l;
}

__device__ inline void load_params(float *s, const float *__restrict__ P, int total) {
  for (int i = threadIdx.x; i < total; i += blockDim.x) s[i] = P[i];
  __syncthreads();
}

__global__ __launch_bounds__(RB) void mlp_forward_kernel(const float *__restrict__ P, int din,
                                                         int dout, const float *__restrict__ X,
                                                         int64_t n, float *__restrict__ Y,
                                                         float *__restrict__ A1,
                                                         float *__restrict__ A2) {
  extern __shared__ float sp[];
  const Lay L = layout(din, dout);
  load_params(sp, P, L.total);
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float x[MAXIN];
#pragma unroll
  for (int k = 0; k < MAXIN; k++) x[k] = k < din ? X[r * din + k] : 0.0f;
  float z1[H];
#pragma unroll
  for (int j = 0; j < H; j++) {
    float a = sp[L.ob1 + j];
    for (int k = 0; k < din; k++) a += sp[L.oW1 + j * din + k] * x[k];
    if (A1) A1[r * H + j] = a;
    z1[j] = a > 0.0f ? a : 0.0f;
  }
  float y[MAXOUT];
#pragma unroll
  for (int o = 0; o < MAXOUT; o++) y[o] = o < dout ? sp[L.ob3 + o] : 0.0f;
  for (int j = 0; j < H; j++) {
    float a = sp[L.ob2 + j];
#pragma unroll
    for (int k = 0; k < H; k++) a += sp[L.oW2 + j * H + k] * z1[k];
    if (A2) A2[r * H + j] = a;
    const float z2 = a > 0.0f ? a : 0.0f;
#pragma unroll
    for (int o = 0; o < MAXOUT; o++)
      if (o < dout) y[o] += sp[L.oW3 + o * H + j] * z2;
  }
#pragma unroll
  for (int o = 0; o < MAXOUT; o++)
    if (o < dout) Y[r * dout + o] = y[o];
}

__global__ __launch_bounds__(RB) void mlp_backward_kernel(
    const float *__restrict__ P, int din, int dout, const float *__restrict__ T, int64_t n,
    const float *__restrict__ Y, const float *__restrict__ A1, const float *__restrict__ A2,
    float *__restrict__ G3, float *__restrict__ GA1, float *__restrict__ GA2,
    double *__restrict__ loss_part) {
  extern __shared__ float sp[];
  __shared__ double s_red[RB / 64];
  const Lay L = layout(din, dout);
  load_params(sp, P, L.total);
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double loss = 0.0;
  if (r < n) {
    const float inv_n = 1.0f / (float)n;
    float g3[MAXOUT];
#pragma unroll
    for (int o = 0; o < MAXOUT; o++) {
      g3[o] = 0.0f;
      if (o < dout) {
        const float e = Y[r * dout + o] - T[r * dout + o];
        loss += 0.5 * (double)e * (double)e;
        g3[o] = e * inv_n;
        G3[r * dout + o] = g3[o];
      }
    }
    float ga2[H];
#pragma unroll
    for (int j = 0; j < H; j++) {
      float s = 0.0f;
#pragma unroll
      for (int o = 0; o < MAXOUT; o++)
        if (o < dout) s += sp[L.oW3 + o * H + j] * g3[o];
      ga2[j] = A2[r * H + j] > 0.0f ? s : 0.0f;
      GA2[r * H + j] = ga2[j];
    }
    for (int k = 0; k < H; k++) {
      float s = 0.0f;
#pragma unroll
      for (int j = 0; j < H; j++) s += sp[L.oW2 + j * H + k] * ga2[j];
      GA1[r * H + k] = A1[r * H + k] > 0.0f ? s : 0.0f;
    }
  }
  loss = gb_wave_sum_f64(loss);
  if (gb_lane() == 0) s_red[threadIdx.x >> 6] = loss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < RB / 64; k++) t += s_red[k];
    loss_part[blockIdx.x] = t;
  }
}

// grid.x: parameters (one thread each), grid.y: row chunk
__global__ void mlp_grad_reduce_kernel(int din, int dout, const float *__restrict__ X, int64_t n,
                                       const float *__restrict__ A1, const float *__restrict__ A2,
                                       const float *__restrict__ G3, const float *__restrict__ GA1,
                                       const float *__restrict__ GA2, float *__restrict__ part) {
  const Lay L = layout(din, dout);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= L.total) return;
  const int64_t r0 = (int64_t)blockIdx.y * CHUNK;
  const int64_t r1 = std::min<int64_t>(n, r0 + CHUNK);
  float acc = 0.0f;
  if (p < L.ob1) {  // W1[j][k] = sum GA1[r][j] * X[r][k]
    const int j = p / din, k = p % din;
    for (int64_t r = r0; r < r1; r++) acc += GA1[r * H + j] * X[r * din + k];
  } else if (p < L.oW2) {  // b1[j]
    const int j = p - L.ob1;
    for (int64_t r = r0; r < r1; r++) acc += GA1[r * H + j];
  } else if (p < L.ob2) {  // W2[j][k] = sum GA2[r][j] * relu(A1[r][k])
    const int q = p - L.oW2, j = q / H, k = q % H;
    for (int64_t r = r0; r < r1; r++) {
      const float z = A1[r * H + k];
      acc += GA2[r * H + j] * (z > 0.0f ? z : 0.0f);
    }
  } else if (p < L.oW3) {  // b2[j]
    const int j = p - L.ob2;
    for (int64_t r = r0; r < r1; r++) acc += GA2[r * H + j];
  } else if (p < L.ob3) {  // W3[o][j] = sum G3[r][o] * relu(A2[r][j])
    const int q = p - L.oW3, o = q / H, j = q % H;
    for (int64_t r = r0; r < r1; r++) {
      const float z = A2[r * H + j];
      acc += G3[r * dout + o] * (z > 0.0f ? z : 0.0f);
    }
  } else {  // b3[o]
    const int o = p - L.ob3;
    for (int64_t r = r0; r < r1; r++) acc += G3[r * dout + o];
  }
  part[(int64_t)blockIdx.y * L.total + p] = acc;
}

__global__ void mlp_grad_finish_kernel(int total, int64_t nchunks, const float *__restrict__ part,
                                       float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  float acc = 0.0f;
  for (int64_t c = 0; c < nchunks; c++) acc += part[c * total + p];
  grad[p] = acc;
}

__global__ void mlp_adam_kernel(int total, float *__restrict__ P, const float *__restrict__ g,
                                float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.0f - b1) * gp;
  const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

struct MlpWork {
  DevBuf Y, A1, A2, G3, GA1, GA2, part, lossp;
  void release() {
    for (DevBuf *b : {&Y, &A1, &A2, &G3, &GA1, &GA2, &part, &lossp}) b->release();
  }
};
MlpWork &mlpw() {
  static MlpWork *w = new MlpWork();
  return *w;
}

bool dims_ok(int din, int dout) { return din >= 1 && din <= MAXIN && dout >= 1 && dout <= MAXOUT; }

// ---- inference with the hidden 64 x 64 layer on MFMA (bf16 in, FP32 sums) --
// For the warm start's one row per vertex (millions of rows) the 64 x 64
// layer is a GEMM: a wave takes 32 rows, computes layer 1 in FP32 straight
// into the A-operand layout, runs 2 x 8 v_mfma_f32_32x32x8_bf16 against W2
// held in registers as B operands, adds b2 and applies ReLU in FP32, passes
// the 32 x 64 activations through LDS and finishes layer 3 in FP32.
// Operand layouts of v_mfma_f32_32x32x8_bf16 (lane l, element e):
//   A (32 x 8):  row l % 32, k = 4 (l / 32) + e, e < 4
//   B (8 x 32):  col l % 32, k = 4 (l / 32) + e
//   D (32 x 32): col l % 32, row 8 (r / 4) + 4 (l / 32) + r % 4 in register r
typedef short mf_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mf_f32x16 __attribute__((ext_vector_type(16)));
constexpr int MF_ROWS = 32;         // rows per wave tile
constexpr int MF_LD = H + 1;        // LDS row stride of the activations (no bank conflicts)
constexpr int MF_BLOCK = 256;

__device__ __forceinline__ short mf_bf16(float f) {  // round to nearest even (finite f)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__global__ __launch_bounds__(MF_BLOCK) void mlp_infer_mfma_kernel(
    const float *__restrict__ P, int din, int dout, const float *__restrict__ X, int64_t n,
    float *__restrict__ Y) {
  __shared__ float s_p[H * MAXIN + H + H + MAXOUT * H + MAXOUT];  // W1 b1 b2 W3 b3
  __shared__ float s_h[MF_BLOCK / GB_WAVE][MF_ROWS * MF_LD];
  const Lay L = layout(din, dout);
  float *sW1 = s_p, *sb1 = sW1 + H * din, *sb2 = sb1 + H, *sW3 = sb2 + H, *sb3 = sW3 + dout * H;
  for (int i = threadIdx.x; i < H * din; i += blockDim.x) sW1[i] = P[L.oW1 + i];
  for (int i = threadIdx.x; i < H; i += blockDim.x) {
    sb1[i] = P[L.ob1 + i];
    sb2[i] = P[L.ob2 + i];
  }
  for (int i = threadIdx.x; i < dout * H; i += blockDim.x) sW3[i] = P[L.oW3 + i];
  for (int i = threadIdx.x; i < dout; i += blockDim.x) sb3[i] = P[L.ob3 + i];
  __syncthreads();
  const int lane = gb_lane(), half = lane >> 5, c32 = lane & 31;
  float *sh = s_h[threadIdx.x / GB_WAVE];
  // W2 as B operands: tile nt (hidden outputs 32 nt ..), k-step ks (inputs 8 ks ..)
  mf_bf16x4 bw[2][8];
#pragma unroll
  for (int nt = 0; nt < 2; nt++)
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
#pragma unroll
      for (int e = 0; e < 4; e++)
        bw[nt][ks][e] = mf_bf16(P[L.oW2 + (32 * nt + c32) * H + 8 * ks + 4 * half + e]);
  const int64_t tiles = (n + MF_ROWS - 1) / MF_ROWS;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t tl = wave; tl < tiles; tl += nwaves) {
    const int64_t r0 = tl * MF_ROWS;
    int64_t row = r0 + c32;
    if (row > n - 1) row = n - 1;  // rows past the end recompute the last row
    float x[MAXIN];
#pragma unroll
    for (int d = 0; d < MAXIN; d++) x[d] = d < din ? X[row * din + d] : 0.0f;
    // layer 1 into A operands: hidden unit k = 8 ks + 4 half + e of this row
    mf_bf16x4 a[8];
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int k = 8 * ks + 4 * half + e;
        float z = sb1[k];
        for (int d = 0; d < din; d++) z += sW1[k * din + d] * x[d];
        a[ks][e] = mf_bf16(z > 0.0f ? z : 0.0f);
      }
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
      mf_f32x16 acc = {};
#pragma unroll
      for (int ks = 0; ks < 8; ks++)
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[ks], bw[nt][ks], acc, 0, 0, 0);
      const int col = 32 * nt + c32;
      const float b = sb2[col];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int i = 8 * (r / 4) + 4 * half + (r % 4);
        const float z = acc[r] + b;
        sh[i * MF_LD + col] = z > 0.0f ? z : 0.0f;
      }
    }
    // a wave's LDS accesses execute in order: its lanes' writes above are seen
    // by the reads below
    __builtin_amdgcn_wave_barrier();
    // layer 3: lane (half, c32) sums hidden units [32 half, 32 half + 32) of row c32
    float y[MAXOUT];
#pragma unroll
    for (int o = 0; o < MAXOUT; o++) {
      float s = 0.0f;
      if (o < dout)
        for (int j = 0; j < 32; j++) s += sW3[o * H + 32 * half + j] * sh[c32 * MF_LD + 32 * half + j];
      y[o] = s + __shfl_xor(s, 32, GB_WAVE);
    }
    __builtin_amdgcn_wave_barrier();  // the tile is rewritten next round
    const int64_t out_row = r0 + c32;
    if (half == 0 && out_row < n)
      for (int o = 0; o < dout; o++) Y[out_row * dout + o] = y[o] + sb3[o];
  }
}

// ---- the same inference in FP32 on MFMA -----------------------------------
// v_mfma_f32_32x32x2f32 multiplies FP32 operands exactly and sums in FP32, so
// the 64 x 64 layer stays an FP32 layer: 2 output tiles x 32 k-steps of 2
// inputs per 32 rows.  W2 lives in registers as B operands (64 per lane) and
// the layer-1 activations of the lane's row as A operands (32 per lane).
// Operand layouts of v_mfma_f32_32x32x2f32 (lane l):
//   A (32 x 2):  row l % 32, k = l / 32
//   B (2 x 32):  col l % 32, k = l / 32
//   D (32 x 32): as for the bf16 form above
__global__ __launch_bounds__(MF_BLOCK) void mlp_infer_mfma_f32_kernel(
    const float *__restrict__ P, int din, int dout, const float *__restrict__ X, int64_t n,
    float *__restrict__ Y) {
  __shared__ float s_p[H * MAXIN + H + H + MAXOUT * H + MAXOUT];  // W1 b1 b2 W3 b3
  __shared__ float s_h[MF_BLOCK / GB_WAVE][MF_ROWS * MF_LD];
  const Lay L = layout(din, dout);
  float *sW1 = s_p, *sb1 = sW1 + H * din, *sb2 = sb1 + H, *sW3 = sb2 + H, *sb3 = sW3 + dout * H;
  for (int i = threadIdx.x; i < H * din; i += blockDim.x) sW1[i] = P[L.oW1 + i];
  for (int i = threadIdx.x; i < H; i += blockDim.x) {
    sb1[i] = P[L.ob1 + i];
    sb2[i] = P[L.ob2 + i];
  }
  for (int i = threadIdx.x; i < dout * H; i += blockDim.x) sW3[i] = P[L.oW3 + i];
  for (int i = threadIdx.x; i < dout; i += blockDim.x) sb3[i] = P[L.ob3 + i];
  __syncthreads();
  const int lane = gb_lane(), half = lane >> 5, c32 = lane & 31;
  float *sh = s_h[threadIdx.x / GB_WAVE];
  // W2 as B operands: tile nt (hidden outputs 32 nt ..), k-step ks (inputs 2 ks ..)
  float bw[2][32];
#pragma unroll
  for (int nt = 0; nt < 2; nt++)
#pragma unroll
    for (int ks = 0; ks < 32; ks++) bw[nt][ks] = P[L.oW2 + (32 * nt + c32) * H + 2 * ks + half];
  const int64_t tiles = (n + MF_ROWS - 1) / MF_ROWS;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t tl = wave; tl < tiles; tl += nwaves) {
    const int64_t r0 = tl * MF_ROWS;
    int64_t row = r0 + c32;
    if (row > n - 1) row = n - 1;  // rows past the end recompute the last row
    float x[MAXIN];
#pragma unroll
    for (int d = 0; d < MAXIN; d++) x[d] = d < din ? X[row * din + d] : 0.0f;
    // layer 1 into A operands: hidden unit k = 2 ks + half of this row
    float a[32];
#pragma unroll
    for (int ks = 0; ks < 32; ks++) {
      const int k = 2 * ks + half;
      float z = sb1[k];
      for (int d = 0; d < din; d++) z += sW1[k * din + d] * x[d];
      a[ks] = z > 0.0f ? z : 0.0f;
    }
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
      mf_f32x16 acc = {};
#pragma unroll
      for (int ks = 0; ks < 32; ks++)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ks], bw[nt][ks], acc, 0, 0, 0);
      const int col = 32 * nt + c32;
      const float b = sb2[col];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int i = 8 * (r / 4) + 4 * half + (r % 4);
        const float z = acc[r] + b;
        sh[i * MF_LD + col] = z > 0.0f ? z : 0.0f;
      }
    }
    // a wave's LDS accesses execute in order: its lanes' writes above are seen
    // by the reads below
    __builtin_amdgcn_wave_barrier();
    float y[MAXOUT];
#pragma unroll
    for (int o = 0; o < MAXOUT; o++) {
      float s = 0.0f;
      if (o < dout)
        for (int j = 0; j < 32; j++) s += sW3[o * H + 32 * half + j] * sh[c32 * MF_LD + 32 * half + j];
      y[o] = s + __shfl_xor(s, 32, GB_WAVE);
    }
    __builtin_amdgcn_wave_barrier();  // the tile is rewritten next round
    const int64_t out_row = r0 + c32;
    if (half == 0 && out_row < n)
      for (int o = 0; o < dout; o++) Y[out_row * dout + o] = y[o] + sb3[o];
  }
}

}  // namespace

extern "C" gb_status gb_mlp_forward_mfma(const float *d_params, int32_t din, int32_t dout,
                                         const float *d_X, int64_t n, float *d_Y, void *stream) {
  if (!dims_ok(din, dout) || !d_params || (n > 0 && (!d_X || !d_Y)))
    return gb_fail(GB_ERR_INVALID, "gb_mlp_forward_mfma: bad arguments");
  if (n == 0) return GB_OK;
  const int64_t tiles = (n + MF_ROWS - 1) / MF_ROWS;
  const int64_t blocks = std::max<int64_t>(
      1, std::min<int64_t>((tiles + 3) / 4, (int64_t)gb_num_cus() * 8));
  hipLaunchKernelGGL(mlp_infer_mfma_f32_kernel, dim3((unsigned)blocks), dim3(MF_BLOCK), 0,
                     gb_stream(stream), d_params, din, dout, d_X, n, d_Y);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

extern "C" gb_status gb_mlp_forward_bf16(const float *d_params, int32_t din, int32_t dout,
                                         const float *d_X, int64_t n, float *d_Y, void *stream) {
  if (!dims_ok(din, dout) || !d_params || (n > 0 && (!d_X || !d_Y)))
    return gb_fail(GB_ERR_INVALID, "gb_mlp_forward_bf16: bad arguments");
  if (n == 0) return GB_OK;
  const int64_t tiles = (n + MF_ROWS - 1) / MF_ROWS;
  const int64_t blocks = std::max<int64_t>(
      1, std::min<int64_t>((tiles + 3) / 4, (int64_t)gb_num_cus() * 8));
  hipLaunchKernelGGL(mlp_infer_mfma_kernel, dim3((unsigned)blocks), dim3(MF_BLOCK), 0,
                     gb_stream(stream), d_params, din, dout, d_X, n, d_Y);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

void gb_mlp_release() { mlpw().release(); }

extern "C" int32_t gb_mlp_param_count(int32_t din, int32_t dout) {
  return dims_ok(din, dout) ? layout(din, dout).total : -1;
}

extern "C" gb_status gb_mlp_forward(const float *d_params, int32_t din, int32_t dout,
                                    const float *d_X, int64_t n, float *d_Y, void *stream) {
  if (!dims_ok(din, dout) || !d_params || (n > 0 && (!d_X || !d_Y)))
    return gb_fail(GB_ERR_INVALID, "gb_mlp_forward: bad arguments");
  if (n == 0) return GB_OK;
  const Lay L = layout(din, dout);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + RB - 1) / RB)), dim3(RB),
                     sizeof(float) * L.total, gb_stream(stream), d_params, din, dout, d_X, n, d_Y,
                     (float *)nullptr, (float *)nullptr);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

extern "C" gb_status gb_mlp_loss_grad(const float *d_params, int32_t din, int32_t dout,
                                      const float *d_X, const float *d_T, int64_t n,
                                      float *d_grad, double *loss, void *stream) {
  if (!dims_ok(din, dout) || !d_params || !d_grad || n <= 0 || !d_X || !d_T)
    return gb_fail(GB_ERR_INVALID, "gb_mlp_loss_grad: bad arguments");
  hipStream_t st = gb_stream(stream);
  const Lay L = layout(din, dout);
  MlpWork &w = mlpw();
  const int64_t nblk = (n + RB - 1) / RB;
  const int64_t nch = (n + CHUNK - 1) / CHUNK;
  GB_TRY(w.Y.ensure(sizeof(float) * n * dout));
  GB_TRY(w.A1.ensure(sizeof(float) * n * H));
  GB_TRY(w.A2.ensure(sizeof(float) * n * H));
  GB_TRY(w.G3.ensure(sizeof(float) * n * dout));
  GB_TRY(w.GA1.ensure(sizeof(float) * n * H));
  GB_TRY(w.GA2.ensure(sizeof(float) * n * H));
  GB_TRY(w.part.ensure(sizeof(float) * nch * L.total));
  GB_TRY(w.lossp.ensure(sizeof(double) * nblk));
  const size_t shm = sizeof(float) * L.total;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)nblk), dim3(RB), shm, st, d_params, din,
                     dout, d_X, n, w.Y.as<float>(), w.A1.as<float>(), w.A2.as<float>());
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((unsigned)nblk), dim3(RB), shm, st, d_params, din,
                     dout, d_T, n, w.Y.as<float>(), w.A1.as<float>(), w.A2.as<float>(),
                     w.G3.as<float>(), w.GA1.as<float>(), w.GA2.as<float>(),
                     w.lossp.as<double>());
  hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((L.total + 255) / 256, (unsigned)nch), dim3(256),
                     0, st, din, dout, d_X, n, w.A1.as<float>(), w.A2.as<float>(),
                     w.G3.as<float>(), w.GA1.as<float>(), w.GA2.as<float>(), w.part.as<float>());
  hipLaunchKernelGGL(mlp_grad_finish_kernel, dim3((L.total + 255) / 256), dim3(256), 0, st,
                     L.total, nch, w.part.as<float>(), d_grad);
  GB_HIP(hipGetLastError());
  if (loss) {
    std::vector<double> lp(nblk);
    GB_HIP(hipMemcpyAsync(lp.data(), w.lossp.p, sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    double t = 0.0;
    for (double x : lp) t += x;
    *loss = t / (double)n;
  }
  return GB_OK;
}

extern "C" gb_status gb_mlp_adam(float *d_params, const float *d_grad, float *d_m, float *d_v,
                                 int32_t total, float lr, float beta1, float beta2, float eps,
                                 int32_t step, void *stream) {
  if (!d_params || !d_grad || !d_m || !d_v || total <= 0 || step < 1)
    return gb_fail(GB_ERR_INVALID, "gb_mlp_adam: bad arguments");
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((total + 255) / 256), dim3(256), 0, gb_stream(stream),
                     total, d_params, d_grad, d_m, d_v, lr, beta1, beta2, eps, c1, c2);
  GB_HIP(hipGetLastError());
  return GB_OK;
}
