// gb_device.hpp -- wave64 primitives shared by the CDNA4 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GB_WAVE 64

__device__ __forceinline__ int gb_lane() { return (int)__lane_id(); }

// Exclusive prefix sum over the 64 lanes of a wave; *total gets the sum.
__device__ __forceinline__ uint32_t gb_wave_excl_sum(uint32_t x, uint32_t *total) {
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < GB_WAVE; d <<= 1) {
    uint32_t y = __shfl_up(incl, d, GB_WAVE);
    if (gb_lane() >= d) incl += y;
  }
  *total = __shfl(incl, GB_WAVE - 1, GB_WAVE);
  return incl - x;
}

__device__ __forceinline__ uint64_t gb_shfl_u64(uint64_t v, int src) {
  uint32_t lo = __shfl((uint32_t)v, src, GB_WAVE);
  uint32_t hi = __shfl((uint32_t)(v >> 32), src, GB_WAVE);
  return ((uint64_t)hi << 32) | lo;
}
// Reserve `cnt` consecutive slots of a device queue for this lane.  One
// atomicAdd per wave.  Must be reached by every active lane of the wave.
__device__ __forceinline__ uint64_t gb_wave_reserve(unsigned long long *counter, uint32_t cnt) {
  uint32_t total;
  uint32_t excl = gb_wave_excl_sum(cnt, &total);
  uint64_t base = 0;
  if (total) {
    // the lowest active lane does the atomic
    uint64_t active = __ballot(1);
    int leader = __ffsll((unsigned long long)active) - 1;
    if (gb_lane() == leader) base = atomicAdd(counter, (unsigned long long)total);
    base = gb_shfl_u64(base, leader);
  }
  return base + excl;
}

// Appends of one wave to a device queue, staged in a wave-private LDS buffer
// of CAP elements and flushed with one counter atomicAdd per flush.  When
// every wave of a launch appends to the same queue, one atomic per wave and
// step on the single counter address serialises in the L2; staging divides
// the atomics by the buffer fill.  Order within the queue is not kept.  All 64
// lanes must call push and flush together (the fill n is wave-uniform).
template <class T, int CAP>
struct gb_wave_queue {
  T *buf;  // this wave's LDS slice, CAP elements
  T *q;
  unsigned long long *counter;
  uint32_t n;
  __device__ __forceinline__ gb_wave_queue(T *b, T *queue, unsigned long long *c)
      : buf(b), q(queue), counter(c), n(0) {}
  __device__ __forceinline__ uint64_t reserve_global(uint32_t total) {
    uint64_t base = 0;
    if (gb_lane() == 0) base = atomicAdd(counter, (unsigned long long)total);
    return gb_shfl_u64(base, 0);
  }
  __device__ __forceinline__ void flush() {
    if (n == 0) return;
    __builtin_amdgcn_wave_barrier();
    const uint64_t base = reserve_global(n);
    for (uint32_t k = gb_lane(); k < n; k += GB_WAVE) q[base + k] = buf[k];
    __builtin_amdgcn_wave_barrier();
    n = 0;
  }
  // this lane appends cnt elements gen(0), ..., gen(cnt - 1)
  template <class F>
  __device__ __forceinline__ void push(uint32_t cnt, F gen) {
    uint32_t total;
    const uint32_t excl = gb_wave_excl_sum(cnt, &total);
    if (total == 0) return;
    if (n + total > CAP) flush();
    if (total <= CAP) {
      for (uint32_t k = 0; k < cnt; k++) buf[n + excl + k] = gen(k);
      n += total;
    } else {  // larger than the buffer: straight to the queue
      const uint64_t base = reserve_global(total);
      for (uint32_t k = 0; k < cnt; k++) q[base + excl + k] = gen(k);
    }
  }
};

// Final flush of the wave queues of a block, by every thread of the block at
// the same point (a block barrier; no wave may have left early): one global
// reservation per queue and block instead of one per wave.  At the end of a
// phase nearly every wave has entries, and thousands of reservations on one
// counter serialise on its cache line.  Entries keep their order within a
// wave; waves are laid out in wave order.
template <class... Q>
__device__ __forceinline__ void gb_block_flush(Q &...qs) {
  constexpr int NQ = sizeof...(Q);
  constexpr int MAXW = 16;  // waves per block (1024 threads)
  __shared__ uint32_t s_n[NQ][MAXW];
  __shared__ unsigned long long s_base[NQ];
  const int w = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6), lane = gb_lane();
  const uint32_t ns[NQ] = {qs.n...};
  unsigned long long *const ctr[NQ] = {qs.counter...};
  if (lane == 0)
    for (int i = 0; i < NQ; i++) s_n[i][w] = ns[i];
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int i = (int)threadIdx.x;
    uint32_t tot = 0;
    for (int k = 0; k < nw; k++) tot += s_n[i][k];
    s_base[i] = tot ? atomicAdd(ctr[i], (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  int i = 0;
  (
      [&] {
        uint64_t off = s_base[i];
        for (int k = 0; k < w; k++) off += s_n[i][k];
        for (uint32_t k = lane; k < qs.n; k += GB_WAVE) qs.q[off + k] = qs.buf[k];
        qs.n = 0;
        ++i;
      }(),
      ...);
  __syncthreads();  // the LDS counts may be reused by a later flush
}

// Block minimum of v, then one atomicMin on *dst by thread 0 when it is below
// `none` (every thread of the block calls it at the same point).
__device__ __forceinline__ void gb_block_atomic_min(uint64_t v, uint64_t none,
                                                    unsigned long long *dst) {
  __shared__ uint64_t s_min[16];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const uint64_t o = gb_shfl_u64(v, gb_lane() ^ d);
    v = o < v ? o : v;
  }
  if (gb_lane() == 0) s_min[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) v = s_min[k] < v ? s_min[k] : v;
    if (v != none) atomicMin(dst, (unsigned long long)v);
  }
}

// Per-wave statistics without a shared address: wave w of a launch adds its
// counts to slots[w * GB_WSTAT + k] with no-return atomics (each slot is one
// wave's), and the slots are folded into the counters once, after the solve
// (gb_wstat_fold).  Thousands of waves adding to one counter per phase would
// serialise on its cache line.
constexpr int GB_WSTAT = 4;
__device__ __forceinline__ void gb_wstat_add(unsigned long long *slots, int k, uint32_t v) {
  if (v) {
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    atomicAdd(&slots[wave * GB_WSTAT + k], (unsigned long long)v);
  }
}

// Deterministic wave sum (fixed butterfly order).
__device__ __forceinline__ double gb_wave_sum_f64(double v) {
#pragma unroll
  for (int d = GB_WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, GB_WAVE);
  return v;
}

// Compensated (double-float) accumulation: value ~= s + c.
struct F2 {
  float s, c;
};
__device__ __forceinline__ void f2_add(F2 &a, float v) {
  float t = a.s + v;
  float bp = t - a.s;
  float err = (a.s - (t - bp)) + (v - bp);
  a.s = t;
  a.c += err;
}
__device__ __forceinline__ F2 f2_merge(F2 a, F2 b) {
  float t = a.s + b.s;
  float bp = t - a.s;
  float err = (a.s - (t - bp)) + (b.s - bp);
  F2 r;
  r.s = t;
  r.c = (a.c + b.c) + err;
  return r;
}
__device__ __forceinline__ float f2_value(F2 a) { return a.s + a.c; }
