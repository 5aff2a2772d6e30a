// gb_msbc.hip -- batched multi-source BFS and Brandes betweenness for gfx950.
//
// Replaces the per-source loop of run_bc (btwcentral/csr_bc.c:40-63) for
// unweighted graphs: 64 sources share one wave, one source per lane, and the
// per-source state is stored vertex-major so that a wave reads one vertex's
// state for all 64 sources with one coalesced load:
//   sigma[v][64], delta[v][64]  FP64 path counts and dependencies
//   fm[d][v]                    64-bit mask: sources for which v is at level d
//   visited[v]                  64-bit mask: sources that have reached v
// Forward step d -> d+1 (pull, comp_num_spath csr_bc.c:70-96): one wave per
// candidate vertex x (some source has not reached it yet).  Lanes read 64 of
// x's in-edges at a time and test fm[d][u]; for each edge whose mask meets the
// unreached lanes, lane s adds sigma[u][s].  Each lane sums its in-edges in
// CSR order, so sigma is identical to a sequential pull.  Lanes whose sum is
// non-zero set level d+1: fm[d+1][x], visited[x] and the level list L_{d+1}.
// Backward step d (comp_delta csr_bc.c:98-123, pull form): one wave per vertex
// of L_d; lane s sums ((1+delta[w][s])*sigma[v][s])/sigma[w][s] over out-edges
// (v,w) with bit s of fm[d+1][w] set, in CSR order.
// Accumulation (csr_bc.c:57-61): bc[v] += delta[v][s] for v != source(s), in
// source order; so bc matches the sequential oracle bit for bit.
// Batches of 64 sources are independent: G of them run side by side in
// grid.y, so one launch per level serves up to 64*G sources.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {

constexpr int MB = 256;   // threads per block (4 waves)
constexpr int MBQ = 256;  // staged list appends per wave

__global__ void mbc_cand_init_kernel(int32_t V, const int32_t *__restrict__ rrp, int32_t hub,
                                     int32_t *__restrict__ cand, unsigned long long *count) {
  // vertices with 1..hub in-edges (none: never reached; more: the hub kernels)
  const int lane = gb_lane();
  for (int64_t b0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ll; b0 < V;
       b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = b0 + lane;
    const int32_t deg = v < V ? rrp[v + 1] - rrp[v] : 0;
    const bool keep = deg > 0 && deg <= hub;
    const uint64_t pos = gb_wave_reserve(count, keep ? 1u : 0u);
    if (keep) cand[pos] = (int32_t)v;
  }
}

// vertices with more than `hub` edges in a CSR, and the largest such degree
__global__ void mbc_hub_list_kernel(int32_t V, const int32_t *__restrict__ rp, int32_t hub,
                                    int32_t *__restrict__ list, unsigned long long *count,
                                    unsigned int *maxdeg) {
  const int lane = gb_lane();
  for (int64_t b0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ll; b0 < V;
       b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = b0 + lane;
    const int32_t deg = v < V ? rp[v + 1] - rp[v] : 0;
    const bool keep = deg > hub;
    const uint64_t pos = gb_wave_reserve(count, keep ? 1u : 0u);
    if (keep) {
      list[pos] = (int32_t)v;
      atomicMax(maxdeg, (unsigned int)deg);
    }
  }
}

// Pull over in-edges [beg, end) of one vertex: lane s adds sigma[u][s] for every
// in-edge u whose level-d mask has bit s within `need`, in CSR order, 8 loads
// in flight.  Without SIGMA only the union of matching bits is collected.
template <bool SIGMA>
__device__ __forceinline__ void pull_in(int lane, int32_t beg, int32_t end,
                                        const int32_t *__restrict__ rcol,
                                        const unsigned long long *__restrict__ fmd, uint64_t need,
                                        const double *__restrict__ sigma, int64_t gV, double &acc,
                                        uint64_t &any) {
  for (int32_t e0 = beg; e0 < end; e0 += GB_WAVE) {
    const int32_t e = e0 + lane;
    int32_t u = 0;
    uint64_t m = 0;
    if (e < end) {
      u = rcol[e];
      m = fmd[u] & need;
    }
    if (!SIGMA) {
#pragma unroll
      for (int k = 32; k >= 1; k >>= 1) m |= gb_shfl_u64(m, lane ^ k);
      any |= m;
      continue;
    }
    uint64_t hit = __ballot(m != 0);
    while (hit) {
      double val[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const bool ok = hit != 0;
        const int j = ok ? __ffsll((unsigned long long)hit) - 1 : 0;
        hit &= hit - 1;
        const int32_t uj = __shfl(u, j, GB_WAVE);
        const uint64_t mj = ok ? gb_shfl_u64(m, j) : 0ull;
        val[k] = ((mj >> lane) & 1ull) ? sigma[((int64_t)gV + uj) * 64 + lane] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) acc += val[k];  // +0.0 leaves acc unchanged
    }
  }
}

// Pull over out-edges [beg, end): lane s adds ((1+delta[w][s])*sv)/sigma[w][s]
// for every successor w one level deeper for source s, in CSR order.
__device__ __forceinline__ double pull_out(int lane, int32_t beg, int32_t end,
                                           const int32_t *__restrict__ col,
                                           const unsigned long long *__restrict__ fmd1,
                                           uint64_t mine, double sv,
                                           const double *__restrict__ sigma,
                                           const double *__restrict__ delta, int64_t gV) {
  double acc = 0.0;
  for (int32_t e0 = beg; e0 < end; e0 += GB_WAVE) {
    const int32_t e = e0 + lane;
    int32_t w = 0;
    uint64_t m = 0;
    if (e < end) {
      w = col[e];
      m = fmd1[w] & mine;
    }
    uint64_t hit = __ballot(m != 0);
    while (hit) {
      double dw[8], sw[8];
      bool on[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const bool ok = hit != 0;
        const int j = ok ? __ffsll((unsigned long long)hit) - 1 : 0;
        hit &= hit - 1;
        const int32_t wj = __shfl(w, j, GB_WAVE);
        const uint64_t mj = ok ? gb_shfl_u64(m, j) : 0ull;
        on[k] = (mj >> lane) & 1ull;
        const int64_t r = ((int64_t)gV + wj) * 64 + lane;
        dw[k] = on[k] ? delta[r] : 0.0;
        sw[k] = on[k] ? sigma[r] : 1.0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (on[k]) acc += ((1.0 + dw[k]) * sv) / sw[k];
    }
  }
  return acc;
}

// copy the shared candidate list to every group
__global__ void mbc_cand_copy_kernel(int64_t n, int32_t G, int64_t stride,
                                     int32_t *__restrict__ cand) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    for (int32_t g = 1; g < G; g++) cand[g * stride + i] = cand[i];
}

// sources of group g: slots [64g, 64g+64)
__global__ void mbc_seed_kernel(int32_t V, int32_t nsrc, const int32_t *__restrict__ src,
                                double *__restrict__ sigma, unsigned long long *__restrict__ fm0,
                                unsigned long long *__restrict__ visited,
                                int32_t *__restrict__ list0, unsigned long long *list0_count,
                                bool with_sigma) {
  const int32_t g = blockIdx.x;
  const int lane = threadIdx.x;  // blockDim = 64
  const int32_t s = g * 64 + lane;
  const int64_t gV = (int64_t)g * V;
  if (s < nsrc) {
    const int32_t v = src[s];
    atomicOr(&fm0[gV + v], 1ull << lane);
    atomicOr(&visited[gV + v], 1ull << lane);
    if (with_sigma) sigma[(gV + v) * 64 + lane] = 1.0;
  }
  __syncthreads();
  // level-0 list: each distinct source vertex once (the lowest lane holding it)
  if (s < nsrc) {
    const int32_t v = src[s];
    const uint64_t m = atomicOr(&fm0[gV + v], 0ull);  // read back from L2
    if ((int)(__ffsll((unsigned long long)m) - 1) == lane) {
      unsigned long long p = atomicAdd(&list0_count[g], 1ull);
      list0[gV + p] = v;
    }
  }
}

template <bool SIGMA>
__global__ __launch_bounds__(MB) void mbc_forward_kernel(
    int32_t V, int32_t d, const int32_t *__restrict__ rrp, const int32_t *__restrict__ rcol,
    const int32_t *__restrict__ cand_in, const unsigned long long *__restrict__ cand_in_count,
    int32_t *__restrict__ cand_out, unsigned long long *cand_out_count,
    const unsigned long long *__restrict__ fm_d, unsigned long long *__restrict__ fm_d1,
    unsigned long long *__restrict__ visited, const unsigned long long *__restrict__ active,
    double *__restrict__ sigma, int32_t *__restrict__ list_d1,
    unsigned long long *list_d1_count, int32_t *__restrict__ level_out) {
  const int32_t g = blockIdx.y;
  const int64_t gV = (int64_t)g * V;
  const int lane = gb_lane();
  const uint64_t n = cand_in_count[g];
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t act = active[g];
  const unsigned long long *fmd = fm_d + gV;
  // the level list and the candidate list are appended to by every wave:
  // stage the appends per wave in LDS
  __shared__ int32_t s_list[MB / GB_WAVE][MBQ], s_cand[MB / GB_WAVE][MBQ];
  gb_wave_queue<int32_t, MBQ> lq(s_list[threadIdx.x / GB_WAVE], list_d1 + gV, &list_d1_count[g]);
  gb_wave_queue<int32_t, MBQ> cq(s_cand[threadIdx.x / GB_WAVE], cand_out + gV,
                                 &cand_out_count[g]);
  for (uint64_t i = wave; i < n; i += nwaves) {
    const int32_t x = cand_in[gV + i];
    const uint64_t need = ~visited[gV + x] & act;
    if (need == 0) continue;  // dropped: every source has reached x
    double acc = 0.0;
    uint64_t any = 0;
    pull_in<SIGMA>(lane, rrp[x], rrp[x + 1], rcol, fmd, need, sigma, gV, acc, any);
    const uint64_t found = SIGMA ? (__ballot(acc > 0.0) & need) : (any & need);
    if (found) {
      if ((found >> lane) & 1ull) {
        if (SIGMA) sigma[(gV + x) * 64 + lane] = acc;
        if (level_out) level_out[((int64_t)g * 64 + lane) * V + x] = d + 1;
      }
      if (lane == 0) {
        fm_d1[gV + x] = found;
        visited[gV + x] |= found;
      }
    }
    lq.push(lane == 0 && found ? 1u : 0u, [&](uint32_t) { return x; });
    // still unreached by some source: a candidate again at the next level
    cq.push(lane == 0 && (need & ~found) ? 1u : 0u, [&](uint32_t) { return x; });
  }
  lq.flush();
  cq.flush();
}

__global__ __launch_bounds__(MB) void mbc_backward_kernel(
    int32_t V, int32_t hub, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
    const int32_t *__restrict__ list_d, const unsigned long long *__restrict__ list_d_count,
    const unsigned long long *__restrict__ fm_d, const unsigned long long *__restrict__ fm_d1,
    const double *__restrict__ sigma, double *__restrict__ delta) {
  const int32_t g = blockIdx.y;
  const int64_t gV = (int64_t)g * V;
  const int lane = gb_lane();
  const uint64_t n = list_d_count[g];
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wave; i < n; i += nwaves) {
    const int32_t v = list_d[gV + i];
    const int32_t beg = rp[v], end = rp[v + 1];
    if (end - beg > hub) continue;  // split over several waves by the hub kernels
    const uint64_t mine = fm_d[gV + v];
    const bool me = (mine >> lane) & 1ull;
    const double sv = me ? sigma[(gV + v) * 64 + lane] : 0.0;
    const double acc = pull_out(lane, beg, end, col, fm_d1 + gV, mine, sv, sigma, delta, gV);
    if (me) delta[(gV + v) * 64 + lane] = acc;
  }
}

// Forward step for in-degree hubs: one wave per (hub, chunk of `hub` edges).
// Chunk sums of sigma are integers, so atomicAdd keeps them exact.
template <bool SIGMA>
__global__ __launch_bounds__(MB) void mbc_forward_hub_kernel(
    int32_t V, int32_t hub, const int32_t *__restrict__ rrp, const int32_t *__restrict__ rcol,
    const int32_t *__restrict__ hubs, int32_t nhubs, int32_t maxch,
    const unsigned long long *__restrict__ fm_d, unsigned long long *__restrict__ fm_d1,
    const unsigned long long *__restrict__ visited, const unsigned long long *__restrict__ active,
    double *__restrict__ sigma, int32_t *__restrict__ list_d1, unsigned long long *list_d1_count) {
  const int32_t g = blockIdx.y;
  const int64_t gV = (int64_t)g * V;
  const int lane = gb_lane();
  const uint64_t items = (uint64_t)nhubs * maxch;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t act = active[g];
  for (uint64_t it = wave; it < items; it += nwaves) {
    const int32_t x = hubs[it / maxch];
    const int32_t k = (int32_t)(it % maxch);
    const int32_t beg = rrp[x] + k * hub;
    const int32_t end = min(rrp[x + 1], beg + hub);
    if (beg >= end) continue;
    const uint64_t need = ~visited[gV + x] & act;
    if (need == 0) continue;
    double acc = 0.0;
    uint64_t any = 0;
    pull_in<SIGMA>(lane, beg, end, rcol, fm_d + gV, need, sigma, gV, acc, any);
    const uint64_t found = SIGMA ? (__ballot(acc > 0.0) & need) : (any & need);
    if (found) {
      if (SIGMA && ((found >> lane) & 1ull)) atomicAdd(&sigma[(gV + x) * 64 + lane], acc);
      if (lane == 0) {
        const unsigned long long old = atomicOr(&fm_d1[gV + x], found);
        if (old == 0) {  // the first chunk to reach x enters it in level d+1
          const unsigned long long p = atomicAdd(&list_d1_count[g], 1ull);
          list_d1[gV + p] = x;
        }
      }
    }
  }
}

// After the hub chunks of a level: record the hubs' new sources as visited
// and, for BFS, write their levels.
__global__ void mbc_hub_finalize_kernel(int32_t V, int32_t d, const int32_t *__restrict__ hubs,
                                        int32_t nhubs,
                                        const unsigned long long *__restrict__ fm_d1,
                                        unsigned long long *__restrict__ visited,
                                        int32_t *__restrict__ level_out) {
  const int32_t g = blockIdx.y;
  const int64_t gV = (int64_t)g * V;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nhubs) return;
  const int32_t x = hubs[i];
  const uint64_t f = fm_d1[gV + x];
  if (!f) return;
  visited[gV + x] |= f;
  if (level_out)
    for (int s = 0; s < 64; s++)
      if ((f >> s) & 1ull) level_out[((int64_t)g * 64 + s) * V + x] = d + 1;
}

// Backward step for out-degree hubs: per-chunk partials ...
__global__ __launch_bounds__(MB) void mbc_backward_hub_kernel(
    int32_t V, int32_t hub, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
    const int32_t *__restrict__ hubs, int32_t nhubs, int32_t maxch,
    const unsigned long long *__restrict__ fm_d, const unsigned long long *__restrict__ fm_d1,
    const double *__restrict__ sigma, const double *__restrict__ delta,
    double *__restrict__ part) {
  const int32_t g = blockIdx.y;
  const int64_t gV = (int64_t)g * V;
  const int lane = gb_lane();
  const uint64_t items = (uint64_t)nhubs * maxch;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t it = wave; it < items; it += nwaves) {
    const int32_t x = hubs[it / maxch];
    const int32_t k = (int32_t)(it % maxch);
    const int32_t beg = rp[x] + k * hub;
    const int32_t end = min(rp[x + 1], beg + hub);
    if (beg >= end) continue;
    const uint64_t mine = fm_d[gV + x];
    if (mine == 0) continue;
    const bool me = (mine >> lane) & 1ull;
    const double sv = me ? sigma[(gV + x) * 64 + lane] : 0.0;
    const double acc = pull_out(lane, beg, end, col, fm_d1 + gV, mine, sv, sigma, delta, gV);
    part[((int64_t)g * items + it) * 64 + lane] = acc;
  }
}

// ... summed in chunk order (deterministic), one wave per hub.
__global__ __launch_bounds__(MB) void mbc_backward_combine_kernel(
    int32_t V, int32_t hub, const int32_t *__restrict__ rp, const int32_t *__restrict__ hubs,
    int32_t nhubs, int32_t maxch, const unsigned long long *__restrict__ fm_d,
    const double *__restrict__ part, double *__restrict__ delta) {
  const int32_t g = blockIdx.y;
  const int64_t gV = (int64_t)g * V;
  const int lane = gb_lane();
  const uint64_t items = (uint64_t)nhubs * maxch;
  const int32_t i = (int32_t)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= nhubs) return;
  const int32_t x = hubs[i];
  const uint64_t mine = fm_d[gV + x];
  if (!((mine >> lane) & 1ull)) return;
  const int32_t nch = (rp[x + 1] - rp[x] + hub - 1) / hub;
  double acc = 0.0;
  for (int32_t k = 0; k < nch; k++)
    acc += part[((int64_t)g * items + (int64_t)i * maxch + k) * 64 + lane];
  delta[(gV + x) * 64 + lane] = acc;
}

// bc[v] += delta[v][s] (v != source s) over groups and lanes in source order
__global__ void mbc_accum_kernel(int32_t V, int32_t G, int32_t nsrc,
                                 const int32_t *__restrict__ src,
                                 const double *__restrict__ delta, double *__restrict__ bc) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  double acc = bc[v];
  for (int32_t g = 0; g < G; g++) {
    const double *row = delta + ((int64_t)g * V + v) * 64;
    for (int k = 0; k < 64; k++) {
      const int32_t s = g * 64 + k;
      if (s < nsrc && src[s] != (int32_t)v) acc += row[k];
    }
  }
  bc[v] = acc;
}

struct LevelBuf {
  DevBuf fm, list;
};

struct MsWork {
  DevBuf sigma, delta, visited, cand[2], counts, src, active, allcand, hub_in, hub_out, hub_meta,
      part;
  std::vector<LevelBuf *> levels;
  void release() {
    for (DevBuf *b : {&sigma, &delta, &visited, &cand[0], &cand[1], &counts, &src, &active,
                      &allcand, &hub_in, &hub_out, &hub_meta, &part})
      b->release();
    for (auto *l : levels) delete l;
    levels.clear();
  }
};
MsWork &msw() {
  static MsWork *w = new MsWork();  // never destroyed (no frees at exit)
  return *w;
}

// edges per wave for high-degree vertices (GB_BC_HUB overrides)
int32_t hub_edges() {
  const char *e = getenv("GB_BC_HUB");
  const int v = e ? atoi(e) : 0;
  return v >= 64 ? v : 1024;
}

// Edges per backward hub chunk (one wave each): GB_BC_BCHUNK, else the hub
// threshold.
int32_t bc_bchunk(int32_t hub) {
  const char *e = getenv("GB_BC_BCHUNK");
  const int v = e ? atoi(e) : 0;
  return v >= 64 ? v : hub;
}

int wgrid(uint64_t n) {
  uint64_t b = (n + 3) / 4;
  uint64_t cap = (uint64_t)gb_num_cus() * 8;
  if (b > cap) b = cap;
  return b < 1 ? 1 : (int)b;
}

// counts layout (u64): [0,G) cand_in, [G,2G) cand_out, [2G,3G) list
enum { CNT_STRIDE = 3 };

// Runs one batch of up to 64*G sources.  with_sigma: Brandes (sigma, delta,
// accumulate into d_bc); otherwise BFS levels into d_level [nsrc][V].
gb_status ms_run(gb_graph g, const int32_t *h_src, int32_t nsrc, bool with_sigma,
                 double *d_bc, int32_t *d_level, hipStream_t st, int32_t *depth_out) {
  MsWork &w = msw();
  const int32_t V = g->V;
  const int32_t G = (nsrc + 63) / 64;
  const int64_t GV = (int64_t)G * V;
  const int32_t *rp = g->rp ? g->rp : g->rrp;
  const int32_t *col = g->rp ? g->col : g->rcol;
  const int32_t *rrp = g->rrp ? g->rrp : g->rp;
  const int32_t *rcol = g->rrp ? g->rcol : g->col;
  if (with_sigma) {
    GB_TRY(w.sigma.ensure(sizeof(double) * GV * 64));
    GB_TRY(w.delta.ensure(sizeof(double) * GV * 64));
    GB_HIP(hipMemsetAsync(w.sigma.p, 0, sizeof(double) * GV * 64, st));
    GB_HIP(hipMemsetAsync(w.delta.p, 0, sizeof(double) * GV * 64, st));
  }
  GB_TRY(w.visited.ensure(sizeof(uint64_t) * GV));
  GB_TRY(w.cand[0].ensure(sizeof(int32_t) * GV));
  GB_TRY(w.cand[1].ensure(sizeof(int32_t) * GV));
  GB_TRY(w.counts.ensure(sizeof(unsigned long long) * CNT_STRIDE * G));
  GB_TRY(w.src.ensure(sizeof(int32_t) * nsrc));
  GB_TRY(w.active.ensure(sizeof(uint64_t) * G));
  GB_HIP(hipMemsetAsync(w.visited.p, 0, sizeof(uint64_t) * GV, st));
  GB_HIP(hipMemcpyAsync(w.src.p, h_src, sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
  std::vector<uint64_t> act(G);
  for (int32_t k = 0; k < G; k++) {
    const int32_t n = std::min(64, nsrc - 64 * k);
    act[k] = n == 64 ? ~0ull : ((1ull << n) - 1);
  }
  GB_HIP(hipMemcpyAsync(w.active.p, act.data(), sizeof(uint64_t) * G, hipMemcpyHostToDevice, st));
  auto level = [&](int d) -> gb_status {
    while ((int)w.levels.size() <= d) w.levels.push_back(new LevelBuf());
    LevelBuf *L = w.levels[d];
    GB_TRY(L->fm.ensure(sizeof(uint64_t) * GV));
    GB_TRY(L->list.ensure(sizeof(int32_t) * GV));
    GB_HIP(hipMemsetAsync(L->fm.p, 0, sizeof(uint64_t) * GV, st));
    return GB_OK;
  };
  unsigned long long *cnt = w.counts.as<unsigned long long>();
  GB_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * CNT_STRIDE * G, st));
  GB_TRY(level(0));
  // vertices above `hub` edges are split into chunks of `hub` edges, one wave
  // each (in-degree for the forward pulls, out-degree for the backward ones)
  const int32_t hub = hub_edges();
  const int vgrid = std::max(1, std::min((V + 255) / 256, 2048));
  GB_TRY(w.hub_in.ensure(sizeof(int32_t) * (V + 1)));
  GB_TRY(w.hub_out.ensure(sizeof(int32_t) * (V + 1)));
  GB_TRY(w.hub_meta.ensure(sizeof(unsigned long long) * 4));
  GB_HIP(hipMemsetAsync(w.hub_meta.p, 0, sizeof(unsigned long long) * 4, st));
  unsigned long long *hm = w.hub_meta.as<unsigned long long>();
  hipLaunchKernelGGL(mbc_hub_list_kernel, dim3(vgrid), dim3(256), 0, st, V, rrp, hub,
                     w.hub_in.as<int32_t>(), hm + 0, (unsigned int *)(hm + 1));
  hipLaunchKernelGGL(mbc_hub_list_kernel, dim3(vgrid), dim3(256), 0, st, V, rp, hub,
                     w.hub_out.as<int32_t>(), hm + 2, (unsigned int *)(hm + 3));
  unsigned long long hmh[4];
  GB_HIP(hipMemcpyAsync(hmh, hm, sizeof(hmh), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  const int32_t nhub_in = (int32_t)hmh[0], nhub_out = (int32_t)hmh[2];
  const int32_t maxch_in = ((int32_t)(uint32_t)hmh[1] + hub - 1) / hub;
  const int32_t bch = bc_bchunk(hub);
  const int32_t maxch_out = ((int32_t)(uint32_t)hmh[3] + bch - 1) / bch;
  if (with_sigma && nhub_out > 0)
    GB_TRY(w.part.ensure(sizeof(double) * 64 * (size_t)G * nhub_out * maxch_out));
  // candidates: vertices with 1..hub in-edges, the same list for every group
  hipLaunchKernelGGL(mbc_cand_init_kernel, dim3(vgrid), dim3(256), 0, st, V, rrp, hub,
                     w.cand[0].as<int32_t>(), cnt);
  unsigned long long nc = 0;
  GB_HIP(hipMemcpyAsync(&nc, cnt, sizeof(nc), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  if (G > 1 && nc > 0) {
    hipLaunchKernelGGL(mbc_cand_copy_kernel, dim3(std::max<int64_t>(1, std::min<int64_t>((nc + 255) / 256, 4096))),
                       dim3(256), 0, st, (int64_t)nc, G, (int64_t)V, w.cand[0].as<int32_t>());
    std::vector<unsigned long long> c0(G, nc);
    GB_HIP(hipMemcpyAsync(cnt, c0.data(), sizeof(unsigned long long) * G, hipMemcpyHostToDevice,
                          st));
  }
  hipLaunchKernelGGL(mbc_seed_kernel, dim3(G), dim3(64), 0, st, V, nsrc, w.src.as<int32_t>(),
                     with_sigma ? w.sigma.as<double>() : nullptr,
                     w.levels[0]->fm.as<unsigned long long>(), w.visited.as<unsigned long long>(),
                     w.levels[0]->list.as<int32_t>(), cnt + 2 * G, with_sigma);
  GB_HIP(hipGetLastError());
  if (d_level) {
    // level 0 of each source: handled by a host-side fill below (sources are known)
  }
  std::vector<std::vector<unsigned long long>> list_n;  // [level][group]
  std::vector<unsigned long long> h(CNT_STRIDE * G);
  GB_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(unsigned long long) * CNT_STRIDE * G,
                        hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  list_n.push_back(std::vector<unsigned long long>(h.begin() + 2 * G, h.begin() + 3 * G));
  int cur = 0;
  int d = 0;
  for (;;) {
    uint64_t maxc = 0;
    for (int32_t k = 0; k < G; k++) maxc = std::max<uint64_t>(maxc, h[k]);
    if (maxc == 0 && nhub_in == 0) break;
    GB_TRY(level(d + 1));
    // zero cand_out and list counts for the next level
    GB_HIP(hipMemsetAsync(cnt + G, 0, sizeof(unsigned long long) * 2 * G, st));
    dim3 grid(wgrid(std::max<uint64_t>(maxc, 1)), G);
    if (with_sigma)
      hipLaunchKernelGGL(mbc_forward_kernel<true>, grid, dim3(MB), 0, st, V, d, rrp, rcol,
                         w.cand[cur].as<int32_t>(), cnt, w.cand[cur ^ 1].as<int32_t>(), cnt + G,
                         w.levels[d]->fm.as<unsigned long long>(),
                         w.levels[d + 1]->fm.as<unsigned long long>(),
                         w.visited.as<unsigned long long>(), w.active.as<unsigned long long>(),
                         w.sigma.as<double>(), w.levels[d + 1]->list.as<int32_t>(), cnt + 2 * G,
                         (int32_t *)nullptr);
    else
      hipLaunchKernelGGL(mbc_forward_kernel<false>, grid, dim3(MB), 0, st, V, d, rrp, rcol,
                         w.cand[cur].as<int32_t>(), cnt, w.cand[cur ^ 1].as<int32_t>(), cnt + G,
                         w.levels[d]->fm.as<unsigned long long>(),
                         w.levels[d + 1]->fm.as<unsigned long long>(),
                         w.visited.as<unsigned long long>(), w.active.as<unsigned long long>(),
                         (double *)nullptr, w.levels[d + 1]->list.as<int32_t>(), cnt + 2 * G,
                         d_level);
    GB_HIP(hipGetLastError());
    if (nhub_in > 0) {
      dim3 hgrid(wgrid((uint64_t)nhub_in * maxch_in), G);
      if (with_sigma)
        hipLaunchKernelGGL(mbc_forward_hub_kernel<true>, hgrid, dim3(MB), 0, st, V, hub, rrp, rcol,
                           w.hub_in.as<int32_t>(), nhub_in, maxch_in,
                           w.levels[d]->fm.as<unsigned long long>(),
                           w.levels[d + 1]->fm.as<unsigned long long>(),
                           w.visited.as<unsigned long long>(), w.active.as<unsigned long long>(),
                           w.sigma.as<double>(), w.levels[d + 1]->list.as<int32_t>(), cnt + 2 * G);
      else
        hipLaunchKernelGGL(mbc_forward_hub_kernel<false>, hgrid, dim3(MB), 0, st, V, hub, rrp,
                           rcol, w.hub_in.as<int32_t>(), nhub_in, maxch_in,
                           w.levels[d]->fm.as<unsigned long long>(),
                           w.levels[d + 1]->fm.as<unsigned long long>(),
                           w.visited.as<unsigned long long>(), w.active.as<unsigned long long>(),
                           (double *)nullptr, w.levels[d + 1]->list.as<int32_t>(), cnt + 2 * G);
      hipLaunchKernelGGL(mbc_hub_finalize_kernel, dim3((nhub_in + 255) / 256, G), dim3(256), 0, st,
                         V, d, w.hub_in.as<int32_t>(), nhub_in,
                         w.levels[d + 1]->fm.as<unsigned long long>(),
                         w.visited.as<unsigned long long>(), d_level);
      GB_HIP(hipGetLastError());
    }
    GB_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(unsigned long long) * CNT_STRIDE * G,
                          hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    std::vector<unsigned long long> ln(h.begin() + 2 * G, h.begin() + 3 * G);
    uint64_t tot = 0;
    for (auto x : ln) tot += x;
    if (tot == 0) break;
    list_n.push_back(ln);
    // the next level reads the surviving candidates
    GB_HIP(hipMemcpyAsync(cnt, cnt + G, sizeof(unsigned long long) * G, hipMemcpyDeviceToDevice,
                          st));
    for (int32_t k = 0; k < G; k++) h[k] = h[G + k];
    cur ^= 1;
    d++;
  }
  const int depth = (int)list_n.size();  // levels 0..depth-1 are non-empty
  if (depth_out) *depth_out = depth;
  if (!with_sigma) return GB_OK;
  // backward: levels depth-2 .. 0 (the deepest level has delta 0)
  for (int dd = depth - 2; dd >= 0; dd--) {
    uint64_t maxn = 0;
    for (auto x : list_n[dd]) maxn = std::max<uint64_t>(maxn, x);
    if (maxn == 0) continue;
    GB_HIP(hipMemcpyAsync(cnt + 2 * G, list_n[dd].data(), sizeof(unsigned long long) * G,
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mbc_backward_kernel, dim3(wgrid(maxn), G), dim3(MB), 0, st, V, hub, rp,
                       col, w.levels[dd]->list.as<int32_t>(), cnt + 2 * G,
                       w.levels[dd]->fm.as<unsigned long long>(),
                       w.levels[dd + 1]->fm.as<unsigned long long>(), w.sigma.as<double>(),
                       w.delta.as<double>());
    GB_HIP(hipGetLastError());
    if (nhub_out > 0) {
      hipLaunchKernelGGL(mbc_backward_hub_kernel, dim3(wgrid((uint64_t)nhub_out * maxch_out), G),
                         dim3(MB), 0, st, V, bch, rp, col, w.hub_out.as<int32_t>(), nhub_out,
                         maxch_out, w.levels[dd]->fm.as<unsigned long long>(),
                         w.levels[dd + 1]->fm.as<unsigned long long>(), w.sigma.as<double>(),
                         w.delta.as<double>(), w.part.as<double>());
      hipLaunchKernelGGL(mbc_backward_combine_kernel, dim3((nhub_out + 3) / 4, G), dim3(MB), 0, st,
                         V, bch, rp, w.hub_out.as<int32_t>(), nhub_out, maxch_out,
                         w.levels[dd]->fm.as<unsigned long long>(), w.part.as<double>(),
                         w.delta.as<double>());
      GB_HIP(hipGetLastError());
    }
  }
  hipLaunchKernelGGL(mbc_accum_kernel, dim3((V + 255) / 256), dim3(256), 0, st, V, G, nsrc,
                     w.src.as<int32_t>(), w.delta.as<double>(), d_bc);
  GB_HIP(hipGetLastError());
  return GB_OK;
}

__global__ void fill_i32_kernel(int64_t n, int32_t *__restrict__ a, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void set_sources_level_kernel(int32_t V, int32_t nsrc, const int32_t *__restrict__ src,
                                         int32_t *__restrict__ level) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsrc) level[(int64_t)s * V + src[s]] = 0;
}

}  // namespace

void gb_msbc_release() { msw().release(); }

extern "C" gb_status gb_bfs(gb_graph g, const int32_t *h_sources, int32_t nsrc, int32_t *d_level,
                            void *stream) {
  if (!g || (!h_sources && nsrc > 0) || nsrc < 0 || !d_level)
    return gb_fail(GB_ERR_INVALID, "gb_bfs: null argument");
  const int32_t V = g->V;
  for (int32_t i = 0; i < nsrc; i++)
    if (h_sources[i] < 0 || h_sources[i] >= V)
      return gb_fail(GB_ERR_INVALID, "gb_bfs: source %d out of range", h_sources[i]);
  if (nsrc == 0 || V == 0) return GB_OK;
  hipStream_t st = gb_stream(stream);
  const int64_t n = (int64_t)nsrc * V;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
                     dim3(256), 0, st, n, d_level, -1);
  // batches of up to 64*32 sources share a launch
  const int32_t per = 64 * 32;
  for (int32_t b0 = 0; b0 < nsrc; b0 += per) {
    const int32_t ns = std::min(per, nsrc - b0);
    GB_TRY(ms_run(g, h_sources + b0, ns, false, nullptr, d_level + (int64_t)b0 * V, st, nullptr));
    hipLaunchKernelGGL(set_sources_level_kernel, dim3((ns + 255) / 256), dim3(256), 0, st, V, ns,
                       msw().src.as<int32_t>(), d_level + (int64_t)b0 * V);
    GB_HIP(hipGetLastError());
  }
  GB_HIP(hipStreamSynchronize(st));
  return GB_OK;
}

extern "C" gb_status gb_bc(gb_graph g, const int32_t *h_sources, int32_t nsrc, int32_t batch,
                           double *d_bc, gb_bc_stats *stats, void *stream) {
  if (!g || (!h_sources && nsrc > 0) || nsrc < 0 || !d_bc)
    return gb_fail(GB_ERR_INVALID, "gb_bc: null argument");
  const int32_t V = g->V;
  for (int32_t i = 0; i < nsrc; i++)
    if (h_sources[i] < 0 || h_sources[i] >= V)
      return gb_fail(GB_ERR_INVALID, "gb_bc: source %d out of range", h_sources[i]);
  hipStream_t st = gb_stream(stream);
  hipEvent_t e0, e1;
  GB_HIP(hipEventCreate(&e0));
  GB_HIP(hipEventCreate(&e1));
  GB_HIP(hipEventRecord(e0, st));
  GB_HIP(hipMemsetAsync(d_bc, 0, sizeof(double) * (V > 0 ? V : 1), st));
  // sources per launch group: multiple of 64, default all (memory permitting)
  int64_t per = batch > 0 ? ((batch + 63) / 64) * 64 : nsrc;
  const int64_t cap_entries = (int64_t)3 << 28;  // sigma+delta entries budget (~12 GB)
  if (V > 0 && per * (int64_t)V > cap_entries) per = std::max<int64_t>(64, (cap_entries / V) / 64 * 64);
  int32_t max_depth = 0, batches = 0;
  for (int32_t b0 = 0; b0 < nsrc; b0 += (int32_t)per) {
    const int32_t ns = (int32_t)std::min<int64_t>(per, nsrc - b0);
    int32_t depth = 0;
    GB_TRY(ms_run(g, h_sources + b0, ns, true, d_bc, nullptr, st, &depth));
    max_depth = std::max(max_depth, depth - 1);
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
