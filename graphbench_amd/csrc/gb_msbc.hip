// gb_msbc.hip -- batched multi-source BFS and Brandes betweenness for gfx950.
//
// Replaces the per-source loop of run_bc (btwcentral/csr_bc.c:40-63) for
// unweighted graphs.  64 sources share one wave, one source per lane, and the
// per-source state is vertex-major so that a wave reads one vertex's state for
// all 64 sources with one coalesced load:
//   sigma[v][64], delta[v][64]  FP64 path counts and dependencies
//   visited[v]                  64-bit mask: sources that have reached v
//   fm[2][v]                    64-bit masks of the current and next level
// G groups of 64 sources run side by side in grid.y.
//
// Levels.  The vertices of level d (for some source of the group) form the
// list L_d, kept in a pool of (vertex, source mask) entries; BFS keeps two
// levels in the pool, Brandes every level for the backward pass.  The pool
// grows on the host between replays when a level could overflow it, so the
// memory per group is O(V + sum of level sizes) <= O(65 V) whatever the depth.
// A level d -> d+1 of the forward pass (comp_num_spath, csr_bc.c:70-96):
//   * control (one wave): the previous level's append count gives L_d's size
//     and the start of L_{d+1}; each group chooses push or pull from the
//     out-edges of L_d against the in-edges of vertices some source has not
//     reached (push while the former is below 1/16 of the latter);
//   * push: one wave per item of at most PUSH_CH out-edges of an L_d vertex;
//     an edge to x carries the sources of the vertex's level-d mask that have
//     not reached x: their sigma is added to x's by FP64 atomicAdd (integers:
//     exact in any order), the bits are ORed into x's next-level mask, and the
//     lane whose OR found the mask empty appends x to L_{d+1};
//   * pull: a scan lists the vertices with 1..hub in-edges that some source
//     has not reached; one wave per listed vertex sums, lane s, sigma[u][s]
//     over its in-edges u at level d for s, in CSR order; vertices above hub
//     in-edges are pulled in chunks of hub edges, one wave each, with FP64
//     atomicAdd;
//   * finalize: each L_{d+1} entry takes its final mask, sets visited, writes
//     BFS levels, and queues its out-edges as push items; the L_d masks are
//     cleared for reuse as level d+2.
// Backward (comp_delta, csr_bc.c:98-123, pull form), levels depth-2 .. 0: one
// wave per L_d vertex; lane s sums ((1+delta[w][s])*sigma[v][s])/sigma[w][s]
// over out-edges (v,w) with w at level d+1 for s, in CSR order (the operand
// order of csr_bc.c:116).  The level masks are looked up in the two dense
// mask arrays of the forward pass (zero when it ends): level l's entries are
// written from the pool into array l & 1 before level l - 1 is processed,
// and the entries of level l + 2 are cleared from it first.  Vertices above
// hub out-edges are summed in chunks whose partials are added in chunk order.
// Accumulation (csr_bc.c:57-61): bc[v] += delta[v][s] for v != source(s), in
// source order.
// Both level loops are state machines on the device, replayed from captured
// hipGraphs LPR levels at a time; the host reads one small state per replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {

constexpr int MB = 256;            // threads per block (4 waves)
constexpr int MBQ = 256;           // staged appends per wave
constexpr int32_t PUSH_CH = 256;   // out-edges per push item at most
constexpr int LPR = 16;            // levels per replay

enum { MODE_NONE = 0, MODE_PUSH = 1, MODE_PULL = 2 };
// per-group counters (u64), CN per group
enum {
  C_NEXT = 0,   // entries appended to L_{d+1}
  C_CUR = 1,    // entries of L_d
  C_MF = 2,     // out-edges of the entries of the newest finalized level
  C_MPULL = 3,  // in-edges of vertices some source has not reached
  C_CAND = 4,   // pull candidates of this level
  C_PQ0 = 5,    // push items for even levels (C_PQ0 + 1: odd levels)
  CN = 8
};

struct __align__(16) PushItem {
  int32_t v, beg, end, pad;
};

struct BcLoop {
  int32_t d;      // forward: the level last processed (-1 after seeding)
  int32_t done;   // forward finished
  int32_t stall;  // the pool or the level table needs room: the host grows it
  int32_t depth;  // forward: number of non-empty levels once done
  int32_t bd;     // backward: the level last processed
  int32_t bdone;  // backward finished
  int32_t pad[2];
};

struct BcCtx {
  int32_t V, G, hub, lvl_cap;
  int64_t E;  // edges of the CSR (both directions have E)
  const int32_t *rp, *col, *rrp, *rcol;
  const unsigned long long *act;  // [G] sources of the group
  unsigned long long *visited;    // [G][V]
  unsigned long long *fm[2];      // [G][V]
  int32_t *pool_v;                // [G][pool_cap]
  unsigned long long *pool_m;     // [G][pool_cap]
  int64_t pool_cap;
  int64_t *loff;  // [G][lvl_cap] pool offset of each level (two slots: BFS)
  int32_t keep;   // 1: every level stays in the pool (Brandes)
  PushItem *pq[2];  // [G][pq_cap] push items of even / odd levels
  int64_t pq_cap;
  int32_t *cand;            // [G][V]
  unsigned long long *cnt;  // [G][CN]
  int32_t *mode;            // [G]
  BcLoop *st;
  double *sigma, *delta;  // [G][V][64] (Brandes)
  int32_t *level_out;     // BFS levels, rows g * 64 + s (or null)
  const int32_t *hub_in;
  int32_t nhub_in, maxch_in;
  const int32_t *hub_out;
  int32_t nhub_out, maxch_out, bch;
  double *part;  // backward hub chunk partials
  int32_t alpha;  // push while alpha * (frontier out-edges) < (unvisited in-edges)
};

__device__ __forceinline__ int lslot(const BcCtx &c, int d) { return c.keep ? d : (d & 1); }

__device__ __forceinline__ uint64_t wave_id() {
  return ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
}
__device__ __forceinline__ uint64_t wave_count() {
  return ((uint64_t)gridDim.x * blockDim.x) >> 6;
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

// Sources of group g (blockDim 64, lane s = source 64 g + s): level 0.  The
// lowest lane holding a vertex appends it to L_0 with the mask of every lane
// holding it.
__global__ void mbc_seed_kernel(BcCtx c, const int32_t *__restrict__ src, int32_t nsrc) {
  const int32_t g = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t gV = (int64_t)g * c.V;
  const bool on = g * 64 + lane < nsrc;
  const int32_t v = on ? src[g * 64 + lane] : -1;
  unsigned long long m = 0;
  bool lowest = on;
  for (int j = 0; j < GB_WAVE; j++) {
    const int32_t vj = __shfl(v, j, GB_WAVE);
    if (on && vj == v) {
      m |= 1ull << j;
      if (j < lane) lowest = false;
    }
  }
  if (!on) return;
  if (c.sigma) c.sigma[(gV + v) * 64 + lane] = 1.0;
  if (lowest) {
    c.fm[0][gV + v] = m;
    const unsigned long long p = atomicAdd(&c.cnt[(int64_t)g * CN + C_NEXT], 1ull);
    c.pool_v[(int64_t)g * c.pool_cap + p] = v;
  }
}

// Forward control, one wave: advances to level d = st->d + 1 (see the header).
__global__ void mbc_fwd_ctl_kernel(BcCtx c) {
  BcLoop *st = c.st;
  if (st->done || st->stall) return;
  const int lane = threadIdx.x;
  const int d = st->d + 1;
  bool ok = true;
  unsigned long long total = 0;
  for (int g = lane; g < c.G; g += 64) {
    const unsigned long long n = c.cnt[(int64_t)g * CN + C_NEXT];
    total += n;
    if (c.keep) {
      const int64_t start = c.loff[(int64_t)g * c.lvl_cap + d];
      if (start + (int64_t)n + c.V > c.pool_cap || d + 1 >= c.lvl_cap) ok = false;
    }
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) total += __shfl_xor(total, k, 64);
  const bool all_ok = __ballot(!ok) == 0ull;
  __builtin_amdgcn_wave_barrier();
  if (!all_ok) {
    if (lane == 0) st->stall = 1;
    return;
  }
  if (total == 0) {
    if (lane == 0) {
      st->done = 1;
      st->depth = d;
    }
    return;
  }
  for (int g = lane; g < c.G; g += 64) {
    unsigned long long *cg = c.cnt + (int64_t)g * CN;
    const unsigned long long n = cg[C_NEXT];
    cg[C_CUR] = n;
    int64_t *lo = c.loff + (int64_t)g * c.lvl_cap;
    lo[lslot(c, d + 1)] = c.keep ? lo[d] + (int64_t)n : (int64_t)((d + 1) & 1) * c.V;
    const unsigned long long mf = cg[C_MF], mp = cg[C_MPULL];
    c.mode[g] = n == 0 ? MODE_NONE : ((unsigned long long)c.alpha * mf < mp ? MODE_PUSH : MODE_PULL);
    cg[C_NEXT] = 0;
    cg[C_MF] = 0;
    cg[C_CAND] = 0;
    cg[C_PQ0 + ((d + 1) & 1)] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) st->d = d;
}

// The next level's queue of group g
struct NextList {
  int32_t *base;
  unsigned long long *count;
};
__device__ __forceinline__ NextList next_list(const BcCtx &c, int32_t g, int d) {
  const int64_t start = c.loff[(int64_t)g * c.lvl_cap + lslot(c, d + 1)];
  return NextList{c.pool_v + (int64_t)g * c.pool_cap + start, &c.cnt[(int64_t)g * CN + C_NEXT]};
}

// Pull candidates: vertices with 1..hub in-edges that some source of the
// group has not reached.
__global__ __launch_bounds__(MB) void mbc_scan_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->done || c.st->stall || c.mode[g] != MODE_PULL) return;
  const int64_t gV = (int64_t)g * c.V;
  const uint64_t act = c.act[g];
  const int lane = gb_lane();
  __shared__ int32_t s_q[MB / GB_WAVE][MBQ];
  gb_wave_queue<int32_t, MBQ> q(s_q[threadIdx.x / GB_WAVE], c.cand + gV,
                                &c.cnt[(int64_t)g * CN + C_CAND]);
  for (int64_t b0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ll; b0 < c.V;
       b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = b0 + lane;
    bool keep = false;
    if (v < c.V) {
      const int32_t deg = c.rrp[v + 1] - c.rrp[v];
      keep = deg > 0 && deg <= c.hub && (~c.visited[gV + v] & act) != 0ull;
    }
    q.push(keep ? 1u : 0u, [&](uint32_t) { return (int32_t)v; });
  }
  gb_block_flush(q);
}

// Push over the level's items (see the header).
template <bool SIGMA>
__global__ __launch_bounds__(MB) void mbc_push_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->done || c.st->stall || c.mode[g] != MODE_PUSH) return;
  const int d = c.st->d;
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const uint64_t n = c.cnt[(int64_t)g * CN + C_PQ0 + (d & 1)];
  const PushItem *items = c.pq[d & 1] + (int64_t)g * c.pq_cap;
  const unsigned long long *fmc = c.fm[d & 1] + gV;
  unsigned long long *fmn = c.fm[(d + 1) & 1] + gV;
  const unsigned long long *vis = c.visited + gV;
  const uint64_t act = c.act[g];
  const NextList nl = next_list(c, g, d);
  __shared__ int32_t s_list[MB / GB_WAVE][MBQ];
  gb_wave_queue<int32_t, MBQ> lq(s_list[threadIdx.x / GB_WAVE], nl.base, nl.count);
  const uint64_t nwaves = wave_count();
  for (uint64_t i = wave_id(); i < n; i += nwaves) {
    const PushItem it = items[i];
    const uint64_t mu = fmc[it.v];
    const double su = SIGMA && ((mu >> lane) & 1ull) ? c.sigma[(gV + it.v) * 64 + lane] : 0.0;
    for (int32_t e0 = it.beg; e0 < it.end; e0 += GB_WAVE) {
      const int32_t e = e0 + lane;
      int32_t x = 0;
      uint64_t nb = 0;
      if (e < it.end) {
        x = c.col[e];
        nb = mu & ~vis[x] & act;
      }
      unsigned long long old = 1ull;
      if (nb) old = atomicOr(&fmn[x], (unsigned long long)nb);
      lq.push(nb && old == 0ull ? 1u : 0u, [&](uint32_t) { return x; });
      if (SIGMA) {
        uint64_t hit = __ballot(nb != 0);
        while (hit) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const bool okk = hit != 0;
            const int j = okk ? __ffsll((unsigned long long)hit) - 1 : 0;
            hit &= hit - 1;
            const int32_t xj = __shfl(x, j, GB_WAVE);
            const uint64_t nj = okk ? gb_shfl_u64(nb, j) : 0ull;
            if ((nj >> lane) & 1ull) atomicAdd(&c.sigma[(gV + xj) * 64 + lane], su);
          }
        }
      }
    }
  }
  gb_block_flush(lq);
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

template <bool SIGMA>
__global__ __launch_bounds__(MB) void mbc_pull_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->done || c.st->stall || c.mode[g] != MODE_PULL) return;
  const int d = c.st->d;
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const uint64_t n = c.cnt[(int64_t)g * CN + C_CAND];
  const unsigned long long *fmc = c.fm[d & 1] + gV;
  unsigned long long *fmn = c.fm[(d + 1) & 1] + gV;
  const uint64_t act = c.act[g];
  const NextList nl = next_list(c, g, d);
  __shared__ int32_t s_list[MB / GB_WAVE][MBQ];
  gb_wave_queue<int32_t, MBQ> lq(s_list[threadIdx.x / GB_WAVE], nl.base, nl.count);
  const uint64_t nwaves = wave_count();
  for (uint64_t i = wave_id(); i < n; i += nwaves) {
    const int32_t x = c.cand[gV + i];
    const uint64_t need = ~c.visited[gV + x] & act;
    double acc = 0.0;
    uint64_t any = 0;
    pull_in<SIGMA>(lane, c.rrp[x], c.rrp[x + 1], c.rcol, fmc, need, c.sigma, gV, acc, any);
    const uint64_t found = SIGMA ? (__ballot(acc > 0.0) & need) : (any & need);
    if (found) {
      if (SIGMA && ((found >> lane) & 1ull)) c.sigma[(gV + x) * 64 + lane] = acc;
      if (lane == 0) fmn[x] = found;  // one wave per candidate: a plain store
    }
    lq.push(lane == 0 && found ? 1u : 0u, [&](uint32_t) { return x; });
  }
  gb_block_flush(lq);
}

// Pull for in-degree hubs: one wave per (hub, chunk of `hub` edges).  Chunk
// sums of sigma are integers, so atomicAdd keeps them exact.
template <bool SIGMA>
__global__ __launch_bounds__(MB) void mbc_pull_hub_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->done || c.st->stall || c.mode[g] != MODE_PULL || c.nhub_in == 0) return;
  const int d = c.st->d;
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const uint64_t items = (uint64_t)c.nhub_in * c.maxch_in;
  const unsigned long long *fmc = c.fm[d & 1] + gV;
  unsigned long long *fmn = c.fm[(d + 1) & 1] + gV;
  const uint64_t act = c.act[g];
  const NextList nl = next_list(c, g, d);
  __shared__ int32_t s_list[MB / GB_WAVE][MBQ];
  gb_wave_queue<int32_t, MBQ> lq(s_list[threadIdx.x / GB_WAVE], nl.base, nl.count);
  const uint64_t nwaves = wave_count();
  for (uint64_t it = wave_id(); it < items; it += nwaves) {
    const int32_t x = c.hub_in[it / c.maxch_in];
    const int32_t k = (int32_t)(it % c.maxch_in);
    const int32_t beg = c.rrp[x] + k * c.hub;
    const int32_t end = min(c.rrp[x + 1], beg + c.hub);
    const uint64_t need = beg < end ? ~c.visited[gV + x] & act : 0ull;
    uint64_t found = 0;
    if (need) {
      double acc = 0.0;
      uint64_t any = 0;
      pull_in<SIGMA>(lane, beg, end, c.rcol, fmc, need, c.sigma, gV, acc, any);
      found = SIGMA ? (__ballot(acc > 0.0) & need) : (any & need);
      if (SIGMA && ((found >> lane) & 1ull)) atomicAdd(&c.sigma[(gV + x) * 64 + lane], acc);
    }
    unsigned long long old = 1ull;
    if (found && lane == 0) old = atomicOr(&fmn[x], (unsigned long long)found);
    lq.push(lane == 0 && old == 0ull ? 1u : 0u, [&](uint32_t) { return x; });
  }
  gb_block_flush(lq);
}

// Finalize L_{d+1} (final masks, visited, BFS levels, push items, the counts
// that choose the next mode) and clear the L_d masks (see the header).
__global__ __launch_bounds__(MB) void mbc_fin_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->done || c.st->stall) return;
  const int d = c.st->d;
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const unsigned long long *cg = c.cnt + (int64_t)g * CN;
  const uint64_t n_next = cg[C_NEXT];
  const uint64_t n_cur = d >= 0 ? cg[C_CUR] : 0;
  const int64_t *lo = c.loff + (int64_t)g * c.lvl_cap;
  const int64_t pbase = (int64_t)g * c.pool_cap;
  const int64_t base_next = pbase + lo[lslot(c, d + 1)];
  const int64_t base_cur = d >= 0 ? pbase + lo[lslot(c, d)] : 0;
  unsigned long long *fmn = c.fm[(d + 1) & 1] + gV;
  unsigned long long *fmc = c.fm[d & 1] + gV;
  const uint64_t act = c.act[g];
  __shared__ PushItem s_pq[MB / GB_WAVE][MBQ / 2];
  gb_wave_queue<PushItem, MBQ / 2> pq(s_pq[threadIdx.x / GB_WAVE],
                                      c.pq[(d + 1) & 1] + (int64_t)g * c.pq_cap,
                                      &c.cnt[(int64_t)g * CN + C_PQ0 + ((d + 1) & 1)]);
  unsigned long long mf = 0, mp = 0;
  const uint64_t n = n_next > n_cur ? n_next : n_cur;
  for (uint64_t b0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; b0 < n;
       b0 += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = b0 + lane;
    int32_t v = 0, r0 = 0, r1 = 0;
    if (i < n_next) {
      v = c.pool_v[base_next + i];
      const unsigned long long m = fmn[v];
      c.pool_m[base_next + i] = m;
      const unsigned long long old = c.visited[gV + v];
      const unsigned long long nv = old | m;
      c.visited[gV + v] = nv;
      if (c.level_out) {
        for (unsigned long long b = m; b; b &= b - 1) {
          const int s = __ffsll(b) - 1;
          c.level_out[((int64_t)g * 64 + s) * c.V + v] = d + 1;
        }
      }
      r0 = c.rp[v];
      r1 = c.rp[v + 1];
      mf += (unsigned long long)(r1 - r0);
      if (old != act && nv == act) mp += (unsigned long long)(c.rrp[v + 1] - c.rrp[v]);
    }
    const uint32_t k = (uint32_t)((r1 - r0 + PUSH_CH - 1) / PUSH_CH);
    pq.push(k, [&](uint32_t q) {
      const int32_t b = r0 + (int32_t)q * PUSH_CH;
      return PushItem{v, b, min(r1, b + PUSH_CH), 0};
    });
    if (i < n_cur) fmc[c.pool_v[base_cur + i]] = 0ull;
  }
  gb_block_flush(pq);
  __shared__ unsigned long long s_mf[MB / GB_WAVE], s_mp[MB / GB_WAVE];
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) {
    mf += __shfl_xor(mf, k, 64);
    mp += __shfl_xor(mp, k, 64);
  }
  if (lane == 0) {
    s_mf[threadIdx.x / GB_WAVE] = mf;
    s_mp[threadIdx.x / GB_WAVE] = mp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tf = 0, tp = 0;
    for (int w = 0; w < MB / GB_WAVE; w++) {
      tf += s_mf[w];
      tp += s_mp[w];
    }
    if (tf) atomicAdd(&c.cnt[(int64_t)g * CN + C_MF], tf);
    if (tp) atomicAdd(&c.cnt[(int64_t)g * CN + C_MPULL], (unsigned long long)0 - tp);
  }
}

// ---- backward -------------------------------------------------------------

// Backward control (one thread): level st->bd - 1, done below 0.
__global__ void mbc_bwd_ctl_kernel(BcCtx c) {
  BcLoop *st = c.st;
  if (st->bdone) return;
  const int bd = st->bd - 1;
  if (bd < 0)
    st->bdone = 1;
  else
    st->bd = bd;
}

// Level l's entries into mask array l & 1 (write) or their masks cleared
// (!write), for l = st->bd + off.
template <bool WRITE>
__global__ void mbc_bwd_masks_kernel(BcCtx c, int off) {
  const int32_t g = blockIdx.y;
  if (c.st->bdone) return;
  const int l = c.st->bd + off;
  if (l < 0 || l >= c.st->depth) return;
  const int64_t gV = (int64_t)g * c.V;
  const int64_t *lo = c.loff + (int64_t)g * c.lvl_cap;
  const int64_t base = (int64_t)g * c.pool_cap + lo[l];
  const int64_t n = lo[l + 1] - lo[l];
  unsigned long long *m = c.fm[l & 1] + gV;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m[c.pool_v[base + i]] = WRITE ? c.pool_m[base + i] : 0ull;
}

// The backward pass sums, per vertex v and source s, psi[w][s] over the
// successors w one level deeper, psi = (1 + delta) / sigma: delta[v][s] =
// sigma[v][s] * sum psi (csr_bc.c:98-123 adds (1 + delta[w]) * sigma[v] /
// sigma[w] per successor; the same sum, factored, so a successor costs one
// 512-byte row read instead of two).  Once delta[v][s] is known, psi[v][s]
// replaces sigma[v][s], which no later level reads.
//
// Pull over out-edges [beg, end): lane s adds psi[w][s] (held in the sigma
// array) for every successor w one level deeper for source s (its mask in
// fmd1), in CSR order.
__device__ __forceinline__ double pull_out(int lane, int32_t beg, int32_t end,
                                           const int32_t *__restrict__ col,
                                           const unsigned long long *__restrict__ fmd1,
                                           uint64_t mine, const double *__restrict__ psi,
                                           int64_t gV) {
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
      double pw[8];
      bool on[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const bool ok = hit != 0;
        const int j = ok ? __ffsll((unsigned long long)hit) - 1 : 0;
        hit &= hit - 1;
        const int32_t wj = __shfl(w, j, GB_WAVE);
        const uint64_t mj = ok ? gb_shfl_u64(m, j) : 0ull;
        on[k] = (mj >> lane) & 1ull;
        pw[k] = on[k] ? psi[((int64_t)gV + wj) * 64 + lane] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (on[k]) acc += pw[k];
    }
  }
  return acc;
}

__global__ __launch_bounds__(MB) void mbc_bwd_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->bdone) return;
  const int bd = c.st->bd;
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const int64_t *lo = c.loff + (int64_t)g * c.lvl_cap;
  const int64_t base = (int64_t)g * c.pool_cap + lo[bd];
  const uint64_t n = (uint64_t)(lo[bd + 1] - lo[bd]);
  const unsigned long long *fm1 = c.fm[(bd + 1) & 1] + gV;
  const bool deepest = bd == c.st->depth - 1;
  const uint64_t nwaves = wave_count();
  for (uint64_t i = wave_id(); i < n; i += nwaves) {
    const int32_t v = c.pool_v[base + i];
    const int32_t beg = c.rp[v], end = c.rp[v + 1];
    if (end - beg > c.bch) continue;  // split over several waves by the hub kernels
    const uint64_t mine = c.pool_m[base + i];
    const bool me = (mine >> lane) & 1ull;
    // the deepest level has no successors
    const double acc = deepest ? 0.0 : pull_out(lane, beg, end, c.col, fm1, mine, c.sigma, gV);
    if (me) {
      const int64_t r = (gV + v) * 64 + lane;
      const double sv = c.sigma[r];
      const double d = sv * acc;
      c.delta[r] = d;
      c.sigma[r] = (1.0 + d) / sv;  // psi
    }
  }
}

// Backward for out-degree hubs: per-chunk partials ...
__global__ __launch_bounds__(MB) void mbc_bwd_hub_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->bdone || c.nhub_out == 0) return;
  const int bd = c.st->bd;
  if (bd == c.st->depth - 1) return;  // no successors: the combine writes 0
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const uint64_t items = (uint64_t)c.nhub_out * c.maxch_out;
  const unsigned long long *fm0 = c.fm[bd & 1] + gV;
  const unsigned long long *fm1 = c.fm[(bd + 1) & 1] + gV;
  const uint64_t nwaves = wave_count();
  for (uint64_t it = wave_id(); it < items; it += nwaves) {
    const int32_t x = c.hub_out[it / c.maxch_out];
    const int32_t k = (int32_t)(it % c.maxch_out);
    const int32_t beg = c.rp[x] + k * c.bch;
    const int32_t end = min(c.rp[x + 1], beg + c.bch);
    if (beg >= end) continue;
    const uint64_t mine = fm0[x];
    if (mine == 0) continue;
    const double acc = pull_out(lane, beg, end, c.col, fm1, mine, c.sigma, gV);
    c.part[((int64_t)g * items + it) * 64 + lane] = acc;
  }
}

// ... summed in chunk order (deterministic), one wave per hub.
__global__ __launch_bounds__(MB) void mbc_bwd_combine_kernel(BcCtx c) {
  const int32_t g = blockIdx.y;
  if (c.st->bdone || c.nhub_out == 0) return;
  const int bd = c.st->bd;
  const int64_t gV = (int64_t)g * c.V;
  const int lane = gb_lane();
  const uint64_t items = (uint64_t)c.nhub_out * c.maxch_out;
  const unsigned long long *fm0 = c.fm[bd & 1] + gV;
  const uint64_t nwaves = wave_count();
  for (uint64_t i = wave_id(); i < (uint64_t)c.nhub_out; i += nwaves) {
    const int32_t x = c.hub_out[i];
    const uint64_t mine = fm0[x];
    if (!((mine >> lane) & 1ull)) continue;
    const int32_t nch = bd == c.st->depth - 1 ? 0 : (c.rp[x + 1] - c.rp[x] + c.bch - 1) / c.bch;
    double acc = 0.0;
    for (int32_t k = 0; k < nch; k++)
      acc += c.part[((int64_t)g * items + (int64_t)i * c.maxch_out + k) * 64 + lane];
    const int64_t r = (gV + x) * 64 + lane;
    const double sv = c.sigma[r];
    const double d = sv * acc;
    c.delta[r] = d;
    c.sigma[r] = (1.0 + d) / sv;  // psi
  }
}

// bc[v] += delta[v][s] (v != source s) over groups and lanes in source order.
// A wave takes ACC_V consecutive vertices per group: it reads their delta rows
// with one coalesced 512-byte load each (lane j: source j) into an LDS tile
// (row stride 65 doubles, so that the column reads below hit distinct
// banks), then lane i sums vertex i's row in source order.  Reading the rows
// straight from each vertex's thread touched 64 lines per load and re-fetched
// every line up to 16 times.
constexpr int ACC_V = 32;
constexpr int ACC_BLOCK = 256;
__global__ __launch_bounds__(ACC_BLOCK) void mbc_accum_kernel(int32_t V, int32_t G, int32_t nsrc,
                                                              const int32_t *__restrict__ src,
                                                              const double *__restrict__ delta,
                                                              double *__restrict__ bc) {
  __shared__ double s_tile[ACC_BLOCK / GB_WAVE][ACC_V][GB_WAVE + 1];
  const int lane = gb_lane(), wave = threadIdx.x / GB_WAVE;
  const int64_t v0 = ((int64_t)blockIdx.x * (ACC_BLOCK / GB_WAVE) + wave) * ACC_V;
  if (v0 >= V) return;  // wave-uniform: no barrier below
  double (*tile)[GB_WAVE + 1] = s_tile[wave];
  const int64_t v = v0 + lane;  // lanes < ACC_V sum vertex v
  const bool mine = lane < ACC_V && v < V;
  double acc = mine ? bc[v] : 0.0;
  for (int32_t g = 0; g < G; g++) {
    const int32_t sj = g * 64 + lane < nsrc ? src[g * 64 + lane] : -1;
    double r[ACC_V];
#pragma unroll
    for (int i = 0; i < ACC_V; i++) {
      const int64_t vi = v0 + i < V ? v0 + i : V - 1;
      r[i] = delta[((int64_t)g * V + vi) * 64 + lane];
    }
#pragma unroll
    for (int i = 0; i < ACC_V; i++) tile[i][lane] = r[i];
    // the tile is this wave's alone: order its LDS writes before the reads
    // of other lanes' values, and those reads before the next group's writes
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = 0; k < 64; k++) {
      const int32_t sk = __shfl(sj, k, GB_WAVE);  // -1 past nsrc
      if (mine && sk >= 0 && sk != (int32_t)v) acc += tile[lane][k];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (mine) bc[v] = acc;
}

struct MsWork {
  DevBuf sigma, delta, visited, fm[2], pool_v, pool_m, loff, pq[2], cand, counts, mode, loop, src,
      act, hub_in, hub_out, hub_meta, part;
  void release() {
    for (DevBuf *b : {&sigma, &delta, &visited, &fm[0], &fm[1], &pool_v, &pool_m, &loff, &pq[0],
                      &pq[1], &cand, &counts, &mode, &loop, &src, &act, &hub_in, &hub_out,
                      &hub_meta, &part})
      b->release();
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

// Grows a [G][old] array to [G][want] entries of `elem` bytes, keeping each
// group's first `old` entries.
gb_status grow_rows(DevBuf &buf, size_t elem, int32_t G, int64_t old, int64_t want,
                    hipStream_t st) {
  DevBuf nb;
  GB_TRY(nb.ensure(elem * (size_t)G * (size_t)want));
  GB_HIP(hipMemcpy2DAsync(nb.p, elem * want, buf.p, elem * old, elem * old, G,
                          hipMemcpyDeviceToDevice, st));
  GB_HIP(hipStreamSynchronize(st));
  std::swap(buf.p, nb.p);
  std::swap(buf.bytes, nb.bytes);
  return GB_OK;  // nb frees the old buffer
}

// One replayable chunk of LPR levels: forward or backward.
struct LevelGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~LevelGraph() { reset(); }
};

template <class F>
gb_status capture(LevelGraph &lg, hipStream_t cs, F launches) {
  lg.reset();
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    launches();
    e = hipStreamEndCapture(cs, &lg.graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&lg.exec, lg.graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return gb_fail(GB_ERR_HIP, "bc: graph capture failed: %s", hipGetErrorString(e));
  return GB_OK;
}

// A launch shape: grid blocks per group (x) as blocks per CU, and the
// push/pull ratio alpha.
struct BcShape {
  int per_cu, alpha;
};

// The rules when the autotuner does not run: 8 blocks per CU per group, at
// least 16 and at most 32768 blocks over all groups; push while 16 times the
// frontier's out-edges stay below the unvisited vertices' in-edges.
int bc_grid(const BcShape &sh, int32_t G) {
  return std::max(16, std::min(gb_num_cus() * sh.per_cu, 32768 / G));
}
constexpr BcShape BC_RULES{8, 16};

// Runs one batch of up to 64*G sources.  with_sigma: Brandes (sigma, delta,
// accumulate into d_bc); otherwise BFS levels into d_level [nsrc][V].
gb_status ms_run(gb_graph g, const int32_t *h_src, int32_t nsrc, bool with_sigma, double *d_bc,
                 int32_t *d_level, hipStream_t st, int32_t *depth_out, const BcShape &sh) {
  MsWork &w = msw();
  const int32_t V = g->V;
  const int32_t G = (nsrc + 63) / 64;
  const int64_t GV = (int64_t)G * V;
  BcCtx c{};
  c.V = V;
  c.G = G;
  c.E = g->E;
  c.rp = g->rp ? g->rp : g->rrp;
  c.col = g->rp ? g->col : g->rcol;
  c.rrp = g->rrp ? g->rrp : g->rp;
  c.rcol = g->rrp ? g->rcol : g->col;
  c.keep = with_sigma ? 1 : 0;
  c.hub = hub_edges();
  c.bch = bc_bchunk(c.hub);
  c.alpha = sh.alpha;
  if (with_sigma) {
    GB_TRY(w.sigma.ensure(sizeof(double) * GV * 64));
    GB_TRY(w.delta.ensure(sizeof(double) * GV * 64));
    GB_HIP(hipMemsetAsync(w.sigma.p, 0, sizeof(double) * GV * 64, st));
    GB_HIP(hipMemsetAsync(w.delta.p, 0, sizeof(double) * GV * 64, st));
    c.sigma = w.sigma.as<double>();
    c.delta = w.delta.as<double>();
  }
  c.level_out = d_level;
  GB_TRY(w.visited.ensure(sizeof(uint64_t) * GV));
  GB_HIP(hipMemsetAsync(w.visited.p, 0, sizeof(uint64_t) * GV, st));
  c.visited = w.visited.as<unsigned long long>();
  for (int k = 0; k < 2; k++) {
    GB_TRY(w.fm[k].ensure(sizeof(uint64_t) * GV));
    GB_HIP(hipMemsetAsync(w.fm[k].p, 0, sizeof(uint64_t) * GV, st));
    c.fm[k] = w.fm[k].as<unsigned long long>();
  }
  // the pool: two levels (BFS) or every level, starting at 6 V entries per group
  int64_t pool_cap = with_sigma ? 6 * (int64_t)V + 64 : 2 * (int64_t)V;
  int32_t lvl_cap = with_sigma ? 64 : 2;
  GB_TRY(w.pool_v.ensure(sizeof(int32_t) * (size_t)G * pool_cap));
  GB_TRY(w.pool_m.ensure(sizeof(uint64_t) * (size_t)G * pool_cap));
  GB_TRY(w.loff.ensure(sizeof(int64_t) * (size_t)G * lvl_cap));
  GB_HIP(hipMemsetAsync(w.loff.p, 0, sizeof(int64_t) * (size_t)G * lvl_cap, st));
  c.pq_cap = (int64_t)V + g->E / PUSH_CH + 1;
  for (int k = 0; k < 2; k++) {
    GB_TRY(w.pq[k].ensure(sizeof(PushItem) * (size_t)G * c.pq_cap));
    c.pq[k] = w.pq[k].as<PushItem>();
  }
  GB_TRY(w.cand.ensure(sizeof(int32_t) * GV));
  c.cand = w.cand.as<int32_t>();
  GB_TRY(w.counts.ensure(sizeof(unsigned long long) * CN * G));
  c.cnt = w.counts.as<unsigned long long>();
  GB_TRY(w.mode.ensure(sizeof(int32_t) * G));
  c.mode = w.mode.as<int32_t>();
  GB_TRY(w.loop.ensure(sizeof(BcLoop)));
  c.st = w.loop.as<BcLoop>();
  GB_TRY(w.src.ensure(sizeof(int32_t) * nsrc));
  GB_TRY(w.act.ensure(sizeof(uint64_t) * G));
  c.act = w.act.as<unsigned long long>();
  GB_HIP(hipMemcpyAsync(w.src.p, h_src, sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
  {
    std::vector<uint64_t> act(G);
    std::vector<unsigned long long> cnt0((size_t)CN * G, 0ull);
    for (int32_t k = 0; k < G; k++) {
      const int32_t n = std::min(64, nsrc - 64 * k);
      act[k] = n == 64 ? ~0ull : ((1ull << n) - 1);
      cnt0[(size_t)k * CN + C_MPULL] = (unsigned long long)g->E;  // every in-edge
    }
    GB_HIP(hipMemcpyAsync(w.act.p, act.data(), sizeof(uint64_t) * G, hipMemcpyHostToDevice, st));
    GB_HIP(hipMemcpyAsync(c.cnt, cnt0.data(), sizeof(unsigned long long) * CN * G,
                          hipMemcpyHostToDevice, st));
    BcLoop l{};
    l.d = -1;
    GB_HIP(hipMemcpyAsync(c.st, &l, sizeof(l), hipMemcpyHostToDevice, st));
    // the host copies read stack memory: wait before leaving the scope
    GB_HIP(hipStreamSynchronize(st));
  }
  // vertices above `hub` edges are pulled in chunks of `hub` edges, one wave
  // each (in-degree for the forward pulls, out-degree for the backward ones)
  const int vgrid = std::max(1, std::min((V + 255) / 256, 2048));
  GB_TRY(w.hub_in.ensure(sizeof(int32_t) * (V + 1)));
  GB_TRY(w.hub_out.ensure(sizeof(int32_t) * (V + 1)));
  GB_TRY(w.hub_meta.ensure(sizeof(unsigned long long) * 4));
  GB_HIP(hipMemsetAsync(w.hub_meta.p, 0, sizeof(unsigned long long) * 4, st));
  unsigned long long *hm = w.hub_meta.as<unsigned long long>();
  hipLaunchKernelGGL(mbc_hub_list_kernel, dim3(vgrid), dim3(256), 0, st, V, c.rrp, c.hub,
                     w.hub_in.as<int32_t>(), hm + 0, (unsigned int *)(hm + 1));
  hipLaunchKernelGGL(mbc_hub_list_kernel, dim3(vgrid), dim3(256), 0, st, V, c.rp, c.bch,
                     w.hub_out.as<int32_t>(), hm + 2, (unsigned int *)(hm + 3));
  unsigned long long hmh[4];
  GB_HIP(hipMemcpyAsync(hmh, hm, sizeof(hmh), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  c.hub_in = w.hub_in.as<int32_t>();
  c.nhub_in = (int32_t)hmh[0];
  c.maxch_in = std::max<int32_t>(1, ((int32_t)(uint32_t)hmh[1] + c.hub - 1) / c.hub);
  c.hub_out = w.hub_out.as<int32_t>();
  c.nhub_out = with_sigma ? (int32_t)hmh[2] : 0;
  c.maxch_out = std::max<int32_t>(1, ((int32_t)(uint32_t)hmh[3] + c.bch - 1) / c.bch);
  if (with_sigma && c.nhub_out > 0) {
    GB_TRY(w.part.ensure(sizeof(double) * 64 * (size_t)G * c.nhub_out * c.maxch_out));
    c.part = w.part.as<double>();
  }
  auto set_pool = [&]() {
    c.pool_v = w.pool_v.as<int32_t>();
    c.pool_m = w.pool_m.as<unsigned long long>();
    c.pool_cap = pool_cap;
    c.loff = w.loff.as<int64_t>();
    c.lvl_cap = lvl_cap;
  };
  set_pool();
  // level 0 and its finalize (as the end of level -1)
  const int gx = bc_grid(sh, G);
  const dim3 grid(gx, G), block(MB);
  hipLaunchKernelGGL(mbc_seed_kernel, dim3(G), dim3(64), 0, st, c, w.src.as<int32_t>(), nsrc);
  hipLaunchKernelGGL(mbc_fin_kernel, grid, block, 0, st, c);
  GB_HIP(hipGetLastError());
  hipStream_t cs = st, own = nullptr;
  if (!st) {  // the null stream cannot be captured
    GB_HIP(hipStreamCreate(&own));
    cs = own;
  }
  struct OwnStream {
    hipStream_t s;
    ~OwnStream() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } own_guard{own};
  if (own) {
    hipEvent_t ready;
    GB_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    GB_HIP(hipEventRecord(ready, st));
    GB_HIP(hipStreamWaitEvent(cs, ready, 0));
    (void)hipEventDestroy(ready);
  }
  auto fwd_launches = [&]() {
    for (int k = 0; k < LPR; k++) {
      hipLaunchKernelGGL(mbc_fwd_ctl_kernel, dim3(1), dim3(64), 0, cs, c);
      hipLaunchKernelGGL(mbc_scan_kernel, grid, block, 0, cs, c);
      if (with_sigma) {
        hipLaunchKernelGGL(mbc_push_kernel<true>, grid, block, 0, cs, c);
        hipLaunchKernelGGL(mbc_pull_kernel<true>, grid, block, 0, cs, c);
        hipLaunchKernelGGL(mbc_pull_hub_kernel<true>, grid, block, 0, cs, c);
      } else {
        hipLaunchKernelGGL(mbc_push_kernel<false>, grid, block, 0, cs, c);
        hipLaunchKernelGGL(mbc_pull_kernel<false>, grid, block, 0, cs, c);
        hipLaunchKernelGGL(mbc_pull_hub_kernel<false>, grid, block, 0, cs, c);
      }
      hipLaunchKernelGGL(mbc_fin_kernel, grid, block, 0, cs, c);
    }
  };
  LevelGraph fg;
  GB_TRY(capture(fg, cs, fwd_launches));
  BcLoop hl{};
  int32_t last_d = -2;
  for (;;) {
    GB_HIP(hipGraphLaunch(fg.exec, cs));
    GB_HIP(hipMemcpyAsync(&hl, c.st, sizeof(hl), hipMemcpyDeviceToHost, cs));
    GB_HIP(hipStreamSynchronize(cs));
    if (hl.done) break;
    if (hl.stall) {
      // room for one more level of V entries per group, and for the level table
      int64_t used = 0;
      std::vector<int64_t> lo((size_t)G * lvl_cap);
      std::vector<unsigned long long> cn((size_t)CN * G);
      GB_HIP(hipMemcpy(lo.data(), c.loff, sizeof(int64_t) * lo.size(), hipMemcpyDeviceToHost));
      GB_HIP(hipMemcpy(cn.data(), c.cnt, sizeof(unsigned long long) * cn.size(),
                       hipMemcpyDeviceToHost));
      const int dn = hl.d + 1;
      for (int32_t k = 0; k < G; k++)
        used = std::max<int64_t>(used, lo[(size_t)k * lvl_cap + dn] +
                                           (int64_t)cn[(size_t)k * CN + C_NEXT]);
      if (used + V > pool_cap) {
        const int64_t ncap = std::max<int64_t>(2 * pool_cap, used + 2 * (int64_t)V);
        GB_TRY(grow_rows(w.pool_v, sizeof(int32_t), G, pool_cap, ncap, cs));
        GB_TRY(grow_rows(w.pool_m, sizeof(uint64_t), G, pool_cap, ncap, cs));
        pool_cap = ncap;
      }
      if (dn + 1 >= lvl_cap) {
        GB_TRY(grow_rows(w.loff, sizeof(int64_t), G, lvl_cap, 2 * lvl_cap, cs));
        lvl_cap *= 2;
      }
      set_pool();
      hl.stall = 0;
      GB_HIP(hipMemcpyAsync(c.st, &hl, sizeof(hl), hipMemcpyHostToDevice, cs));
      GB_HIP(hipStreamSynchronize(cs));
      GB_TRY(capture(fg, cs, fwd_launches));
      continue;
    }
    if (hl.d == last_d) return gb_fail(GB_ERR_HIP, "bc: level loop made no progress");
    last_d = hl.d;
  }
  const int depth = hl.depth;  // levels 0..depth-1 are non-empty
  if (depth_out) *depth_out = depth;
  if (!with_sigma) return GB_OK;
  if (depth >= 2) {
    // levels depth-1 .. 0 (the forward pass left both mask arrays zero): clear
    // level bd + 2, write level bd, process; the deepest level only turns
    // sigma into psi
    hl.bd = depth;
    hl.bdone = 0;
    GB_HIP(hipMemcpyAsync(c.st, &hl, sizeof(hl), hipMemcpyHostToDevice, cs));
    GB_HIP(hipStreamSynchronize(cs));
    LevelGraph bg;
    GB_TRY(capture(bg, cs, [&]() {
      for (int k = 0; k < LPR; k++) {
        hipLaunchKernelGGL(mbc_bwd_ctl_kernel, dim3(1), dim3(1), 0, cs, c);
        hipLaunchKernelGGL(mbc_bwd_masks_kernel<false>, grid, block, 0, cs, c, 2);
        hipLaunchKernelGGL(mbc_bwd_masks_kernel<true>, grid, block, 0, cs, c, 0);
        hipLaunchKernelGGL(mbc_bwd_kernel, grid, block, 0, cs, c);
        hipLaunchKernelGGL(mbc_bwd_hub_kernel, grid, block, 0, cs, c);
        hipLaunchKernelGGL(mbc_bwd_combine_kernel, grid, block, 0, cs, c);
      }
    }));
    for (;;) {
      GB_HIP(hipGraphLaunch(bg.exec, cs));
      GB_HIP(hipMemcpyAsync(&hl, c.st, sizeof(hl), hipMemcpyDeviceToHost, cs));
      GB_HIP(hipStreamSynchronize(cs));
      if (hl.bdone) break;
    }
  }
  hipLaunchKernelGGL(mbc_accum_kernel,
                     dim3((V + ACC_V * (ACC_BLOCK / GB_WAVE) - 1) / (ACC_V * (ACC_BLOCK / GB_WAVE))),
                     dim3(ACC_BLOCK), 0, cs, V, G, nsrc,
                     w.src.as<int32_t>(), c.delta, d_bc);
  GB_HIP(hipGetLastError());
  if (own) GB_HIP(hipStreamSynchronize(cs));
  return GB_OK;
}

__global__ void fill_i32_kernel(int64_t n, int32_t *__restrict__ a, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// The shape for a launch of G groups on graph g: GB_BC_GRID (4, 8 or 16
// blocks per CU) and GB_BC_ALPHA (4 .. 64) fix parts; else the shape cached
// in the handle for G groups, else the rules.
const gb_graph_s::BcTuned *bc_cached(gb_graph g, int32_t G) {
  for (const auto &t : g->bc_tuned)
    if (t.groups == G) return &t;
  return nullptr;
}
BcShape bc_shape(gb_graph g, int32_t G, int *fixed_parts) {
  BcShape sh = BC_RULES;
  if (const auto *t = bc_cached(g, G)) sh = BcShape{t->per_cu, t->alpha};
  int fixed = 0;
  const char *e = getenv("GB_BC_GRID");
  if (e && (atoi(e) == 4 || atoi(e) == 8 || atoi(e) == 16)) {
    sh.per_cu = atoi(e);
    fixed |= 1;
  }
  e = getenv("GB_BC_ALPHA");
  if (e && atoi(e) >= 4 && atoi(e) <= 64) {
    sh.alpha = atoi(e);
    fixed |= 2;
  }
  if (fixed_parts) *fixed_parts = fixed;
  return sh;
}

// Autotuner for the Brandes engine, run on a graph's first batch of G groups
// when the handle has no shape for G yet (GB_BC_TUNE=0: rules only): the
// batch is solved in the rules' shape, then with one change at a time (the
// other blocks-per-CU counts, then alpha halved and doubled), into a scratch
// score array; the fastest shape is kept in the handle.  Every shape gives
// the same scores.
gb_status bc_tune(gb_graph g, const int32_t *h_src, int32_t nsrc, hipStream_t st) {
  const int32_t G = (nsrc + 63) / 64;
  int fixed = 0;
  const BcShape base = bc_shape(g, G, &fixed);
  const char *e = getenv("GB_BC_TUNE");
  if ((e && e[0] == '0') || bc_cached(g, G) || fixed == 3) return GB_OK;
  DevBuf scratch;
  GB_TRY(scratch.ensure(sizeof(double) * (g->V + 1)));
  auto timed = [&](const BcShape &s, double &ms) -> gb_status {
    GB_HIP(hipMemsetAsync(scratch.p, 0, sizeof(double) * g->V, st));
    GB_HIP(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    int32_t depth = 0;
    GB_TRY(ms_run(g, h_src, nsrc, true, scratch.as<double>(), nullptr, st, &depth, s));
    GB_HIP(hipStreamSynchronize(st));
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GB_OK;
  };
  BcShape best = base;
  double best_ms = 0;
  GB_TRY(timed(best, best_ms));
  auto consider = [&](BcShape s) -> gb_status {
    double ms = 0;
    GB_TRY(timed(s, ms));
    if (ms < best_ms) {
      best = s;
      best_ms = ms;
    }
    return GB_OK;
  };
  if (!(fixed & 1)) {
    const int pc0 = best.per_cu;
    for (int pc : {4, 8, 16})
      if (pc != pc0) GB_TRY(consider(BcShape{pc, best.alpha}));
  }
  if (!(fixed & 2)) {
    const int a0 = best.alpha;
    for (int a : {a0 / 2, a0 * 2})
      if (a >= 4 && a <= 64) GB_TRY(consider(BcShape{best.per_cu, a}));
  }
  g->bc_tuned.push_back(gb_graph_s::BcTuned{G, best.per_cu, best.alpha});
  return GB_OK;
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
    GB_TRY(ms_run(g, h_sources + b0, ns, false, nullptr, d_level + (int64_t)b0 * V, st, nullptr,
                  bc_shape(g, (ns + 63) / 64, nullptr)));
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
  for (int32_t b0 = 0; b0 < nsrc && V > 0; b0 += (int32_t)per) {
    const int32_t ns = (int32_t)std::min<int64_t>(per, nsrc - b0);
    int32_t depth = 0;
    GB_TRY(bc_tune(g, h_sources + b0, ns, st));
    GB_TRY(ms_run(g, h_sources + b0, ns, true, d_bc, nullptr, st, &depth,
                  bc_shape(g, (ns + 63) / 64, nullptr)));
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
