// gb_msssp.hip -- delta-stepping for many sources of one graph at once.
//
// The per-source engine (gb_sssp.hip) keeps one queue entry per (source,
// vertex) and reads a vertex's edges once per source.  Here up to 64 sources
// form a group and share everything but their distances:
//   * distances are vertex-major, dist[v][W] (W = group width, 8..64), so the
//     W sources' values of one vertex are one or two cache lines;
//   * a queue entry is a 32/64-edge chunk of a vertex (vertex, edge range and
//     the mask of sources for which it must be expanded, 32 bytes);
//   * a wave takes one entry and loads its col/w with one coalesced load; its
//     lanes are then (edge slot, source) pairs, 64 / W edges per step, the
//     slot's target and weight taken from the loading lane by a cross-lane
//     read, so the W relaxations of one target are one coalesced line;
//   * a group of steps is issued together: all target reads, then the
//     atomicMin of every improving pair and one atomicOr per (step, slot) on
//     the targets' next-phase masks, issued by one lane each in a single
//     instruction; neither returns a value (the decision uses the value read),
//     so a group costs one memory round trip whatever its size, and the next
//     entry's loads are in flight meanwhile;
//   * an improved (target, source) pair sets the source's bit in the target's
//     next-phase mask (same bucket) or far mask (later bucket); a far mask
//     turned from empty to non-empty appends the target to the far queue;
//   * between two phases a compaction pass scans the next-phase masks of all
//     vertices (coalesced), zeroes them and writes one queue entry per 32/64-
//     edge chunk of every marked vertex with out-edges, carrying the vertex's
//     source mask, so the step kernel never waits for an atomic's return or
//     reads row bounds inside its inner loop (the scan form, SCAN = true).
// The scan costs a pass over V per phase.  Small graphs with many short
// phases use the append form instead (SCAN = false): the lane whose atomicOr
// turns a target's next-phase mask from empty to non-empty (its return value)
// reads the target's row bounds and appends its chunks at once; the masks
// are double-buffered (the phase reads its input mask array, which is
// cleared by queue entry after the phase) and entries carry no mask.
// All sources step through bucket indices in lockstep: light phases while the
// near queue is non-empty, then a far scan that moves the next non-empty
// bucket's (vertex, source) pairs to the near queue.  There is no light/heavy
// split: an expansion relaxes all edges of the vertex.  Delta-stepping is
// exact for any bucket width, so distances equal the per-source engine's.
// The bucket loop is the state machine ms_ctl_advance (gb_sssp_ctl.hpp),
// replayed from a captured hipGraph as in the per-source engine.  A phase of
// the scan form is three launches (step, compaction, state advance); of the
// append form two (the state advance fused with the clear of the previous
// phase's input masks, then the step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"
#include "gb_sssp_ctl.hpp"

namespace {

constexpr int MS_BLOCK = 256;
// Default edges per queue entry: 32 for groups of up to 16 sources (many
// edge slots per step: a short chunk keeps the per-entry critical path short),
// 64 for wider groups.  The chunk of a solve (32 or 64, MsCtx::cshift) is a
// launch parameter the autotuner may change; the unrolling below follows the
// default.
template <int W>
struct MsShape {
  static constexpr int CHUNK_SHIFT = W <= 16 ? 5 : 6;
  static constexpr int32_t CHUNK = 1 << CHUNK_SHIFT;
  static constexpr int ES = GB_WAVE / W;                // edge slots per step
  static constexpr int STEPS = CHUNK / ES;              // steps per full chunk
  static constexpr int UNROLL = STEPS < 4 ? STEPS : 4;  // steps per group
  static constexpr int G = ES * UNROLL;                 // edges per group (<= 64)
  // waves per SIMD the step kernel's registers must allow (1: no bound)
  static constexpr int MIN_WAVES = W >= 16 ? 6 : 1;
};
constexpr int32_t MS_MIN_CHUNK = 32;  // for queue capacities
constexpr int MS_NCAP = 128, MS_FCAP = 128;  // LDS append buffers per wave (entries)

// near-queue entry: one chunk [beg, end) of vertex v's edges and (scan form)
// the mask of sources for which it is expanded
struct __align__(16) MsEnt {
  unsigned long long m;
  uint32_t v;
  int32_t beg, end;
  uint32_t pad;
};

__device__ __forceinline__ uint32_t ms_nchunks(int32_t deg, int cshift) {
  const uint32_t k = (uint32_t)((deg + (1 << cshift) - 1) >> cshift);
  return k ? k : 1u;
}

// chunk q of the vertex whose edges are [r0, r1)
__device__ __forceinline__ MsEnt ms_chunk(unsigned long long m, uint32_t v, int32_t r0, int32_t r1,
                                          uint32_t q, int cshift) {
  const int32_t b = r0 + (int32_t)(q << cshift);
  return MsEnt{m, v, b, min(r1, b + (1 << cshift)), 0u};
}

struct MsCtx {
  const int32_t *rp;
  const int32_t *col;
  const uint32_t *w;
  int32_t V;
  int64_t E;
  int32_t ntask;                   // sources in this group (<= W)
  uint32_t *dist;                  // [V][W]
  // next-phase masks: scan form mask[0] only (zero between phases), append
  // form one per near queue
  unsigned long long *mask[2];
  unsigned long long *far;         // far masks
  MsEnt *qn[2];                    // near queues 0/1 (chunks)
  uint64_t *qf[2];                 // far queues 0/1 (vertices)
  unsigned long long *cnt;         // MS_* counters
  unsigned long long *wstat;       // per-wave statistics slots (GB_WSTAT per wave)
  uint64_t delta, magic;           // bucket width and its division constant
  int32_t cshift;                  // log2 of the edges per queue entry (5 or 6)
};

__device__ __forceinline__ uint64_t ms_bucket(uint32_t d, uint64_t magic) {
  return magic ? __umul64hi((uint64_t)d, magic) : (uint64_t)d;
}

__global__ void ms_fill_kernel(int64_t n, uint32_t *__restrict__ dist) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dist[i] = 0xFFFFFFFFu;
}

// One wave (blockDim 64): lane t seeds source t; the lowest lane holding a
// vertex writes its chunks with the mask of every lane holding it.
template <int W, bool SCAN>
__global__ void ms_seed_kernel(MsCtx c, const int32_t *__restrict__ src) {
  const int t = threadIdx.x;
  const bool on = t < c.ntask;
  const int32_t s = on ? src[t] : -1;
  unsigned long long m = 0;
  bool lowest = on;
  for (int j = 0; j < GB_WAVE; j++) {
    const int32_t sj = __shfl(s, j, GB_WAVE);
    if (on && sj == s) {
      m |= 1ull << j;
      if (j < t) lowest = false;
    }
  }
  if (!on) return;
  c.dist[(int64_t)s * W + t] = 0;
  if (!SCAN) atomicOr(&c.mask[0][s], 1ull << t);
  const int32_t r0 = c.rp[s], r1 = c.rp[s + 1];
  if (lowest && r1 > r0) {
    const uint32_t k = ms_nchunks(r1 - r0, c.cshift);
    const unsigned long long b = atomicAdd(&c.cnt[MS_NEAR0], (unsigned long long)k);
    for (uint32_t i = 0; i < k; i++) c.qn[0][b + i] = ms_chunk(m, (uint32_t)s, r0, r1, i, c.cshift);
  }
}

// What a wave loads for one near entry: the entry's source mask, this lane's
// source distance at the vertex, and the chunk's edges (lane r holds edge
// beg + r).
struct MsIn {
  unsigned long long m;
  uint32_t du;
  int32_t x;
  uint32_t w;
};

template <int W, bool SCAN>
__device__ __forceinline__ MsIn ms_load(const MsCtx &c, const unsigned long long *min_,
                                        const MsEnt &e, int lane, int t) {
  MsIn r;
  r.m = SCAN ? e.m : min_[e.v];
  r.du = t < c.ntask ? c.dist[(int64_t)e.v * W + t] : 0xFFFFFFFFu;
  const bool in = lane < e.end - e.beg;
  r.x = in ? c.col[e.beg + lane] : 0;
  r.w = in ? c.w[e.beg + lane] : 0u;
  return r;
}

// Light phase: expand the entries of near queue a_in.  Entries are software
// pipelined: entry i is expanded while entry i + nwaves's loads and entry
// i + 2 nwaves's descriptor are in flight.
template <int W, bool SCAN>
__device__ __forceinline__ void ms_light(const MsCtx &c, const MsCtl &a, MsEnt *lds_near,
                                         uint64_t *lds_far, uint32_t &relax, uint32_t &imp_near,
                                         uint32_t &imp_far, uint32_t &expanded, bool &ovf) {
  using S = MsShape<W>;
  constexpr int ES = S::ES, UNROLL = S::UNROLL, G = S::G;
  const int lane = gb_lane();
  const int es = lane / W, t = lane % W;
  // the lane that issues the mask update of (step lane / ES, slot lane % ES)
  const int kj = lane / ES, ej = lane % ES;
  const uint64_t wmask = W == 64 ? ~0ull : ((1ull << W) - 1);
  const MsEnt *in = c.qn[a.a_in - MS_NEAR0];
  const unsigned long long *min_ = c.mask[a.a_in - MS_NEAR0];
  unsigned long long *mout = SCAN ? c.mask[0] : c.mask[a.a_near - MS_NEAR0];
  gb_wave_queue<MsEnt, MS_NCAP> nq(lds_near, c.qn[a.a_near - MS_NEAR0], &c.cnt[a.a_near]);
  gb_wave_queue<uint64_t, MS_FCAP> fq(lds_far, c.qf[a.a_far - MS_FAR0], &c.cnt[a.a_far]);
  const uint64_t n = c.cnt[a.a_in];
  const uint64_t b = a.a_bucket;
  // wave-uniform indices (readfirstlane): the entries are read with scalar
  // loads and live in scalar registers
  const uint64_t wave = __builtin_amdgcn_readfirstlane(
      (uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t i = wave, i1 = wave + nwaves;
  MsEnt ecur{}, enxt{};
  MsIn cur{};
  if (i < n) {
    ecur = in[i];
    cur = ms_load<W, SCAN>(c, min_, ecur, lane, t);
  }
  if (i1 < n) enxt = in[i1];
  while (i < n) {
    MsIn nxt{};
    MsEnt enn{};
    const uint64_t i2 = i1 + nwaves;
    if (i1 < n) nxt = ms_load<W, SCAN>(c, min_, enxt, lane, t);
    if (i2 < n) enn = in[i2];
    const bool act = t < c.ntask && ((cur.m >> t) & 1ull);
    const int cnt = ecur.end - ecur.beg;
    expanded += lane == 0 ? (uint32_t)cnt : 0u;
    for (int g0 = 0; g0 < cnt; g0 += G) {
      int32_t x[UNROLL];
      uint32_t nd[UNROLL], old[UNROLL];
      bool ok[UNROLL];
#pragma unroll
      for (int k = 0; k < UNROLL; k++) {
        const int r = g0 + k * ES + es;  // < chunk <= 64 (the chunk is a multiple of G)
        x[k] = __shfl(cur.x, r, GB_WAVE);
        const uint32_t wk = __shfl(cur.w, r, GB_WAVE);
        // branch-free, so that the reads below stay one batch
        ok[k] = act & (r < cnt);
        const uint64_t s = (uint64_t)cur.du + wk;
        const bool big = s >= 0xFFFFFFFFull;
        ovf |= ok[k] & big;
        nd[k] = big ? 0xFFFFFFFEu : (uint32_t)s;
        relax += ok[k];
      }
      // unconditional: x[k] is a vertex for every lane (0 past the chunk's
      // end) and t < W, so the read is in bounds; a conditional read would be
      // compiled to a branch per step that waits for its own load
#pragma unroll
      for (int k = 0; k < UNROLL; k++) old[k] = c.dist[(int64_t)x[k] * W + t];
      // lane j < G updates the masks of edge g0 + j = (step j / ES, slot j % ES)
      uint64_t nm = 0, fm = 0;
#pragma unroll
      for (int k = 0; k < UNROLL; k++) {
        ok[k] = ok[k] & (nd[k] < old[k]);
        const bool isnear = ms_bucket(nd[k], c.magic) <= b;
        const bool near = ok[k] & isnear;
        const bool far = ok[k] & !isnear;
        imp_near += near;
        imp_far += far;
        const uint64_t nbk = __ballot(near), fbk = __ballot(far);
        nm = kj == k ? nbk : nm;
        fm = kj == k ? fbk : fm;
      }
      // the atomics go after every read of the group: a wait for a later read
      // would otherwise also wait for the atomics issued before it
#pragma unroll
      for (int k = 0; k < UNROLL; k++)
        if (ok[k]) atomicMin(&c.dist[(int64_t)x[k] * W + t], nd[k]);
      nm = (nm >> (ej * W)) & wmask;
      fm = (fm >> (ej * W)) & wmask;
      const uint32_t xj = (uint32_t)__shfl(cur.x, (g0 + lane) & (GB_WAVE - 1), GB_WAVE);
      if (SCAN) {
        // no return value: the compaction pass after the phase queues the
        // marked targets
        if (lane < G && nm) atomicOr(&mout[xj], (unsigned long long)nm);
        // far marks are rarer: the lane that starts a far mask queues the target
        if (__ballot(lane < G && fm) != 0ull) {
          unsigned long long of = 1ull;
          if (lane < G && fm) of = atomicOr(&c.far[xj], (unsigned long long)fm);
          fq.push(of == 0ull ? 1u : 0u, [&](uint32_t) { return (uint64_t)xj; });
        }
      } else {
        unsigned long long on = 1ull, of = 1ull;
        if (lane < G && nm) on = atomicOr(&mout[xj], (unsigned long long)nm);
        if (lane < G && fm) of = atomicOr(&c.far[xj], (unsigned long long)fm);
        // a target without out-edges is not queued: its mask bits stay set,
        // which only suppresses later appends that would expand nothing
        int32_t r0 = 0, r1 = 0;
        const bool newn = on == 0ull, newf = of == 0ull;
        if (newn || newf) {
          r0 = c.rp[xj];
          r1 = c.rp[xj + 1];
        }
        const uint32_t pn = newn && r1 > r0 ? ms_nchunks(r1 - r0, c.cshift) : 0u;
        nq.push(pn, [&](uint32_t q) { return ms_chunk(0ull, xj, r0, r1, q, c.cshift); });
        fq.push(newf && r1 > r0 ? 1u : 0u, [&](uint32_t) { return (uint64_t)xj; });
      }
    }
    ecur = enxt;
    cur = nxt;
    enxt = enn;
    i = i1;
    i1 = i2;
  }
  if (SCAN)
    gb_block_flush(fq);
  else
    gb_block_flush(nq, fq);
}

// Far scan: far queue a_in -> bucket a_bucket's pairs marked in the
// next-phase masks (queued by the compaction), later buckets stay (far queue
// a_far), settled ones are dropped.
template <int W, bool SCAN>
__device__ __forceinline__ void ms_far(const MsCtx &c, const MsCtl &a, MsEnt *lds_near,
                                       uint64_t *lds_far) {
  constexpr int ES = GB_WAVE / W;
  const int lane = gb_lane();
  const int es = lane / W, t = lane % W;
  const uint64_t wmask = W == 64 ? ~0ull : ((1ull << W) - 1);
  const uint64_t *in = c.qf[a.a_in - MS_FAR0];
  unsigned long long *mout = SCAN ? c.mask[0] : c.mask[a.a_near - MS_NEAR0];
  gb_wave_queue<MsEnt, MS_NCAP> nq(lds_near, c.qn[a.a_near - MS_NEAR0], &c.cnt[a.a_near]);
  gb_wave_queue<uint64_t, MS_FCAP> fq(lds_far, c.qf[a.a_far - MS_FAR0], &c.cnt[a.a_far]);
  const uint64_t n = c.cnt[a.a_in];
  const uint64_t bnext = a.a_bucket;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t minb = ~0ull;
  for (uint64_t i0 = wave * ES; i0 < n; i0 += nwaves * ES) {
    const uint64_t i = i0 + es;
    bool near = false, keep = false;
    uint32_t x = 0;
    if (i < n) {
      x = (uint32_t)in[i];
      const unsigned long long m = c.far[x];
      if ((m >> t) & 1ull) {
        const uint64_t bk = ms_bucket(c.dist[(int64_t)x * W + t], c.magic);
        if (bk == bnext) {
          near = true;
        } else if (bk > bnext) {
          keep = true;
          minb = bk < minb ? bk : minb;
        }
      }
    }
    const uint64_t nb = __ballot(near), kb = __ballot(keep);
    const uint64_t nm = (nb >> (es * W)) & wmask, km = (kb >> (es * W)) & wmask;
    uint32_t pn = 0, pf = 0;
    int32_t r0 = 0, r1 = 0;
    if (t == 0 && i < n) {
      c.far[x] = km;  // the entry is this vertex's only one in the scan
      if (SCAN) {
        if (nm) atomicOr(&mout[x], (unsigned long long)nm);
      } else if (nm && atomicOr(&mout[x], (unsigned long long)nm) == 0ull) {
        r0 = c.rp[x];
        r1 = c.rp[x + 1];
        pn = ms_nchunks(r1 - r0, c.cshift);
      }
      if (km) pf = 1;
    }
    if (SCAN) {
      fq.push(pf, [&](uint32_t) { return (uint64_t)x; });
    } else {
      nq.push(pn, [&](uint32_t q) { return ms_chunk(0ull, x, r0, r1, q, c.cshift); });
      fq.push(pf, [&](uint32_t) { return (uint64_t)x; });
    }
  }
  if (SCAN)
    gb_block_flush(fq);
  else
    gb_block_flush(nq, fq);
  gb_block_atomic_min(minb, ~0ull, &c.cnt[MS_MINB]);
}

template <int W, bool SCAN>
__global__ __launch_bounds__(MS_BLOCK)
__attribute__((amdgpu_waves_per_eu(MsShape<W>::MIN_WAVES))) void ms_step_kernel(MsCtx c,
                                                           const MsCtl *__restrict__ ctl) {
  const int act = ctl->action;
  if (act == ACT_NONE) return;
  const MsCtl a = *ctl;
  __shared__ uint64_t s_far[MS_BLOCK / GB_WAVE][MS_FCAP];
  __shared__ MsEnt s_near[MS_BLOCK / GB_WAVE][SCAN ? 1 : MS_NCAP];
  uint64_t *lf = s_far[threadIdx.x / GB_WAVE];
  MsEnt *ln = s_near[threadIdx.x / GB_WAVE];
  if (act == ACT_LIGHT) {
    uint32_t relax = 0, imp_near = 0, imp_far = 0, expanded = 0;
    bool ovf = false;
    ms_light<W, SCAN>(c, a, ln, lf, relax, imp_near, imp_far, expanded, ovf);
    uint32_t tot, tn, tf, te;
    (void)gb_wave_excl_sum(relax, &tot);
    (void)gb_wave_excl_sum(imp_near, &tn);
    (void)gb_wave_excl_sum(imp_far, &tf);
    (void)gb_wave_excl_sum(expanded, &te);
    if (gb_lane() == 0) {
      gb_wstat_add(c.wstat, 0, tot);
      gb_wstat_add(c.wstat, 1, tn);
      gb_wstat_add(c.wstat, 2, tf);
      gb_wstat_add(c.wstat, 3, te);
    }
    if (__ballot(ovf) && gb_lane() == 0) atomicOr(&c.cnt[MS_OVF], 1ull);
  } else {
    ms_far<W, SCAN>(c, a, ln, lf);
  }
}

// The state machine between two phases (one thread).
__global__ void ms_ctrl_kernel(MsCtx c, MsCtl *__restrict__ ctl) {
  MsCtl s = *ctl;
  ms_ctl_advance(s, c.cnt);
  *ctl = s;
}

// Append form, between two phases, one launch: thread 0 computes the next
// action from the previous state `prev` into `next`, and, when the previous
// action was a light phase, every thread clears that phase's input masks (the
// buffer is the output of the phase after next).  The two states are
// distinct buffers, alternating per phase, and the clear takes its entry
// count from `prev` rather than the counter, which the advance may reset.
__global__ void ms_clear_ctrl_kernel(MsCtx c, const MsCtl *__restrict__ prev,
                                     MsCtl *__restrict__ next) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MsCtl s = *prev;
    ms_ctl_advance(s, c.cnt);
    *next = s;
  }
  if (prev->action != ACT_LIGHT) return;
  const int in = prev->a_in;
  const uint64_t n = prev->a_in_n;
  unsigned long long *m = c.mask[in - MS_NEAR0];
  const MsEnt *q = c.qn[in - MS_NEAR0];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    m[q[i].v] = 0;
}

// After a phase: every vertex whose next-phase mask is set gets one near-queue
// entry per chunk of its edges, with the mask, into queue a_near; the mask is
// zeroed for the phase after.  Vertices without out-edges are dropped.  A block
// takes `range` consecutive vertices (a multiple of MS_CSPAN) in batches of
// MS_CSPAN, a wave 256 consecutive vertices of a batch, four per lane, their
// masks read together and coalesced.  Two passes over the range: the first
// counts its entries and reserves them with one atomic on the queue counter
// (per-wave reservations, one per 128 entries, serialise on the counter's
// cache line when most vertices are marked); the second, run only by blocks
// with entries, writes them at their offsets and clears the masks, re-reading
// the masks and row offsets the first pass has just brought into L2.  A
// wave's entries are laid out vertex slot by slot (j = 0..3), so lanes with
// one entry each store one contiguous run.
constexpr int MS_CPL = 4;  // vertices per lane and batch
constexpr int64_t MS_CSPAN = (int64_t)MS_BLOCK * MS_CPL;
template <int W>
__global__ __launch_bounds__(MS_BLOCK) void ms_compact_kernel(MsCtx c,
                                                              const MsCtl *__restrict__ ctl,
                                                              int64_t range) {
  const int act = ctl->action;
  if (act == ACT_NONE) return;
  const int out = ctl->a_near;
  constexpr int NW = MS_BLOCK / GB_WAVE;
  constexpr int64_t SPAN = (int64_t)GB_WAVE * MS_CPL;
  __shared__ uint32_t s_wt[NW];
  __shared__ unsigned long long s_base;
  __shared__ uint32_t s_tot;
  const int lane = gb_lane(), w = (int)(threadIdx.x >> 6);
  const int64_t lo = (int64_t)blockIdx.x * range;
  const int64_t hi = min((int64_t)c.V, lo + range);
  unsigned long long *mask = c.mask[0];
  // pass 1: count
  uint32_t cnt = 0;
  for (int64_t b0 = lo + w * SPAN; b0 < hi; b0 += MS_CSPAN) {
    unsigned long long m[MS_CPL];
    bool any = false;
#pragma unroll
    for (int j = 0; j < MS_CPL; j++) {
      const int64_t v = b0 + (int64_t)j * GB_WAVE + lane;
      m[j] = v < hi ? mask[v] : 0ull;
      any |= m[j] != 0ull;
    }
    if (__ballot(any) == 0ull) continue;
#pragma unroll
    for (int j = 0; j < MS_CPL; j++) {
      const int64_t v = b0 + (int64_t)j * GB_WAVE + lane;
      if (m[j]) {
        const int32_t r0 = c.rp[v], r1 = c.rp[v + 1];
        cnt += r1 > r0 ? ms_nchunks(r1 - r0, c.cshift) : 0u;
      }
    }
  }
  uint32_t wtot;
  (void)gb_wave_excl_sum(cnt, &wtot);
  if (lane == 0) s_wt[w] = wtot;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < NW; q++) t += s_wt[q];
    s_tot = t;
    s_base = t ? atomicAdd(&c.cnt[out], (unsigned long long)t) : 0ull;
  }
  __syncthreads();
  if (s_tot == 0) return;  // block-uniform
  // pass 2: write, in batch order, each batch's waves in wave order
  MsEnt *q = c.qn[out - MS_NEAR0];
  uint64_t off = s_base;
  for (int64_t bb = lo; bb < hi; bb += MS_CSPAN) {  // block-uniform
    const int64_t b0 = bb + w * SPAN;
    unsigned long long m[MS_CPL];
    int32_t r0[MS_CPL], r1[MS_CPL];
    uint32_t k[MS_CPL], ex[MS_CPL], tj[MS_CPL];
#pragma unroll
    for (int j = 0; j < MS_CPL; j++) {
      const int64_t v = b0 + (int64_t)j * GB_WAVE + lane;
      m[j] = v < hi ? mask[v] : 0ull;
    }
    uint32_t wsum = 0;
#pragma unroll
    for (int j = 0; j < MS_CPL; j++) {
      const int64_t v = b0 + (int64_t)j * GB_WAVE + lane;
      r0[j] = r1[j] = 0;
      if (m[j]) {
        mask[v] = 0ull;
        r0[j] = c.rp[v];
        r1[j] = c.rp[v + 1];
      }
      k[j] = r1[j] > r0[j] ? ms_nchunks(r1[j] - r0[j], c.cshift) : 0u;
      ex[j] = gb_wave_excl_sum(k[j], &tj[j]);
      wsum += tj[j];
    }
    if (lane == 0) s_wt[w] = wsum;
    __syncthreads();
    uint64_t p = off;
    uint32_t btot = 0;
    for (int qq = 0; qq < NW; qq++) {
      const uint32_t t = s_wt[qq];
      if (qq < w) p += t;
      btot += t;
    }
    __syncthreads();  // s_wt is rewritten by the next batch
#pragma unroll
    for (int j = 0; j < MS_CPL; j++) {
      const uint32_t v = (uint32_t)(b0 + (int64_t)j * GB_WAVE + lane);
      for (uint32_t i = 0; i < k[j]; i++)
        q[p + ex[j] + i] = ms_chunk(m[j], v, r0[j], r1[j], i, c.cshift);
      p += tj[j];
    }
    off += btot;
  }
}

// group distances [V][W] -> caller rows out[t][V] for t < ntask.  A block
// takes VT = 4096 / W vertices: it reads their W * VT distances as one
// contiguous run (every lane busy) into an LDS tile whose row stride VT + 1
// keeps the transposed writes on distinct banks, then writes each source's VT
// distances as one run.
template <int W>
__global__ __launch_bounds__(256) void ms_unpack_kernel(int64_t V, int ntask,
                                                         const uint32_t *__restrict__ dist,
                                                         uint32_t *__restrict__ out) {
  constexpr int VT = 4096 / W;
  __shared__ uint32_t tile[W][VT + 1];
  const int64_t v0 = (int64_t)blockIdx.x * VT;
  const int64_t rest = V - v0;
  const int nv = rest < VT ? (int)rest : VT;
  const uint32_t *src = dist + v0 * W;
  for (int i = threadIdx.x; i < W * VT; i += 256) {
    const int vv = i / W, tt = i % W;
    if (vv < nv) tile[tt][vv] = src[i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < W * VT; i += 256) {
    const int tt = i / VT, vv = i % VT;
    if (vv < nv && tt < ntask) out[(int64_t)tt * V + v0 + vv] = tile[tt][vv];
  }
}

// reached (source, vertex) pairs and their out-edges, over the caller's rows
// grid.y strides over source rows; one atomic pair per block (blockDim 256)
__global__ void ms_count_kernel(int64_t V, int32_t nsrc, const int32_t *__restrict__ rp,
                                const uint32_t *__restrict__ rows,
                                unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_r[4], s_e[4];
  unsigned long long reached = 0, edges = 0;
  for (int64_t t = blockIdx.y; t < nsrc; t += gridDim.y) {
    const uint32_t *row = rows + t * V;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
      if (row[v] != 0xFFFFFFFFu) {
        reached++;
        edges += (unsigned long long)(rp[v + 1] - rp[v]);
      }
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    reached += __shfl_xor(reached, d, 64);
    edges += __shfl_xor(edges, d, 64);
  }
  if (gb_lane() == 0) {
    s_r[threadIdx.x >> 6] = reached;
    s_e[threadIdx.x >> 6] = edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], s_r[0] + s_r[1] + s_r[2] + s_r[3]);
    atomicAdd(&out[1], s_e[0] + s_e[1] + s_e[2] + s_e[3]);
  }
}

// Step-kernel blocks per CU when the autotuner does not run: 4 below 2^23
// edges (small phases: fewer idle blocks to launch and flush), 8 up to 2^25,
// 16 beyond (more waves in flight in the large phases).  Measured at 4 / 8 /
// 16 blocks per CU: R-MAT 16 with 8 sources 3.3 / 3.7-3.9 / 4.7 ms; R-MAT 22
// with 32 sources 27.9 / 26.3-27.2 / 26.3-26.5 ms.
int ms_rule_per_cu(int64_t E) {
  return E < ((int64_t)1 << 23) ? 4 : E < ((int64_t)1 << 25) ? 8 : 16;
}
constexpr int MS_MAX_PER_CU = 16;

// The scan form pays a pass over V per phase, the append form a returning
// atomic and a row-bound read per newly marked target inside the step.
// Measured with 8 sources of R-MAT 16 (185 phases): append 3.3 ms, scan
// 4.3 ms; with 32 sources of R-MAT 24 (12 phases): append 61 ms, scan 54 ms.
// This is the rule when the autotuner does not run.
bool ms_rule_scan(int64_t V) { return V >= ((int64_t)1 << 20); }

// A solve's shape: form, step-kernel blocks per CU, log2 edges per entry.
struct MsShapeRt {
  bool scan;
  int per_cu;
  int cshift;
};

// The parts of the shape fixed by the environment (-1: free):
// GB_MSSSP_SCAN (0/1), GB_MSSSP_GRID (blocks per CU: 4, 8 or 16),
// GB_MSSSP_CHUNK (32 or 64); GB_MSSSP_TUNE=0 keeps the rules for the rest.
struct MsFixed {
  int scan = -1, per_cu = -1, cshift = -1;
  bool tune = true;
};
MsFixed ms_fixed() {
  MsFixed f;
  const char *e = getenv("GB_MSSSP_SCAN");
  if (e && e[0]) f.scan = e[0] != '0' ? 1 : 0;
  e = getenv("GB_MSSSP_GRID");
  if (e && (atoi(e) == 4 || atoi(e) == 8 || atoi(e) == 16)) f.per_cu = atoi(e);
  e = getenv("GB_MSSSP_CHUNK");
  if (e && (atoi(e) == 32 || atoi(e) == 64)) f.cshift = atoi(e) == 32 ? 5 : 6;
  e = getenv("GB_MSSSP_TUNE");
  if (e && e[0] == '0') f.tune = false;
  return f;
}

// Compaction blocks: one batch of MS_CPL vertices per lane where that takes
// at most 16 blocks per CU, else several batches per wave.
int ms_compact_grid(int64_t V) {
  const int64_t per_block = (int64_t)MS_BLOCK * MS_CPL;
  const int64_t want = (V + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)gb_num_cus() * 16));
}

struct MsWork {
  DevBuf dist, m0, m1, far, q[4], cnt, ctl, src, wstat;
  void release() {
    for (DevBuf *b : {&dist, &m0, &m1, &far, &cnt, &ctl, &src, &wstat}) b->release();
    for (auto &b : q) b.release();
  }
};
MsWork &g_msw_ref() {
  static MsWork *w = new MsWork();
  return *w;
}
#define g_msw (g_msw_ref())

// Counters of one group's solve.
struct MsStats {
  int64_t phases = 0, buckets = 0, imp_near = 0, imp_far = 0, expanded = 0;
  unsigned long long relax = 0;
  bool ovf = false;
  void add(const MsStats &o) {
    phases += o.phases;
    buckets += o.buckets;
    imp_near += o.imp_near;
    imp_far += o.imp_far;
    expanded += o.expanded;
    relax += o.relax;
    ovf = ovf || o.ovf;
  }
};

template <int W, bool SCAN>
gb_status ms_group(MsCtx c, const int32_t *d_src, int max_phases, int grid, hipStream_t st,
                   uint32_t *out_rows, MsStats &sx) {
  MsWork &w = g_msw;
  const int64_t V = c.V;
  hipLaunchKernelGGL(ms_fill_kernel, dim3((unsigned)std::min<int64_t>((V * W + 255) / 256, 65536)),
                     dim3(256), 0, st, V * W, c.dist);
  GB_HIP(hipMemsetAsync(c.mask[0], 0, sizeof(unsigned long long) * V, st));
  if (!SCAN) GB_HIP(hipMemsetAsync(c.mask[1], 0, sizeof(unsigned long long) * V, st));
  GB_HIP(hipMemsetAsync(c.far, 0, sizeof(unsigned long long) * V, st));
  GB_HIP(hipMemsetAsync(c.cnt, 0, sizeof(unsigned long long) * MS_N, st));
  const int64_t nwaves = (int64_t)grid * (MS_BLOCK / GB_WAVE);
  GB_HIP(hipMemsetAsync(c.wstat, 0, sizeof(unsigned long long) * GB_WSTAT * nwaves, st));
  hipLaunchKernelGGL((ms_seed_kernel<W, SCAN>), dim3(1), dim3(64), 0, st, c, d_src);
  GB_HIP(hipGetLastError());
  // the loop state.  Scan form: one state in ctl[0], advanced to the first
  // action here, then after every phase's compaction.  Append form: two
  // states alternating per phase: phase k acts on ctl[k % 2], which the launch
  // before it computed from ctl[(k + 1) % 2]; the first launch starts from the
  // initial state in ctl[1] (no action taken yet)
  MsCtl hc{};
  hc.max_phases = max_phases > 0 ? max_phases : ((int64_t)1 << 40);
  hc.state = MS_ST_LIGHT;
  hc.buckets = 1;
  hc.action = ACT_NONE;
  MsCtl *ctl = w.ctl.as<MsCtl>();
  MsCtl *hstate = SCAN ? ctl : ctl + 1;  // where a replay leaves the state
  GB_HIP(hipMemcpyAsync(hstate, &hc, sizeof(hc), hipMemcpyHostToDevice, st));
  if (SCAN) {
    hipLaunchKernelGGL(ms_ctrl_kernel, dim3(1), dim3(1), 0, st, c, ctl);
    GB_HIP(hipGetLastError());
  }
  const int cgrid = SCAN ? ms_compact_grid(c.V) : gb_num_cus() * 2;
  // vertices per compaction block: a multiple of MS_CSPAN, cgrid blocks cover V
  const int64_t crange = ((c.V + cgrid - 1) / cgrid + MS_CSPAN - 1) / MS_CSPAN * MS_CSPAN;
  constexpr int PH = 16;  // phases per replay (even: the append form ends on ctl[1])
  hipStream_t cs = st, own = nullptr;
  if (!st) {  // the null stream cannot be captured
    GB_HIP(hipStreamCreate(&own));
    cs = own;
    hipEvent_t ready;
    GB_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    GB_HIP(hipEventRecord(ready, st));
    GB_HIP(hipStreamWaitEvent(cs, ready, 0));
    (void)hipEventDestroy(ready);
  }
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  gb_status gst = GB_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int k = 0; k < PH; k++) {
      if (SCAN) {
        hipLaunchKernelGGL((ms_step_kernel<W, true>), dim3(grid), dim3(MS_BLOCK), 0, cs, c,
                           (const MsCtl *)ctl);
        hipLaunchKernelGGL(ms_compact_kernel<W>, dim3(cgrid), dim3(MS_BLOCK), 0, cs, c,
                           (const MsCtl *)ctl, crange);
        hipLaunchKernelGGL(ms_ctrl_kernel, dim3(1), dim3(1), 0, cs, c, ctl);
      } else {
        MsCtl *cur = ctl + (k % 2), *prev = ctl + ((k + 1) % 2);
        hipLaunchKernelGGL(ms_clear_ctrl_kernel, dim3(cgrid), dim3(256), 0, cs, c,
                           (const MsCtl *)prev, cur);
        hipLaunchKernelGGL((ms_step_kernel<W, false>), dim3(grid), dim3(MS_BLOCK), 0, cs, c,
                           (const MsCtl *)cur);
      }
    }
    e = hipStreamEndCapture(cs, &graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) gst = gb_fail(GB_ERR_HIP, "msssp: graph capture failed: %s", hipGetErrorString(e));
  MsCtl hs{};
  int64_t last = -1;
  while (gst == GB_OK) {
    if (hipGraphLaunch(exec, cs) != hipSuccess ||
        hipMemcpyAsync(&hs, hstate, sizeof(hs), hipMemcpyDeviceToHost, cs) != hipSuccess ||
        hipStreamSynchronize(cs) != hipSuccess) {
      gst = gb_fail(GB_ERR_HIP, "msssp: phase replay failed");
      break;
    }
    if (hs.done) break;
    if (hs.phases == last) {
      gst = gb_fail(GB_ERR_HIP, "msssp: bucket loop made no progress");
      break;
    }
    last = hs.phases;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (own) {
    (void)hipStreamSynchronize(own);
    (void)hipStreamDestroy(own);
  }
  GB_TRY(gst);
  if (hs.capped) return gb_fail(GB_ERR_NOT_CONVERGED, "sssp: phase cap reached");
  constexpr int VT = 4096 / W;
  hipLaunchKernelGGL(ms_unpack_kernel<W>, dim3((unsigned)((V + VT - 1) / VT)), dim3(256), 0, st,
                     V, c.ntask, (const uint32_t *)c.dist, out_rows);
  GB_HIP(hipGetLastError());
  GB_TRY(gb_wstat_fold(nwaves, c.wstat, c.cnt, MS_RELAX, MS_IMP_NEAR, MS_IMP_FAR, st,
                       MS_EXPANDED));
  unsigned long long h[MS_N];
  GB_HIP(hipMemcpyAsync(h, c.cnt, sizeof(h), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  sx.phases = hs.phases;
  sx.buckets = hs.buckets;
  sx.relax = h[MS_RELAX];
  sx.imp_near = (int64_t)h[MS_IMP_NEAR];
  sx.imp_far = (int64_t)h[MS_IMP_FAR];
  sx.expanded = (int64_t)h[MS_EXPANDED];
  sx.ovf = h[MS_OVF] != 0;
  return GB_OK;
}

// One group's solve in shape sh.
template <int W>
gb_status ms_group_shape(MsCtx c, const int32_t *d_src, int max_phases, const MsShapeRt &sh,
                         hipStream_t st, uint32_t *out_rows, MsStats &sx) {
  c.cshift = sh.cshift;
  const int grid = gb_num_cus() * sh.per_cu;
  return sh.scan ? ms_group<W, true>(c, d_src, max_phases, grid, st, out_rows, sx)
                 : ms_group<W, false>(c, d_src, max_phases, grid, st, out_rows, sx);
}

// The shape for a group of width W on graph g: the environment's fixed parts,
// then the shape tuned for (W, delta) cached in the graph handle, else the
// rules.  When tuning is on and the handle has no shape for (W, delta) yet,
// the group is solved in several shapes (every solve is exact, so each leaves
// the same distances): the rules' shape, then one change at a time (the other
// form, the other blocks-per-CU counts, the other chunk), keeping a change
// when it is faster.  A shape's time is the wall time of its solve, graph
// capture and host waits included; below 20 ms, the better of two solves.
// The stats returned are those of the kept shape's solve.
template <int W>
gb_status ms_group_tuned(gb_graph g, MsCtx c, const int32_t *d_src, int max_phases,
                         hipStream_t st, uint32_t *out_rows, MsStats &sx) {
  const MsFixed f = ms_fixed();
  MsShapeRt sh{f.scan >= 0 ? f.scan == 1 : ms_rule_scan(c.V),
               f.per_cu > 0 ? f.per_cu : ms_rule_per_cu(c.E),
               f.cshift > 0 ? f.cshift : MsShape<W>::CHUNK_SHIFT};
  const bool all_fixed = f.scan >= 0 && f.per_cu > 0 && f.cshift > 0;
  const gb_graph_s::MsTuned *cached = nullptr;
  for (const auto &t : g->ms_tuned)
    if (t.W == W && t.delta == c.delta) cached = &t;
  if (cached && !all_fixed) {
    if (f.scan < 0) sh.scan = cached->scan == 1;
    if (f.per_cu < 0) sh.per_cu = cached->per_cu;
    if (f.cshift < 0) sh.cshift = cached->cshift;
  }
  if (!f.tune || all_fixed || cached || max_phases > 0)
    return ms_group_shape<W>(c, d_src, max_phases, sh, st, out_rows, sx);
  int reps = 1;
  auto timed = [&](const MsShapeRt &s, double &ms, MsStats &o) -> gb_status {
    ms = 1e300;
    for (int r = 0; r < reps; r++) {
      GB_HIP(hipStreamSynchronize(st));
      const auto t0 = std::chrono::steady_clock::now();
      GB_TRY(ms_group_shape<W>(c, d_src, max_phases, s, st, out_rows, o));
      GB_HIP(hipStreamSynchronize(st));
      ms = std::min(ms, std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0).count());
    }
    return GB_OK;
  };
  MsShapeRt best = sh;
  double best_ms = 0;
  MsStats best_st;
  GB_TRY(timed(best, best_ms, best_st));
  if (best_ms < 20.0) {  // short solves: the better of two from here on
    reps = 2;
    GB_TRY(timed(best, best_ms, best_st));
  }
  auto consider = [&](MsShapeRt s) -> gb_status {
    double ms = 0;
    MsStats o;
    GB_TRY(timed(s, ms, o));
    if (ms < best_ms) {
      best = s;
      best_ms = ms;
      best_st = o;
    }
    return GB_OK;
  };
  if (f.scan < 0) {
    MsShapeRt s = best;
    s.scan = !s.scan;
    GB_TRY(consider(s));
  }
  if (f.per_cu < 0) {
    const int base = best.per_cu;
    for (int pc : {4, 8, 16})
      if (pc != base) {
        MsShapeRt s = best;
        s.per_cu = pc;
        GB_TRY(consider(s));
      }
  }
  if (f.cshift < 0) {
    MsShapeRt s = best;
    s.cshift = s.cshift == 5 ? 6 : 5;
    GB_TRY(consider(s));
  }
  g->ms_tuned.push_back(gb_graph_s::MsTuned{W, c.delta, (int8_t)(best.scan ? 1 : 0),
                                            (int8_t)best.per_cu, (int8_t)best.cshift});
  sx = best_st;  // out_rows: the last solve's distances, the same in every shape
  return GB_OK;
}

}  // namespace

void gb_msssp_release() { g_msw.release(); }

// Distances from nsrc sources in groups of up to 64 (u32 weights); rows of
// d_dist as gb_sssp.  Called by gb_sssp for nsrc >= 2.
gb_status gb_msssp_count(gb_graph g, int32_t nsrc, const uint32_t *d_dist, int64_t *reached,
                         int64_t *traversed, hipStream_t st) {
  MsWork &w = g_msw;
  GB_TRY(w.cnt.ensure(sizeof(unsigned long long) * MS_N));
  GB_HIP(hipMemsetAsync(w.cnt.p, 0, 2 * sizeof(unsigned long long), st));
  const int64_t bx = std::max<int64_t>(
      1, std::min<int64_t>(((int64_t)g->V + 255) / 256, std::max<int64_t>(1, gb_num_cus() * 4 / nsrc)));
  hipLaunchKernelGGL(ms_count_kernel, dim3((unsigned)bx, (unsigned)std::min(nsrc, 65535)), dim3(256), 0, st,
                     (int64_t)g->V, nsrc, g->rp, d_dist, w.cnt.as<unsigned long long>());
  GB_HIP(hipGetLastError());
  unsigned long long h[2];
  GB_HIP(hipMemcpyAsync(h, w.cnt.p, sizeof(h), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  *reached = (int64_t)h[0];
  *traversed = (int64_t)h[1];
  return GB_OK;
}

gb_status gb_msssp_run(gb_graph g, uint32_t delta, const int32_t *h_sources, int32_t nsrc,
                       uint32_t *d_dist, int max_phases, int64_t *phases, int64_t *buckets,
                       unsigned long long *relax, int64_t *imp_near, int64_t *imp_far,
                       int64_t *expanded, bool *ovf, hipStream_t st) {
  MsWork &w = g_msw;
  const int64_t V = g->V, E = g->E;
  // queue capacity: a vertex enters a near queue at most once per phase with
  // ceil(deg / chunk) entries, a far queue at most once
  const int64_t qn = V + E / MS_MIN_CHUNK + 1;
  const int Wmax = nsrc >= 64 ? 64 : (nsrc > 32 ? 64 : nsrc > 16 ? 32 : nsrc > 8 ? 16 : 8);
  GB_TRY(w.dist.ensure(sizeof(uint32_t) * (size_t)V * Wmax));
  // both forms' buffers: the autotuner may try either
  GB_TRY(w.m0.ensure(sizeof(unsigned long long) * V));
  GB_TRY(w.m1.ensure(sizeof(unsigned long long) * V));
  GB_TRY(w.far.ensure(sizeof(unsigned long long) * V));
  GB_TRY(w.q[0].ensure(sizeof(MsEnt) * qn));
  GB_TRY(w.q[1].ensure(sizeof(MsEnt) * qn));
  GB_TRY(w.q[2].ensure(sizeof(uint64_t) * (V + 1)));
  GB_TRY(w.q[3].ensure(sizeof(uint64_t) * (V + 1)));
  GB_TRY(w.cnt.ensure(sizeof(unsigned long long) * MS_N));
  GB_TRY(w.wstat.ensure(sizeof(unsigned long long) * GB_WSTAT * gb_num_cus() * MS_MAX_PER_CU *
                        (MS_BLOCK / GB_WAVE)));
  GB_TRY(w.ctl.ensure(2 * sizeof(MsCtl)));
  GB_TRY(w.src.ensure(sizeof(int32_t) * 64));
  MsCtx c{};
  c.rp = g->rp;
  c.col = g->col;
  c.w = g->w;
  c.V = g->V;
  c.E = g->E;
  c.dist = w.dist.as<uint32_t>();
  c.mask[0] = w.m0.as<unsigned long long>();
  c.mask[1] = w.m1.as<unsigned long long>();
  c.far = w.far.as<unsigned long long>();
  c.qn[0] = w.q[0].as<MsEnt>();
  c.qn[1] = w.q[1].as<MsEnt>();
  c.qf[0] = w.q[2].as<uint64_t>();
  c.qf[1] = w.q[3].as<uint64_t>();
  c.cnt = w.cnt.as<unsigned long long>();
  c.wstat = w.wstat.as<unsigned long long>();
  c.delta = delta;
  c.magic = delta > 1 ? (~0ull / delta) + 1 : 0;
  MsStats tot;
  for (int32_t g0 = 0; g0 < nsrc; g0 += 64) {
    const int n = std::min<int32_t>(64, nsrc - g0);
    c.ntask = n;
    GB_HIP(hipMemcpyAsync(w.src.p, h_sources + g0, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    uint32_t *rows = d_dist + (int64_t)g0 * V;
    const int W = n > 32 ? 64 : n > 16 ? 32 : n > 8 ? 16 : 8;
    const int32_t *ds = w.src.as<int32_t>();
    MsStats sx;
    switch (W) {
      case 8:
        GB_TRY(ms_group_tuned<8>(g, c, ds, max_phases, st, rows, sx));
        break;
      case 16:
        GB_TRY(ms_group_tuned<16>(g, c, ds, max_phases, st, rows, sx));
        break;
      case 32:
        GB_TRY(ms_group_tuned<32>(g, c, ds, max_phases, st, rows, sx));
        break;
      default:
        GB_TRY(ms_group_tuned<64>(g, c, ds, max_phases, st, rows, sx));
    }
    tot.add(sx);
  }
  *phases += tot.phases;
  *buckets += tot.buckets;
  *relax += tot.relax;
  *imp_near += tot.imp_near;
  *imp_far += tot.imp_far;
  *expanded += tot.expanded;
  if (tot.ovf) *ovf = true;
  return GB_OK;
}
