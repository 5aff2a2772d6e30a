// gb_sssp.hip -- delta-stepping SSSP for gfx950.
//
// Replaces run_sspath_arr (sssp/csr_sspath_array.c:149-363) and the distance
// part of run_sspath (sssp/csr_sspath.c:38-122).  The reference's per-bucket
// pipeline request_arr -> merge_sort -> reorder_arr -> relax_arr becomes:
//   light phase  waves take queue entries (a vertex, or a 256-edge chunk of
//                a hub's light edges); lanes relax consecutive edges with
//                atomicMin on the distance; an improved target whose new
//                distance is still in the current bucket is appended to the
//                next near queue (dedup by a per-vertex generation stamp),
//                otherwise to the far pile (dedup per far epoch).  Queue
//                slots are reserved with one atomicAdd per wave.
//   heavy phase  the same over the bucket's settled set S (vertices expanded
//                in the bucket's light phases, remember_arr 1044-1098), heavy
//                edges only (w > light_max, separate_csr graph/csr.c:75-130).
//   far scan     when the bucket is exhausted, one pass over the far pile
//                moves entries of the next non-empty bucket to the near queue
//                and keeps the rest.
// Sorting by target and first-of-run relaxation are not needed: atomicMin
// resolves concurrent requests to the same target.
// Many tasks -- sources of one graph, or whole graphs -- run together: every
// queue entry carries its task id and all tasks step through bucket indices
// in lockstep, so one launch serves all of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"
#include "gb_sssp_ctl.hpp"

namespace {

constexpr int SS_BLOCK = 256;
constexpr int CHUNK_SHIFT = 8;  // 256 edges per queue entry at most
constexpr int32_t CHUNK = 1 << CHUNK_SHIFT;

__device__ __host__ __forceinline__ uint32_t nchunks(int32_t deg) {
  return (uint32_t)((deg + CHUNK - 1) >> CHUNK_SHIFT);
}
__device__ __forceinline__ uint64_t mk_entry(uint32_t t, uint32_t ch, uint32_t v) {
  return ((uint64_t)t << 48) | ((uint64_t)(ch & 0xFFFFu) << 32) | v;
}

// distance/weight traits
struct TU32 {
  using D = uint32_t;
  using W = uint32_t;
  using DT = uint64_t;
  static __device__ __forceinline__ D inf() { return 0xFFFFFFFFu; }
  static __device__ __forceinline__ D add(D du, W w, bool &ovf) {
    uint64_t nd = (uint64_t)du + w;
    if (nd >= 0xFFFFFFFFull) {
      ovf = true;
      nd = 0xFFFFFFFEull;
    }
    return (D)nd;
  }
  static __device__ __forceinline__ D amin(D *p, D v) { return atomicMin(p, v); }
  static __device__ __forceinline__ void amin_noret(D *p, D v) { (void)atomicMin(p, v); }
  // floor(d / delta) as a multiply-high by magic = floor((2^64 - 1) / delta) + 1,
  // exact for d, delta < 2^32 (delta = 1: magic = 0, the quotient is d)
  static __device__ __forceinline__ uint64_t bucket(D d, DT delta, uint64_t magic) {
    return magic ? __umul64hi((uint64_t)d, magic) : (uint64_t)d;
  }
  static uint64_t host_magic(DT delta) { return delta > 1 ? (~0ull / delta) + 1 : 0; }
};
struct TF64 {
  using D = double;
  using W = double;
  using DT = double;
  static __device__ __forceinline__ D inf() { return __longlong_as_double(0x7FF0000000000000ll); }
  static __device__ __forceinline__ D add(D du, W w, bool &) { return __dadd_rn(du, w); }
  static __device__ __forceinline__ D amin(D *p, D v) {
    // non-negative doubles order like their bit patterns
    unsigned long long old =
        atomicMin((unsigned long long *)p, (unsigned long long)__double_as_longlong(v));
    return __longlong_as_double((long long)old);
  }
  static __device__ __forceinline__ void amin_noret(D *p, D v) {
    (void)atomicMin((unsigned long long *)p, (unsigned long long)__double_as_longlong(v));
  }
  static __device__ __forceinline__ uint64_t bucket(D d, DT delta, uint64_t) {
    return (uint64_t)(long long)(d / delta);
  }
  static uint64_t host_magic(DT) { return 0; }
};

template <class W>
struct SGraph {
  const int32_t *rp;
  const int32_t *le;  // light end per vertex (light edges first in each row)
  const int32_t *col;
  const W *w;
  int32_t V;
  int32_t pad;
  int64_t E;
};

template <class Tr>
struct STask {
  typename Tr::D *dist;
  int64_t off;  // offset of this task's stamps
  typename Tr::DT delta;
  uint64_t magic;  // bucket division constant of delta
  int32_t graph;
  int32_t src;
};


template <class Tr>
struct Ctx {
  const SGraph<typename Tr::W> *graphs;
  const STask<Tr> *tasks;
  uint32_t *nst;  // near generation stamp
  uint32_t *sst;  // settled-set round stamp
  uint32_t *fst;  // far epoch stamp
  unsigned long long *cnt;
  unsigned long long *wstat;  // per-wave statistics slots (GB_WSTAT per wave)
  // one graph and one delta, task t's distances at dist0 + t * V (sources of a
  // single graph): task data then follow from t without per-edge lookups
  typename Tr::D *dist0;
  typename Tr::DT delta0;
  uint64_t magic0;
};

// What the kernels need of task t: its distances and stamps, bucket width
// and graph.  With SINGLE the graph is the kernel-wide G0.
template <class Tr, bool SINGLE>
struct TaskView {
  typename Tr::D *dist;
  int64_t off;
  typename Tr::DT delta;
  uint64_t magic;
  SGraph<typename Tr::W> G;
  __device__ __forceinline__ TaskView(const Ctx<Tr> &c, const SGraph<typename Tr::W> &G0,
                                      uint32_t t) {
    if constexpr (SINGLE) {
      off = (int64_t)t * G0.V;
      dist = c.dist0 + off;
      delta = c.delta0;
      magic = c.magic0;
      G = G0;
    } else {
      const STask<Tr> tk = c.tasks[t];
      dist = tk.dist;
      off = tk.off;
      delta = tk.delta;
      magic = tk.magic;
      G = c.graphs[tk.graph];
    }
  }
  __device__ __forceinline__ typename Tr::D *d(uint32_t v) const { return dist + v; }
  __device__ __forceinline__ int64_t s(uint32_t v) const { return off + v; }
};

template <class Tr, bool SINGLE>
__global__ __launch_bounds__(SS_BLOCK) void sssp_fill_kernel(Ctx<Tr> c, int32_t ntasks) {
  // grid.y = task
  const int32_t t = blockIdx.y;
  if (t >= ntasks) return;
  SGraph<typename Tr::W> G0{};
  if constexpr (SINGLE) G0 = c.graphs[0];
  const TaskView<Tr, SINGLE> tk(c, G0, t);
  const int32_t V = tk.G.V;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    *tk.d(v) = Tr::inf();
    c.nst[tk.s(v)] = 0;
    c.sst[tk.s(v)] = 0;
    c.fst[tk.s(v)] = 0;
  }
}

template <class Tr, bool SINGLE>
__global__ void sssp_seed_kernel(Ctx<Tr> c, int32_t ntasks, uint64_t *near_out,
                                 unsigned long long *n_near, uint32_t gen) {
  const int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= ntasks) return;
  SGraph<typename Tr::W> G0{};
  if constexpr (SINGLE) G0 = c.graphs[0];
  const TaskView<Tr, SINGLE> tk(c, G0, t);
  const SGraph<typename Tr::W> &G = tk.G;
  const int32_t s = c.tasks[t].src;
  if (s < 0 || s >= G.V) return;
  *tk.d(s) = (typename Tr::D)0;
  c.nst[tk.s(s)] = gen;
  uint32_t k = nchunks(G.le[s] - G.rp[s]);
  if (k == 0) k = 1;
  unsigned long long b = atomicAdd(n_near, (unsigned long long)k);
  for (uint32_t i = 0; i < k; i++) near_out[b + i] = mk_entry((uint32_t)t, i, (uint32_t)s);
}

// queue appends are staged per wave in LDS (gb_wave_queue, gb_device.hpp)
constexpr int NEAR_CAP = 384, FAR_CAP = 128, S_CAP = 128;  // 5 KiB of LDS per wave
// edges relaxed per wave and round: EXP_U slots of 64; their improved targets
// are listed in LDS (IMP_CAP per wave) before the stamps are taken
constexpr int EXP_U = 4, IMP_CAP = GB_WAVE * EXP_U;

// Light (HEAVY=false) or heavy (HEAVY=true) expansion of one queue.
// Each wave takes 64 queue entries at a time (one per lane) and spreads their
// edges over its lanes: a wave prefix sum of the entries' edge counts gives
// every edge slot k an owner lane, found by a binary search over the lanes'
// offsets with cross-lane reads.  Lanes therefore stay busy whether entries
// hold 3 edges or 256.  A round covers EXP_U * 64 edge slots in phases, so
// that each phase's memory accesses are in flight together: the slots' col/w
// reads, then their target distance reads, then the atomicMin of every
// improving slot (no return value: improvement is judged on the value read);
// the improved targets are listed in LDS and then stamped and queued 64 at a
// time.
template <class T>
__device__ __forceinline__ T gb_shfl_any(T v, int src);
template <>
__device__ __forceinline__ uint32_t gb_shfl_any<uint32_t>(uint32_t v, int src) {
  return __shfl(v, src, GB_WAVE);
}
template <>
__device__ __forceinline__ double gb_shfl_any<double>(double v, int src) {
  return __longlong_as_double((long long)gb_shfl_u64((uint64_t)__double_as_longlong(v), src));
}

template <class Tr, bool HEAVY, bool SINGLE>
__device__ __forceinline__ void expand_body(
    const Ctx<Tr> &c, const SGraph<typename Tr::W> &G0, uint64_t *lds_near, uint64_t *lds_far,
    uint64_t *lds_s, uint64_t *lds_imp, const uint64_t *__restrict__ in,
    const unsigned long long *__restrict__ n_in, uint64_t *__restrict__ near_out,
    unsigned long long *n_near, uint64_t *__restrict__ s_out, unsigned long long *n_s,
    uint64_t *__restrict__ far_out, unsigned long long *n_far, uint64_t bucket,
    uint32_t gen_next, uint32_t sround, uint32_t fepoch, uint32_t epw) {
  using D = typename Tr::D;
  const int lane = gb_lane();
  gb_wave_queue<uint64_t, NEAR_CAP> nearq(lds_near, near_out, n_near);
  gb_wave_queue<uint64_t, FAR_CAP> farq(lds_far, far_out, n_far);
  gb_wave_queue<uint64_t, S_CAP> sq(lds_s, s_out, n_s);
  const uint64_t n = *n_in;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t relax = 0, imp_near = 0, imp_far = 0;
  bool ovf = false;
  // epw (1..64) entries per wave and round: small queues use fewer entries per
  // wave so that more waves run
  for (uint64_t b0 = wave * epw; b0 < n; b0 += nwaves * epw) {
    // ---- this lane's entry
    const uint64_t i = b0 + lane;
    uint32_t t = 0, u = 0, scnt = 0;
    int32_t beg = 0, end = 0;
    D du = (D)0;
    if ((uint32_t)lane < epw && i < n) {
      const uint64_t ent = in[i];
      t = (uint32_t)(ent >> 48);
      const uint32_t ch = (uint32_t)(ent >> 32) & 0xFFFFu;
      u = (uint32_t)ent;
      const TaskView<Tr, SINGLE> tk(c, G0, t);
      du = *tk.d(u);
      const int32_t rs = tk.G.rp[u], le = tk.G.le[u], re = tk.G.rp[u + 1];
      if (!HEAVY) {
        beg = rs + (int32_t)(ch << CHUNK_SHIFT);
        end = min(le, beg + CHUNK);
        // remember_arr: the first chunk records u in S (heavy edges pending)
        if (ch == 0 && re > le && atomicExch(&c.sst[tk.s(u)], sround) != sround)
          scnt = nchunks(re - le);
      } else {
        beg = le + (int32_t)(ch << CHUNK_SHIFT);
        end = min(re, beg + CHUNK);
      }
    }
    if (!HEAVY) sq.push(scnt, [&](uint32_t k) { return mk_entry(t, k, u); });
    // ---- spread the 64 entries' edges over the lanes
    uint32_t total;
    const uint32_t off = gb_wave_excl_sum((uint32_t)max(end - beg, 0), &total);
    for (uint32_t k0 = 0; k0 < total; k0 += GB_WAVE * EXP_U) {
      uint32_t tj[EXP_U], v[EXP_U];
      int32_t e[EXP_U];
      D duj[EXP_U], nd[EXP_U], old[EXP_U];
      bool ok[EXP_U];
      // slots past the end take slot k0's edge (k0 < total), so that every
      // read below is in bounds and none needs a branch
#pragma unroll
      for (int s = 0; s < EXP_U; s++) {
        const uint32_t k = k0 + (uint32_t)(s * GB_WAVE + lane);
        // owner: the last lane j with off_j <= k (offsets are non-decreasing)
        int j = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
          const int cand = j + step;
          const uint32_t oc = __shfl(off, cand < GB_WAVE ? cand : GB_WAVE - 1, GB_WAVE);
          if (cand < GB_WAVE && oc <= k) j = cand;
        }
        tj[s] = __shfl(t, j, GB_WAVE);
        e[s] = __shfl(beg, j, GB_WAVE) + (int32_t)(k - __shfl(off, j, GB_WAVE));
        duj[s] = gb_shfl_any<D>(du, j);
        ok[s] = k < total;
      }
#pragma unroll
      for (int s = 0; s < EXP_U; s++) {
        const uint32_t t0 = __shfl(tj[0], 0, GB_WAVE);
        const int32_t e0 = __shfl(e[0], 0, GB_WAVE);
        tj[s] = ok[s] ? tj[s] : t0;
        e[s] = ok[s] ? e[s] : e0;
      }
      typename Tr::W wv[EXP_U];
#pragma unroll
      for (int s = 0; s < EXP_U; s++) {
        const TaskView<Tr, SINGLE> tk(c, G0, tj[s]);
        v[s] = (uint32_t)tk.G.col[e[s]];
        wv[s] = tk.G.w[e[s]];
      }
      D *dp[EXP_U];
      bool isfar[EXP_U];
#pragma unroll
      for (int s = 0; s < EXP_U; s++) {
        const TaskView<Tr, SINGLE> tk(c, G0, tj[s]);
        bool o = false;
        nd[s] = Tr::add(duj[s], wv[s], o);
        ovf |= o & ok[s];
        relax += ok[s];
        isfar[s] = Tr::bucket(nd[s], tk.delta, tk.magic) > bucket;
        dp[s] = tk.d(v[s]);
        old[s] = *dp[s];
      }
#pragma unroll
      for (int s = 0; s < EXP_U; s++) ok[s] = ok[s] & (nd[s] < old[s]);
#pragma unroll
      for (int s = 0; s < EXP_U; s++)
        if (ok[s]) Tr::amin_noret(dp[s], nd[s]);
      // list the improved targets: bit 63 = later bucket, task, vertex
      uint32_t nimp = 0;
#pragma unroll
      for (int s = 0; s < EXP_U; s++) {
        imp_near += ok[s] & !isfar[s];
        imp_far += ok[s] & isfar[s];
        uint32_t cnt_s;
        const uint32_t pos = gb_wave_excl_sum(ok[s] ? 1u : 0u, &cnt_s);
        if (ok[s])
          lds_imp[nimp + pos] =
              ((uint64_t)isfar[s] << 63) | ((uint64_t)tj[s] << 32) | (uint64_t)v[s];
        nimp += cnt_s;
      }
      __builtin_amdgcn_wave_barrier();
      // stamp and queue them: a plain read of the stamp first, the exchange
      // only when unset.  A vertex without out-edges is stamped but not
      // queued; one with only heavy edges gets one (empty) chunk, which
      // records it in S.
      for (uint32_t q0 = 0; q0 < nimp; q0 += GB_WAVE) {
        const bool has = q0 + lane < nimp;
        const uint64_t item = has ? lds_imp[q0 + lane] : 0ull;
        const uint32_t ti = (uint32_t)(item >> 32) & 0xFFFFu, vi = (uint32_t)item;
        const bool isfar = (item >> 63) != 0;
        const TaskView<Tr, SINGLE> tk(c, G0, ti);
        uint32_t *sp = isfar ? &c.fst[tk.s(vi)] : &c.nst[tk.s(vi)];
        const uint32_t g = isfar ? fepoch : gen_next;
        bool fresh = false;
        if (has && *sp != g) fresh = atomicExch(sp, g) != g;
        uint32_t ncnt = 0;
        bool far = false;
        if (fresh) {
          const int32_t rs = tk.G.rp[vi], re = tk.G.rp[vi + 1];
          if (isfar) {
            far = re > rs;
          } else {
            ncnt = nchunks(tk.G.le[vi] - rs);
            if (ncnt == 0 && re > rs) ncnt = 1;
          }
        }
        nearq.push(ncnt, [&](uint32_t q) { return mk_entry(ti, q, vi); });
        farq.push(far ? 1u : 0u, [&](uint32_t) { return mk_entry(ti, 0, vi); });
      }
      __builtin_amdgcn_wave_barrier();  // the list is rewritten next round
    }
  }
  gb_block_flush(nearq, farq, sq);
  // per-wave statistics
  uint32_t tot, tn, tf;
  (void)gb_wave_excl_sum(relax, &tot);
  (void)gb_wave_excl_sum(imp_near, &tn);
  (void)gb_wave_excl_sum(imp_far, &tf);
  if (lane == 0) {
    gb_wstat_add(c.wstat, 0, tot);
    gb_wstat_add(c.wstat, 1, tn);
    gb_wstat_add(c.wstat, 2, tf);
  }
  if (__ballot(ovf) && lane == 0) atomicOr(&c.cnt[C_OVF], 1ull);
}

template <class Tr, bool SINGLE>
__device__ __forceinline__ void far_scan_body(
    const Ctx<Tr> &c, const SGraph<typename Tr::W> &G0, uint64_t *lds_near, uint64_t *lds_far,
    const uint64_t *__restrict__ in, const unsigned long long *__restrict__ n_in,
    uint64_t *__restrict__ near_out, unsigned long long *n_near, uint64_t *__restrict__ far_out,
    unsigned long long *n_far, uint64_t bnext, uint32_t gen_next, uint32_t fepoch_next) {
  const int lane = gb_lane();
  gb_wave_queue<uint64_t, NEAR_CAP> nearq(lds_near, near_out, n_near);
  gb_wave_queue<uint64_t, FAR_CAP> farq(lds_far, far_out, n_far);
  const uint64_t n = *n_in;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t minb = ~0ull;
  for (uint64_t b0 = wave * GB_WAVE; b0 < n; b0 += nwaves * GB_WAVE) {
    const uint64_t i = b0 + lane;
    uint32_t ncnt = 0;
    bool far = false;
    uint32_t t = 0, v = 0;
    if (i < n) {
      const uint64_t ent = in[i];
      t = (uint32_t)(ent >> 48);
      v = (uint32_t)ent;
      const TaskView<Tr, SINGLE> tk(c, G0, t);
      const uint64_t bk = Tr::bucket(*tk.d(v), tk.delta, tk.magic);
      if (bk == bnext) {
        if (atomicExch(&c.nst[tk.s(v)], gen_next) != gen_next) {
          const SGraph<typename Tr::W> &G = tk.G;
          ncnt = nchunks(G.le[v] - G.rp[v]);
          if (ncnt == 0) ncnt = 1;
        }
      } else if (bk > bnext) {
        if (atomicExch(&c.fst[tk.s(v)], fepoch_next) != fepoch_next) far = true;
        minb = bk < minb ? bk : minb;
      }
    }
    nearq.push(ncnt, [&](uint32_t k) { return mk_entry(t, k, v); });
    farq.push(far ? 1u : 0u, [&](uint32_t) { return mk_entry(t, 0, v); });
  }
  gb_block_flush(nearq, farq);
  gb_block_atomic_min(minb, ~0ull, &c.cnt[C_MINB]);
}

// ---- device-driven phase loop ----------------------------------------------
// The bucket loop runs as a state machine on the device, so that the host
// neither reads queue sizes nor chooses launches: sssp_ctrl_kernel (one lane)
// looks at the counters left by the previous phase, advances the loop state,
// clears the counters the next phase appends to, and writes the next action;
// sssp_step_kernel (fixed grid) performs that action.  Pairs of the two are
// captured in a hipGraph and replayed until the state says done.
// Queues: 0/1 near (double-buffered), 2 settled set S, 3/4 far (double-
// buffered); queue k's length is counter k.
struct Queues {
  uint64_t *q[5];
};

__global__ void sssp_ctrl_kernel(unsigned long long *__restrict__ cnt, SsspCtl *__restrict__ ctl,
                                 uint32_t cus) {
  if (threadIdx.x != 0) return;
  SsspCtl s = *ctl;
  sssp_ctl_advance(s, cnt, cus);
  *ctl = s;
}

template <class Tr, bool SINGLE>
__global__ __launch_bounds__(SS_BLOCK) void sssp_step_kernel(Ctx<Tr> c, Queues q,
                                                              const SsspCtl *__restrict__ ctl) {
  const int act = ctl->action;
  if (act == ACT_NONE) return;
  unsigned long long *cnt = c.cnt;
  const int in = ctl->a_in, nr = ctl->a_near, fr = ctl->a_far;
  SGraph<typename Tr::W> G0{};
  if constexpr (SINGLE) G0 = c.graphs[0];
  __shared__ uint64_t s_near[SS_BLOCK / GB_WAVE][NEAR_CAP];
  __shared__ uint64_t s_far[SS_BLOCK / GB_WAVE][FAR_CAP];
  __shared__ uint64_t s_set[SS_BLOCK / GB_WAVE][S_CAP];
  __shared__ uint64_t s_imp[SS_BLOCK / GB_WAVE][IMP_CAP];
  uint64_t *ln = s_near[threadIdx.x / GB_WAVE], *lf = s_far[threadIdx.x / GB_WAVE];
  uint64_t *ls = s_set[threadIdx.x / GB_WAVE], *li = s_imp[threadIdx.x / GB_WAVE];
  if (act == ACT_FAR) {
    far_scan_body<Tr, SINGLE>(c, G0, ln, lf, q.q[in], cnt + in, q.q[nr], cnt + nr, q.q[fr],
                              cnt + fr, ctl->a_bucket, ctl->a_gen, ctl->a_fepoch);
  } else if (act == ACT_LIGHT) {
    expand_body<Tr, false, SINGLE>(c, G0, ln, lf, ls, li, q.q[in], cnt + in, q.q[nr], cnt + nr,
                                   q.q[C_S], cnt + C_S, q.q[fr], cnt + fr, ctl->a_bucket,
                                   ctl->a_gen, ctl->a_sround, ctl->a_fepoch, ctl->a_epw);
  } else {
    expand_body<Tr, true, SINGLE>(c, G0, ln, lf, ls, li, q.q[in], cnt + in, q.q[nr], cnt + nr,
                                  q.q[C_S], cnt + C_S, q.q[fr], cnt + fr, ctl->a_bucket,
                                  ctl->a_gen, ctl->a_sround, ctl->a_fepoch, ctl->a_epw);
  }
}

// traversed edges: sum of out-degrees of reached (task, vertex) pairs
template <class Tr>
__global__ __launch_bounds__(SS_BLOCK) void sssp_count_kernel(Ctx<Tr> c, int32_t ntasks,
                                                               unsigned long long *out) {
  const int32_t t = blockIdx.y;
  if (t >= ntasks) return;
  const STask<Tr> tk = c.tasks[t];
  const SGraph<typename Tr::W> G = c.graphs[tk.graph];
  unsigned long long reached = 0, edges = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < G.V;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (tk.dist[v] != Tr::inf()) {
      reached++;
      edges += (unsigned long long)(G.rp[v + 1] - G.rp[v]);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    reached += __shfl_xor(reached, d, 64);
    edges += __shfl_xor(edges, d, 64);
  }
  __shared__ unsigned long long s_r[SS_BLOCK / 64], s_e[SS_BLOCK / 64];
  if (gb_lane() == 0) {
    s_r[threadIdx.x >> 6] = reached;
    s_e[threadIdx.x >> 6] = edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // one atomic pair per block
    unsigned long long r = 0, e = 0;
    for (int k = 0; k < SS_BLOCK / 64; k++) {
      r += s_r[k];
      e += s_e[k];
    }
    atomicAdd(&out[0], r);
    atomicAdd(&out[1], e);
  }
}

struct SsspWork {
  DevBuf graphs, tasks, nst, sst, fst, q[5], cnt, counts, ctl, wstat;
  void release() {
    for (DevBuf *b : {&graphs, &tasks, &nst, &sst, &fst, &cnt, &counts, &ctl, &wstat}) b->release();
    for (auto &b : q) b.release();
  }
};
// Heap singleton, never destroyed: device memory must not be freed from a
// static destructor, which can run after the HIP runtime has shut down.
SsspWork &g_ssw_ref() {
  static SsspWork *w = new SsspWork();
  return *w;
}
#define g_ssw (g_ssw_ref())

template <class Tr>
gb_status sssp_engine(const std::vector<SGraph<typename Tr::W>> &hg,
                      std::vector<STask<Tr>> &ht, int max_phases, gb_sssp_stats *stats,
                      hipStream_t st) {
  using G = SGraph<typename Tr::W>;
  const int32_t ntasks = (int32_t)ht.size();
  if (ntasks <= 0) return GB_OK;
  if (ntasks > 65535) return gb_fail(GB_ERR_INVALID, "sssp: at most 65535 tasks per call");
  int64_t total = 0, qcap = 0;
  int32_t maxV = 0;
  for (auto &t : ht) {
    const G &g = hg[t.graph];
    t.off = total;
    total += g.V;
    maxV = std::max(maxV, g.V);
  }
  // Queue capacity: a (task, vertex) pair enters a queue at most once per
  // generation, with ceil(deg / CHUNK) >= 1 entries, so per task at most
  // V + E / CHUNK + 1 entries.
  for (auto &t : ht) qcap += (int64_t)hg[t.graph].V + hg[t.graph].E / CHUNK + 1;
  for (auto &t : ht) t.magic = Tr::host_magic(t.delta);
  // one graph and delta, distances of task i at dist0 + i * V: the kernels
  // then derive task data from the task id
  bool single = hg.size() == 1;
  for (auto &t : ht)
    single = single && t.graph == 0 && t.delta == ht[0].delta && t.dist == ht[0].dist + t.off;
  SsspWork &w = g_ssw;
  GB_TRY(w.graphs.ensure(sizeof(G) * hg.size()));
  GB_TRY(w.tasks.ensure(sizeof(STask<Tr>) * ht.size()));
  GB_TRY(w.nst.ensure(sizeof(uint32_t) * total));
  GB_TRY(w.sst.ensure(sizeof(uint32_t) * total));
  GB_TRY(w.fst.ensure(sizeof(uint32_t) * total));
  for (int k = 0; k < 5; k++) GB_TRY(w.q[k].ensure(sizeof(uint64_t) * qcap));
  GB_TRY(w.cnt.ensure(sizeof(unsigned long long) * C_N));
  // Step-kernel blocks per CU: 4 while the largest graph has fewer than 2^27
  // edges, 16 beyond.  Measured at 4 / 16: one source of R-MAT 16 / 20 / 22
  // 1.93 / 3.09 / 4.61 ms against 2.30 / 4.11 / 5.85 ms, the 512-graph batch
  // 1569 against 1831 ms; one source of R-MAT 24 14.3 against 13.8 ms.
  int64_t max_e = 0;
  for (const G &gr : hg) max_e = std::max<int64_t>(max_e, gr.E);
  const int grid = gb_num_cus() * (max_e < ((int64_t)1 << 27) ? 4 : 16);
  const int64_t nwaves = (int64_t)grid * (SS_BLOCK / GB_WAVE);
  GB_TRY(w.wstat.ensure(sizeof(unsigned long long) * GB_WSTAT * nwaves));
  GB_TRY(w.ctl.ensure(sizeof(SsspCtl)));
  GB_HIP(hipMemcpyAsync(w.graphs.p, hg.data(), sizeof(G) * hg.size(), hipMemcpyHostToDevice, st));
  GB_HIP(hipMemcpyAsync(w.tasks.p, ht.data(), sizeof(STask<Tr>) * ht.size(),
                        hipMemcpyHostToDevice, st));
  unsigned long long *cnt = w.cnt.as<unsigned long long>();
  GB_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * C_N, st));
  unsigned long long *wstat = w.wstat.as<unsigned long long>();
  GB_HIP(hipMemsetAsync(wstat, 0, sizeof(unsigned long long) * GB_WSTAT * nwaves, st));
  Ctx<Tr> c{w.graphs.as<G>(), w.tasks.as<STask<Tr>>(), w.nst.as<uint32_t>(), w.sst.as<uint32_t>(),
            w.fst.as<uint32_t>(), cnt, wstat, ht[0].dist, ht[0].delta, ht[0].magic};

  hipEvent_t e0, e1;
  GB_HIP(hipEventCreate(&e0));
  GB_HIP(hipEventCreate(&e1));
  GB_HIP(hipEventRecord(e0, st));
  {
    int gx = (int)std::min<int64_t>((maxV + SS_BLOCK - 1) / SS_BLOCK, 1024);
    if (gx < 1) gx = 1;
    if (single)
      hipLaunchKernelGGL((sssp_fill_kernel<Tr, true>), dim3(gx, ntasks), dim3(SS_BLOCK), 0, st, c,
                         ntasks);
    else
      hipLaunchKernelGGL((sssp_fill_kernel<Tr, false>), dim3(gx, ntasks), dim3(SS_BLOCK), 0, st, c,
                         ntasks);
  }
  const uint32_t gen = 1;
  if (single)
    hipLaunchKernelGGL((sssp_seed_kernel<Tr, true>), dim3((ntasks + 255) / 256), dim3(256), 0, st,
                       c, ntasks, w.q[0].as<uint64_t>(), cnt + C_NEAR0, gen);
  else
    hipLaunchKernelGGL((sssp_seed_kernel<Tr, false>), dim3((ntasks + 255) / 256), dim3(256), 0, st,
                       c, ntasks, w.q[0].as<uint64_t>(), cnt + C_NEAR0, gen);
  GB_HIP(hipGetLastError());
  // the loop state starts at bucket 0 with the seeds in near queue 0
  SsspCtl hc{};
  hc.gen = gen;
  hc.sround = 0;
  hc.fepoch = 1;
  hc.max_phases = max_phases > 0 ? max_phases : ((int64_t)1 << 40);
  hc.state = ST_BUCKET;
  GB_HIP(hipMemcpyAsync(w.ctl.p, &hc, sizeof(hc), hipMemcpyHostToDevice, st));
  SsspCtl *ctl = w.ctl.as<SsspCtl>();
  Queues qs{{w.q[0].as<uint64_t>(), w.q[1].as<uint64_t>(), w.q[2].as<uint64_t>(),
             w.q[3].as<uint64_t>(), w.q[4].as<uint64_t>()}};
  const uint32_t cus = (uint32_t)gb_num_cus();
  auto one_phase = [&]() {
    hipLaunchKernelGGL(sssp_ctrl_kernel, dim3(1), dim3(64), 0, st, cnt, ctl, cus);
    if (single)
      hipLaunchKernelGGL((sssp_step_kernel<Tr, true>), dim3(grid), dim3(SS_BLOCK), 0, st, c, qs,
                         (const SsspCtl *)ctl);
    else
      hipLaunchKernelGGL((sssp_step_kernel<Tr, false>), dim3(grid), dim3(SS_BLOCK), 0, st, c, qs,
                         (const SsspCtl *)ctl);
  };
  // Chunks of PH phases are captured once per call and replayed until the
  // state is done; a replay past the end costs two near-empty launches per
  // phase.  The null stream cannot be captured: a blocking stream, ordered
  // with it, is used for the replays instead.
  constexpr int PH = 16;
  hipStream_t cs = st;
  hipStream_t own = nullptr;
  if (!st) {
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
  {
    hipStream_t keep = st;
    st = cs;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int k = 0; k < PH; k++) one_phase();
      e = hipStreamEndCapture(cs, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    st = keep;
    if (e != hipSuccess) gst = gb_fail(GB_ERR_HIP, "sssp: graph capture failed: %s", hipGetErrorString(e));
  }
  SsspCtl hs{};
  int64_t last_phases = -1;
  while (gst == GB_OK) {
    if (hipGraphLaunch(exec, cs) != hipSuccess ||
        hipMemcpyAsync(&hs, ctl, sizeof(hs), hipMemcpyDeviceToHost, cs) != hipSuccess ||
        hipStreamSynchronize(cs) != hipSuccess) {
      gst = gb_fail(GB_ERR_HIP, "sssp: phase replay failed");
      break;
    }
    if (hs.done) break;
    // every replay that is not done issues PH actions; guard the host loop
    if (hs.phases == last_phases) {
      gst = gb_fail(GB_ERR_HIP, "sssp: bucket loop made no progress");
      break;
    }
    last_phases = hs.phases;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (own) {
    (void)hipStreamSynchronize(own);
    (void)hipStreamDestroy(own);
  }
  if (gst == GB_OK && hs.capped) gst = gb_fail(GB_ERR_NOT_CONVERGED, "sssp: phase cap reached");
  if (gst != GB_OK) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return gst;
  }
  GB_HIP(hipGetLastError());
  gb_status status = GB_OK;
  const int64_t phases = hs.phases, buckets = hs.buckets;
  GB_TRY(gb_wstat_fold(nwaves, wstat, cnt, C_RELAX, C_IMP_NEAR, C_IMP_FAR, st));
  GB_HIP(hipEventRecord(e1, st));
  unsigned long long h[C_N];
  GB_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  if (h[C_OVF]) status = gb_fail(GB_ERR_OVERFLOW, "sssp: uint32 distance overflow");
  float ms = 0.0f;
  GB_HIP(hipEventSynchronize(e1));
  GB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (stats) {
    stats->buckets = buckets;
    stats->phases = phases;
    stats->relax_edges = (int64_t)h[C_RELAX];
    stats->improved_near = (int64_t)h[C_IMP_NEAR];
    stats->improved_far = (int64_t)h[C_IMP_FAR];
    stats->expanded_edges = (int64_t)h[C_RELAX];  // one source per expansion
    stats->ms = ms;
    GB_TRY(w.counts.ensure(2 * sizeof(unsigned long long)));
    GB_HIP(hipMemsetAsync(w.counts.p, 0, 2 * sizeof(unsigned long long), st));
    int gx = (int)std::min<int64_t>((maxV + SS_BLOCK - 1) / SS_BLOCK,
                                    std::max<int64_t>(1, (int64_t)gb_num_cus() * 4 / ntasks));
    if (gx < 1) gx = 1;
    hipLaunchKernelGGL(sssp_count_kernel<Tr>, dim3(gx, ntasks), dim3(SS_BLOCK), 0, st, c, ntasks,
                       w.counts.as<unsigned long long>());
    unsigned long long hc[2];
    GB_HIP(hipMemcpyAsync(hc, w.counts.p, sizeof(hc), hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    stats->reached = (int64_t)hc[0];
    stats->traversed = (int64_t)hc[1];
  }
  return status;
}

// Queue entries carry a 16-bit chunk index: a vertex may have at most
// 65535 chunks of CHUNK edges.
gb_status check_degree(gb_graph g, hipStream_t st) {
  if (g->E <= (int64_t)0xFFFF * CHUNK) return GB_OK;
  int32_t md = 0;
  GB_TRY(gb_graph_max_outdeg(g, &md, st));
  if ((int64_t)md > (int64_t)0xFFFF * CHUNK)
    return gb_fail(GB_ERR_UNSUPPORTED, "sssp: out-degree %d above %d", md, 0xFFFF * CHUNK);
  return GB_OK;
}

// Several sources of one graph go to the batched engine unless
// GB_SSSP_PER_SOURCE is set (the per-source engine then runs them in lockstep
// with one queue entry per source and vertex).
bool ms_enabled() {
  const char *e = getenv("GB_SSSP_PER_SOURCE");
  return !(e && e[0] && e[0] != '0');
}

uint32_t plain_delta(uint32_t max_w, int32_t V, int64_t E) {
  // max_weight / average degree, as an integer bucket width >= 1
  if (E <= 0 || V <= 0) return 1;
  double d = (double)max_w * (double)V / (double)E;
  uint32_t r = (uint32_t)std::floor(d);
  return r < 1 ? 1 : r;
}

// The graph's rows as light and heavy parts for threshold lm.  When no weight
// exceeds lm every edge is light: the graph's own arrays serve, with the row
// end as the light end, and nothing is copied.
gb_status light_rows(gb_graph g, uint32_t lm, hipStream_t st, SGraph<uint32_t> &out,
                     bool sync = true) {
  uint32_t mw = 0;
  GB_TRY(gb_graph_max_w(g, &mw, st));
  if (lm >= mw) {
    out = SGraph<uint32_t>{g->rp, g->rp + 1, g->col, g->w, g->V, 0, g->E};
    return GB_OK;
  }
  GB_TRY(gb_graph_split(g, lm, st, sync));
  out = SGraph<uint32_t>{g->rp, g->split_le.as<int32_t>(), g->split_col.as<int32_t>(),
                         g->split_w.as<uint32_t>(), g->V, 0, g->E};
  return GB_OK;
}

}  // namespace

void gb_sssp_release() { g_ssw.release(); }

extern "C" gb_status gb_sssp(gb_graph g, const gb_sssp_params *p, const int32_t *h_sources,
                             int32_t nsrc, uint32_t *d_dist, gb_sssp_stats *stats,
                             void *stream) {
  if (!g || !h_sources || nsrc < 0 || !d_dist) return gb_fail(GB_ERR_INVALID, "gb_sssp: null");
  if (!g->w && g->E > 0) return gb_fail(GB_ERR_INVALID, "gb_sssp: graph has no uint32 weights");
  for (int32_t i = 0; i < nsrc; i++)
    if (h_sources[i] < 0 || h_sources[i] >= g->V)
      return gb_fail(GB_ERR_INVALID, "gb_sssp: source %d out of range", h_sources[i]);
  hipStream_t st = gb_stream(stream);
  uint32_t delta = p ? p->delta : 0, lmax = p ? p->light_max : 0;
  if (delta == 0) {
    uint32_t mw = 0;
    GB_TRY(gb_graph_max_w(g, &mw, st));
    delta = plain_delta(mw, g->V, g->E);
  }
  if (lmax == 0) lmax = delta;
  if (nsrc >= 2 && ms_enabled()) {
    // several sources: the batched engine (one entry per vertex and source mask)
    gb_sssp_stats s{};
    bool ovf = false;
    unsigned long long relax = 0;
    hipEvent_t e0, e1;
    GB_HIP(hipEventCreate(&e0));
    GB_HIP(hipEventCreate(&e1));
    GB_HIP(hipEventRecord(e0, st));
    gb_status r = gb_msssp_run(g, delta, h_sources, nsrc, d_dist, p ? p->max_phases : 0,
                               &s.phases, &s.buckets, &relax, &s.improved_near,
                               &s.improved_far, &s.expanded_edges, &ovf, st);
    if (r == GB_OK) r = hipEventRecord(e1, st) == hipSuccess ? GB_OK : gb_fail(GB_ERR_HIP, "event");
    if (r == GB_OK && stats) {
      float ms = 0.0f;
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&ms, e0, e1);
      s.ms = ms;
      s.relax_edges = (int64_t)relax;
      r = gb_msssp_count(g, nsrc, d_dist, &s.reached, &s.traversed, st);
      *stats = s;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    GB_TRY(r);
    if (ovf) return gb_fail(GB_ERR_OVERFLOW, "sssp: uint32 distance overflow");
    return GB_OK;
  }
  GB_TRY(check_degree(g, st));
  std::vector<SGraph<uint32_t>> hg(1);
  GB_TRY(light_rows(g, lmax, st, hg[0]));
  std::vector<STask<TU32>> ht(nsrc);
  for (int32_t i = 0; i < nsrc; i++)
    ht[i] = STask<TU32>{d_dist + (int64_t)i * g->V, 0, (uint64_t)delta, 0, 0, h_sources[i]};
  return sssp_engine<TU32>(hg, ht, p ? p->max_phases : 0, stats, st);
}

extern "C" gb_status gb_sssp_f64(gb_graph g, double delta, const int32_t *h_sources,
                                 int32_t nsrc, double *d_dist, gb_sssp_stats *stats,
                                 void *stream) {
  if (!g || !h_sources || nsrc < 0 || !d_dist) return gb_fail(GB_ERR_INVALID, "gb_sssp_f64: null");
  if (!g->w64 && g->E > 0) return gb_fail(GB_ERR_INVALID, "gb_sssp_f64: graph has no f64 weights");
  if (!(delta > 0.0) || !std::isfinite(delta))
    return gb_fail(GB_ERR_INVALID, "gb_sssp_f64: delta must be positive");
  for (int32_t i = 0; i < nsrc; i++)
    if (h_sources[i] < 0 || h_sources[i] >= g->V)
      return gb_fail(GB_ERR_INVALID, "gb_sssp_f64: source %d out of range", h_sources[i]);
  hipStream_t st = gb_stream(stream);
  GB_TRY(check_degree(g, st));
  GB_TRY(gb_graph_split64(g, delta, st));
  std::vector<SGraph<double>> hg(1);
  hg[0] = SGraph<double>{g->rp, g->split64_le.as<int32_t>(), g->split64_col.as<int32_t>(),
                         g->split64_w.as<double>(), g->V, 0, g->E};
  std::vector<STask<TF64>> ht(nsrc);
  for (int32_t i = 0; i < nsrc; i++)
    ht[i] = STask<TF64>{d_dist + (int64_t)i * g->V, 0, delta, 0, 0, h_sources[i]};
  return sssp_engine<TF64>(hg, ht, 0, stats, st);
}

extern "C" gb_status gb_sssp_multi(int32_t ngraphs, const gb_graph *gs, const int32_t *sources,
                                   const uint32_t *deltas, const uint32_t *light_max,
                                   uint32_t *const *d_dist, gb_sssp_stats *stats, void *stream) {
  if (ngraphs < 0 || (ngraphs > 0 && (!gs || !sources || !d_dist)))
    return gb_fail(GB_ERR_INVALID, "gb_sssp_multi: null");
  // the engine's task limit, checked before any split is queued
  if (ngraphs > 65535) return gb_fail(GB_ERR_INVALID, "gb_sssp_multi: at most 65535 graphs per call");
  hipStream_t st = gb_stream(stream);
  std::vector<SGraph<uint32_t>> hg(ngraphs);
  std::vector<STask<TU32>> ht(ngraphs);
  // a handle caches one split: a handle repeated with another threshold that
  // needs a split would find its rows overwritten
  std::unordered_map<gb_graph, uint32_t> split_lm;
  auto prep = [&](int32_t i) -> gb_status {
    gb_graph g = gs[i];
    if (!g || (!g->w && g->E > 0) || sources[i] < 0 || sources[i] >= g->V || !d_dist[i])
      return gb_fail(GB_ERR_INVALID, "gb_sssp_multi: bad graph %d", i);
    uint32_t mw = 0;
    GB_TRY(gb_graph_max_w(g, &mw, st));
    uint32_t delta = deltas ? deltas[i] : 0;
    if (delta == 0) delta = plain_delta(mw, g->V, g->E);
    uint32_t lm = light_max && light_max[i] ? light_max[i] : delta;
    if (lm < mw) {
      auto it = split_lm.find(g);
      if (it != split_lm.end() && it->second != lm)
        return gb_fail(GB_ERR_INVALID,
                       "gb_sssp_multi: graph %d repeats a handle with another light threshold", i);
      split_lm.emplace(g, lm);
    }
    GB_TRY(check_degree(g, st));
    // the splits stay queued on st, ahead of the solve (no wait per graph)
    GB_TRY(light_rows(g, lm, st, hg[i], false));
    ht[i] = STask<TU32>{d_dist[i], 0, (uint64_t)delta, 0, i, sources[i]};
    return GB_OK;
  };
  for (int32_t i = 0; i < ngraphs; i++) {
    const gb_status s = prep(i);
    if (s != GB_OK) {
      (void)hipStreamSynchronize(st);  // queued splits end before the scratch is free
      return s;
    }
  }
  const gb_status r = sssp_engine<TU32>(hg, ht, 0, stats, st);
  // on an early error the splits may still be queued: let them end before
  // the caller reuses the split buffers or the shared scratch
  if (r != GB_OK) (void)hipStreamSynchronize(st);
  return r;
}
