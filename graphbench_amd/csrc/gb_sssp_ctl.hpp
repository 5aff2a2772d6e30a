// gb_sssp_ctl.hpp -- the delta-stepping bucket loop as a state machine.
//
// sssp_ctl_advance runs between two phases: it reads the queue counters the
// previous phase left, advances the loop (bucket, round, light/heavy/far
// phase), clears the counters the next phase appends to and records the next
// action in the state.  The device runs it in a one-lane kernel; being
// host-callable, it is also exercised against a host model of the phases.
#ifndef GB_SSSP_CTL_HPP
#define GB_SSSP_CTL_HPP
#include <hip/hip_runtime.h>

#include <cstdint>

// counters (u64): near A/B, S, far A/B, min bucket, relaxations, overflow.
// Queue k (0/1 near, 2 settled set S, 3/4 far) has length counter k.
// C_IMP_NEAR / C_IMP_FAR count relaxations that improved a distance, with
// the new distance in the current bucket or in a later one.
enum { C_NEAR0 = 0, C_NEAR1 = 1, C_S = 2, C_FAR0 = 3, C_FAR1 = 4, C_MINB = 5, C_RELAX = 6,
       C_OVF = 7, C_IMP_NEAR = 8, C_IMP_FAR = 9, C_N = 10 };
constexpr uint32_t SSSP_MAX_EPW = 64;  // queue entries per wave at most

enum { ACT_NONE = 0, ACT_LIGHT = 1, ACT_HEAVY = 2, ACT_FAR = 3 };
enum { ST_BUCKET = 0, ST_ROUND, ST_LIGHT, ST_AFTER_HEAVY, ST_AFTER_BUCKET, ST_FAR, ST_AFTER_FAR };

struct SsspCtl {
  uint64_t b, bnext;          // current bucket, next candidate bucket
  int64_t phases, buckets, max_phases;
  uint32_t gen, sround, fepoch;
  int32_t cur, fcur, state;
  int32_t done, capped;
  // the action for the next step
  int32_t action, a_in, a_near, a_far;
  uint64_t a_bucket;
  uint32_t a_gen, a_sround, a_fepoch, a_epw;
};

__host__ __device__ inline uint32_t sssp_entries_per_wave(uint64_t n, uint32_t cus) {
  const uint64_t target = (uint64_t)cus * 32;
  uint64_t e = (n + target - 1) / target;
  return (uint32_t)(e < 1 ? 1 : (e > SSSP_MAX_EPW ? SSSP_MAX_EPW : e));
}

__host__ __device__ inline void sssp_ctl_advance(SsspCtl &s, unsigned long long *cnt,
                                                uint32_t cus) {
  s.action = ACT_NONE;
  // at most a handful of transitions separate two actions; the bound only
  // guarantees termination of this loop
  for (int guard = 0; guard < 16 && s.action == ACT_NONE && !s.done; ++guard) {
    switch (s.state) {
      case ST_BUCKET:
        s.buckets++;
        s.state = ST_ROUND;
        break;
      case ST_ROUND:
        s.sround++;
        cnt[C_S] = 0;
        s.state = ST_LIGHT;
        break;
      case ST_LIGHT:
        if (cnt[C_NEAR0 + s.cur] > 0 || cnt[C_S] > 0) {
          if (++s.phases > s.max_phases) {
            s.capped = 1;
            s.done = 1;
            break;
          }
          s.gen++;
          s.a_bucket = s.b;
          s.a_gen = s.gen;
          s.a_sround = s.sround;
          s.a_fepoch = s.fepoch;
          s.a_far = C_FAR0 + s.fcur;
          if (cnt[C_NEAR0 + s.cur] > 0) {  // light phase over the near queue
            s.action = ACT_LIGHT;
            s.a_in = C_NEAR0 + s.cur;
            s.a_near = C_NEAR0 + (s.cur ^ 1);
            cnt[s.a_near] = 0;
            s.cur ^= 1;
          } else {  // heavy phase over the bucket's settled set
            s.action = ACT_HEAVY;
            s.a_in = C_S;
            s.a_near = C_NEAR0 + s.cur;
            cnt[s.a_near] = 0;
            s.state = ST_AFTER_HEAVY;
          }
          s.a_epw = sssp_entries_per_wave(cnt[s.a_in], cus);
        } else {
          s.state = ST_AFTER_BUCKET;
        }
        break;
      case ST_AFTER_HEAVY:  // a heavy edge may land back in this bucket
        s.state = cnt[C_NEAR0 + s.cur] > 0 ? ST_ROUND : ST_AFTER_BUCKET;
        break;
      case ST_AFTER_BUCKET:
        if (cnt[C_FAR0 + s.fcur] == 0) {
          s.done = 1;
        } else {
          s.bnext = s.b + 1;
          s.state = ST_FAR;
        }
        break;
      case ST_FAR:
        if (++s.phases > s.max_phases) {
          s.capped = 1;
          s.done = 1;
          break;
        }
        s.gen++;
        s.action = ACT_FAR;
        s.a_in = C_FAR0 + s.fcur;
        s.a_near = C_NEAR0 + s.cur;
        s.a_far = C_FAR0 + (s.fcur ^ 1);
        s.a_bucket = s.bnext;
        s.a_gen = s.gen;
        s.a_fepoch = s.fepoch + 1;
        s.a_epw = SSSP_MAX_EPW;
        cnt[s.a_near] = 0;
        cnt[s.a_far] = 0;
        cnt[C_MINB] = ~0ull;
        s.fepoch++;
        s.fcur ^= 1;
        s.state = ST_AFTER_FAR;
        break;
      case ST_AFTER_FAR:
        if (cnt[C_NEAR0 + s.cur] > 0) {
          s.b = s.bnext;
          s.state = ST_BUCKET;
        } else if (cnt[C_FAR0 + s.fcur] == 0) {
          s.done = 1;
        } else {
          s.bnext = cnt[C_MINB];  // skip empty bucket indices
          s.state = ST_FAR;
        }
        break;
      default:
        s.done = 1;
    }
  }
}

// ---- batched form: one vertex entry serves a mask of tasks -----------------
// Near queues 0/1 (with per-vertex task masks 0/1), far queues 2/3 (one far
// mask per vertex), counters: MS_NEAR0..1, MS_FAR0..1, MS_MINB, MS_RELAX,
// MS_OVF, the improvement outcomes and MS_EXPANDED (edges of expanded queue
// entries, each read once for the entry's whole source mask).  No light/heavy
// split: an expansion relaxes all edges of a vertex.
enum { MS_NEAR0 = 0, MS_NEAR1 = 1, MS_FAR0 = 2, MS_FAR1 = 3, MS_MINB = 4, MS_RELAX = 5,
       MS_OVF = 6, MS_IMP_NEAR = 7, MS_IMP_FAR = 8, MS_EXPANDED = 9, MS_N = 10 };
enum { MS_ST_LIGHT = 0, MS_ST_AFTER_BUCKET, MS_ST_FAR, MS_ST_AFTER_FAR };

struct MsCtl {
  uint64_t b, bnext;
  int64_t phases, buckets, max_phases;
  int32_t cur, fcur, state, done, capped;
  // next action: ACT_LIGHT (near queue a_in -> a_near, far pushes -> a_far)
  // or ACT_FAR (far queue a_in -> near a_near and far a_far)
  int32_t action, a_in, a_near, a_far;
  uint64_t a_bucket;
  uint64_t a_in_n;  // light action: entries in queue a_in (its counter is reused later)
};

__host__ __device__ inline void ms_ctl_advance(MsCtl &s, unsigned long long *cnt) {
  s.action = ACT_NONE;
  for (int guard = 0; guard < 8 && s.action == ACT_NONE && !s.done; ++guard) {
    switch (s.state) {
      case MS_ST_LIGHT:
        if (cnt[MS_NEAR0 + s.cur] > 0) {
          if (++s.phases > s.max_phases) {
            s.capped = 1;
            s.done = 1;
            break;
          }
          s.action = ACT_LIGHT;
          s.a_in = MS_NEAR0 + s.cur;
          s.a_near = MS_NEAR0 + (s.cur ^ 1);
          s.a_far = MS_FAR0 + s.fcur;
          s.a_bucket = s.b;
          s.a_in_n = cnt[s.a_in];
          cnt[s.a_near] = 0;
          s.cur ^= 1;
        } else {
          s.state = MS_ST_AFTER_BUCKET;
        }
        break;
      case MS_ST_AFTER_BUCKET:
        if (cnt[MS_FAR0 + s.fcur] == 0) {
          s.done = 1;
        } else {
          s.bnext = s.b + 1;
          s.state = MS_ST_FAR;
        }
        break;
      case MS_ST_FAR:
        if (++s.phases > s.max_phases) {
          s.capped = 1;
          s.done = 1;
          break;
        }
        s.action = ACT_FAR;
        s.a_in = MS_FAR0 + s.fcur;
        s.a_near = MS_NEAR0 + s.cur;
        s.a_far = MS_FAR0 + (s.fcur ^ 1);
        s.a_bucket = s.bnext;
        cnt[s.a_near] = 0;
        cnt[s.a_far] = 0;
        cnt[MS_MINB] = ~0ull;
        s.fcur ^= 1;
        s.state = MS_ST_AFTER_FAR;
        break;
      case MS_ST_AFTER_FAR:
        if (cnt[MS_NEAR0 + s.cur] > 0) {
          s.b = s.bnext;
          s.buckets++;
          s.state = MS_ST_LIGHT;
        } else if (cnt[MS_FAR0 + s.fcur] == 0) {
          s.done = 1;
        } else {
          s.bnext = cnt[MS_MINB];
          s.state = MS_ST_FAR;
        }
        break;
      default:
        s.done = 1;
    }
  }
}

#endif
