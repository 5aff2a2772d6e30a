// Host model of the batched (task-mask) SSSP phases driven by ms_ctl_advance.
#include "gb_sssp_ctl.hpp"
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <random>
#include <vector>
#include <algorithm>
using u64 = uint64_t;
constexpr int CH = 64;
static uint32_t nch(int d) { uint32_t k = (uint32_t)((d + CH - 1) / CH); return k ? k : 1; }
int main(int argc, char **argv) {
  int trials = argc > 1 ? atoi(argv[1]) : 200;
  std::mt19937_64 rng(777);
  long long tot_ph = 0;
  for (int tr = 0; tr < trials; tr++) {
    int V = 1 + rng() % 300, E = rng() % (V * 8 + 1);
    uint32_t wmax = 1 + rng() % 1000;
    int W = 1 << (3 + rng() % 4);  // 8..64
    int T = 1 + rng() % W;
    std::vector<std::vector<std::pair<int, uint32_t>>> adj(V);
    for (int e = 0; e < E; e++) adj[rng() % V].push_back({(int)(rng() % V), (uint32_t)(rng() % (wmax + 1))});
    if (tr % 5 == 0 && V > 1) for (int e = 0; e < 500; e++) adj[0].push_back({(int)(rng() % V), (uint32_t)(rng() % (wmax + 1))});
    std::vector<int> rp(V + 1, 0), col; std::vector<uint32_t> w;
    for (int u = 0; u < V; u++) { rp[u + 1] = rp[u] + adj[u].size(); for (auto &p : adj[u]) { col.push_back(p.first); w.push_back(p.second); } }
    uint64_t delta = 1 + rng() % (wmax + 5);
    std::vector<int> src(T); for (auto &x : src) x = rng() % V;
    // model
    std::vector<uint32_t> dist((size_t)V * W, 0xFFFFFFFFu);
    std::vector<u64> mask[2] = {std::vector<u64>(V, 0), std::vector<u64>(V, 0)}, far(V, 0);
    size_t qcap = V + col.size() / CH + 2 + 600;
    std::vector<std::vector<u64>> q(4, std::vector<u64>(qcap));
    unsigned long long cnt[MS_N] = {0};
    auto push = [&](int qi, u64 e) { if (cnt[qi] >= q[qi].size()) { printf("overflow\n"); exit(1); } q[qi][cnt[qi]++] = e; };
    for (int t = 0; t < T; t++) {
      int s = src[t]; dist[(size_t)s * W + t] = 0;
      u64 old = mask[0][s]; mask[0][s] |= 1ull << t;
      if (old == 0) for (uint32_t c = 0; c < nch(rp[s + 1] - rp[s]); c++) push(MS_NEAR0, ((u64)c << 32) | s);
    }
    MsCtl s{}; s.max_phases = 1 << 30; s.state = MS_ST_LIGHT;
    long guard = 0;
    for (;;) {
      ms_ctl_advance(s, cnt);
      if (s.done) break;
      if (++guard > 10000000) { printf("no termination\n"); return 1; }
      if (s.action == ACT_NONE) continue;
      u64 n = cnt[s.a_in];
      if (s.action == ACT_LIGHT) {
        int mi = s.a_in - MS_NEAR0, mo = s.a_near - MS_NEAR0;
        for (u64 i = 0; i < n; i++) {
          u64 en = q[s.a_in][i]; uint32_t u = (uint32_t)en, ch = (uint32_t)(en >> 32);
          u64 m = mask[mi][u];
          int beg = rp[u] + ch * CH, end = std::min(rp[u + 1], beg + CH);
          for (int e = beg; e < end; e++) {
            int x = col[e]; u64 nm = 0, fm = 0;
            for (int t = 0; t < W; t++) {
              if (!((m >> t) & 1) || t >= T) continue;
              u64 nd = (u64)dist[(size_t)u * W + t] + w[e];
              if (nd < dist[(size_t)x * W + t]) { dist[(size_t)x * W + t] = (uint32_t)nd; if (nd / delta <= s.a_bucket) nm |= 1ull << t; else fm |= 1ull << t; }
            }
            if (nm) { u64 old = mask[mo][x]; mask[mo][x] |= nm; if (!old) for (uint32_t c = 0; c < nch(rp[x + 1] - rp[x]); c++) push(s.a_near, ((u64)c << 32) | x); }
            if (fm) { u64 old = far[x]; far[x] |= fm; if (!old) push(s.a_far, x); }
          }
        }
        // clear kernel
        for (u64 i = 0; i < n; i++) mask[mi][(uint32_t)q[s.a_in][i]] = 0;
      } else {
        int mo = s.a_near - MS_NEAR0;
        for (u64 i = 0; i < n; i++) {
          uint32_t x = (uint32_t)q[s.a_in][i]; u64 m = far[x], nm = 0, km = 0;
          for (int t = 0; t < W; t++) {
            if (!((m >> t) & 1)) continue;
            u64 bk = dist[(size_t)x * W + t] / delta;
            if (bk == s.a_bucket) nm |= 1ull << t;
            else if (bk > s.a_bucket) { km |= 1ull << t; if (bk < cnt[MS_MINB]) cnt[MS_MINB] = bk; }
          }
          far[x] = km;
          if (nm) { u64 old = mask[mo][x]; mask[mo][x] |= nm; if (!old) for (uint32_t c = 0; c < nch(rp[x + 1] - rp[x]); c++) push(s.a_near, ((u64)c << 32) | x); }
          if (km) push(s.a_far, x);
        }
      }
    }
    tot_ph += s.phases;
    for (int t = 0; t < T; t++) {
      std::vector<u64> ref(V, ~0ull); ref[src[t]] = 0;
      std::priority_queue<std::pair<u64, int>, std::vector<std::pair<u64, int>>, std::greater<>> pq; pq.push({0, src[t]});
      while (!pq.empty()) { auto [d, u] = pq.top(); pq.pop(); if (d > ref[u]) continue;
        for (int e = rp[u]; e < rp[u + 1]; e++) { u64 nd = d + w[e]; if (nd < ref[col[e]]) { ref[col[e]] = nd; pq.push({nd, col[e]}); } } }
      for (int v = 0; v < V; v++) {
        u64 d = dist[(size_t)v * W + t] == 0xFFFFFFFFu ? ~0ull : dist[(size_t)v * W + t];
        if (d != ref[v]) { printf("MISMATCH trial %d task %d v %d got %llu ref %llu\n", tr, t, v, (unsigned long long)d, (unsigned long long)ref[v]); return 1; }
      }
    }
  }
  printf("ok: %d trials, %lld phases\n", trials, tot_ph);
  return 0;
}
