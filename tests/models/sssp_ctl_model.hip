// Host model of the SSSP phases driven by sssp_ctl_advance (the device state
// machine), checked against Dijkstra on random graphs.
#include "gb_sssp_ctl.hpp"
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <random>
#include <vector>
#include <algorithm>
using u64 = uint64_t;
constexpr int CH = 256;  // edges per queue entry, as in gb_sssp.hip
static uint32_t nch(int d) { return (uint32_t)((d + CH - 1) / CH); }
struct G { int V; std::vector<int> rp, le, col; std::vector<uint32_t> w; };
static u64 ent(uint32_t ch, uint32_t v) { return ((u64)ch << 32) | v; }
int main(int argc, char **argv) {
  int trials = argc > 1 ? atoi(argv[1]) : 200;
  std::mt19937_64 rng(12345);
  long long total_phases = 0;
  for (int tr = 0; tr < trials; tr++) {
    int V = 1 + rng() % 300;
    int E = rng() % (V * 8 + 1);
    uint32_t wmax = 1 + rng() % 1000;
    std::vector<std::vector<std::pair<int, uint32_t>>> adj(V);
    for (int e = 0; e < E; e++) adj[rng() % V].push_back({(int)(rng() % V), (uint32_t)(rng() % (wmax + 1))});
    if (tr % 7 == 0 && V > 1) for (int e = 0; e < 3000; e++) adj[0].push_back({(int)(rng() % V), (uint32_t)(rng() % (wmax + 1))});
    uint32_t delta = 1 + rng() % (wmax + 5), lmax = (tr % 3 == 0) ? delta / 2 : delta;
    G g; g.V = V; g.rp.assign(V + 1, 0); g.le.assign(V, 0);
    for (int u = 0; u < V; u++) {
      std::stable_sort(adj[u].begin(), adj[u].end(), [&](auto a, auto b) { return (a.second <= lmax) > (b.second <= lmax); });
      g.rp[u + 1] = g.rp[u] + adj[u].size();
      int l = 0;
      for (auto &p : adj[u]) { g.col.push_back(p.first); g.w.push_back(p.second); if (p.second <= lmax) l++; }
      g.le[u] = g.rp[u] + l;
    }
    int src = rng() % V;
    // dijkstra
    std::vector<u64> ref(V, ~0ull); ref[src] = 0;
    std::priority_queue<std::pair<u64, int>, std::vector<std::pair<u64, int>>, std::greater<>> pq; pq.push({0, src});
    while (!pq.empty()) { auto [d, u] = pq.top(); pq.pop(); if (d > ref[u]) continue;
      for (int e = g.rp[u]; e < g.rp[u + 1]; e++) { u64 nd = d + g.w[e]; if (nd < ref[g.col[e]]) { ref[g.col[e]] = nd; pq.push({nd, g.col[e]}); } } }
    // model state
    std::vector<uint32_t> dist(V, 0xFFFFFFFFu), nst(V, 0), sst(V, 0), fst(V, 0);
    std::vector<std::vector<u64>> q(5, std::vector<u64>(V + E / CH + 3 + 3000));
    unsigned long long cnt[C_N] = {0};
    // seed (gen 1)
    dist[src] = 0; nst[src] = 1;
    { uint32_t k = nch(g.le[src] - g.rp[src]); if (!k) k = 1; for (uint32_t i = 0; i < k; i++) q[0][cnt[0]++] = ent(i, src); }
    SsspCtl s{}; s.gen = 1; s.fepoch = 1; s.max_phases = 1 << 30; s.state = ST_BUCKET;
    long guard = 0;
    auto push = [&](int qi, u64 e) { if (cnt[qi] >= q[qi].size()) { printf("queue overflow\n"); exit(1); } q[qi][cnt[qi]++] = e; };
    for (;;) {
      sssp_ctl_advance(s, cnt, 256);
      if (s.done) break;
      if (++guard > 10000000) { printf("no termination trial %d\n", tr); return 1; }
      if (s.action == ACT_NONE) continue;
      if (s.a_in < 0 || s.a_in > 4 || s.a_near < 0 || s.a_near > 4 || s.a_far < 0 || s.a_far > 4) { printf("bad queue idx\n"); return 1; }
      u64 n = cnt[s.a_in];
      if (s.action == ACT_FAR) {
        for (u64 i = 0; i < n; i++) {
          uint32_t v = (uint32_t)q[s.a_in][i];
          u64 bk = dist[v] / delta;
          if (bk == s.a_bucket) { if (nst[v] != s.a_gen) { nst[v] = s.a_gen; uint32_t k = nch(g.le[v] - g.rp[v]); if (!k) k = 1; for (uint32_t c = 0; c < k; c++) push(s.a_near, ent(c, v)); } }
          else if (bk > s.a_bucket) { if (fst[v] != s.a_fepoch) { fst[v] = s.a_fepoch; push(s.a_far, ent(0, v)); } if (bk < cnt[C_MINB]) cnt[C_MINB] = bk; }
        }
      } else {
        bool heavy = s.action == ACT_HEAVY;
        for (u64 i = 0; i < n; i++) {
          u64 en = q[s.a_in][i]; uint32_t ch = (uint32_t)(en >> 32) & 0xFFFF, u = (uint32_t)en;
          int rs = g.rp[u], le = g.le[u], re = g.rp[u + 1], beg, end;
          if (!heavy) { beg = rs + ch * CH; end = std::min(le, beg + CH);
            if (ch == 0 && re > le && sst[u] != s.a_sround) { sst[u] = s.a_sround; for (uint32_t k = 0; k < nch(re - le); k++) push(C_S, ent(k, u)); } }
          else { beg = le + ch * CH; end = std::min(re, beg + CH); }
          for (int e = beg; e < end; e++) {
            int v = g.col[e]; u64 nd = (u64)dist[u] + g.w[e];
            if (nd < dist[v]) { dist[v] = (uint32_t)nd;
              if (nd / delta <= s.a_bucket) { if (nst[v] != s.a_gen) { nst[v] = s.a_gen; uint32_t k = nch(g.le[v] - g.rp[v]); if (!k) k = 1; for (uint32_t c = 0; c < k; c++) push(s.a_near, ent(c, v)); } }
              else { if (fst[v] != s.a_fepoch) { fst[v] = s.a_fepoch; push(s.a_far, ent(0, v)); } } }
          }
        }
      }
    }
    total_phases += s.phases;
    for (int v = 0; v < V; v++) {
      u64 d = dist[v] == 0xFFFFFFFFu ? ~0ull : dist[v];
      if (d != ref[v]) { printf("MISMATCH trial %d v %d got %llu ref %llu (V=%d E=%d delta=%u lmax=%u)\n", tr, v, (unsigned long long)d, (unsigned long long)ref[v], V, E, delta, lmax); return 1; }
    }
  }
  printf("ok: %d trials, %lld phases\n", trials, total_phases);
  return 0;
}
