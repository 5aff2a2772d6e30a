// gb_gen.cpp -- deterministic synthetic graph generators (host side).
//
// Counter-based: the draws for edge e depend only on (seed, e), so the result
// is independent of the thread count.  Per edge the generator runs a
// SplitMix64 stream seeded with mix(seed, e): `scale` quadrant draws for R-MAT
// (thresholds a, a+b, a+b+c scaled to 2^64), then one weight draw.
// Uniform integers use Lemire's multiply-shift bounded draw.
// Rows are sorted by (destination, weight), so the CSR does not depend on
// the scatter order either.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "gb_internal.hpp"

namespace {

inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Stream {
  uint64_t s;
  explicit Stream(uint64_t seed, uint64_t ctr) : s(mix64(seed ^ mix64(ctr))) {}
  inline uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // Lemire bounded draw in [0, n)
  inline uint64_t below(uint64_t n) {
    uint64_t x = next();
    __uint128_t m = (__uint128_t)x * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
      uint64_t t = (0 - n) % n;
      while (l < t) {
        x = next();
        m = (__uint128_t)x * n;
        l = (uint64_t)m;
      }
    }
    return (uint64_t)(m >> 64);
  }
  inline double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int nthreads() {
  const char *e = getenv("GB_GEN_THREADS");
  int n = e ? atoi(e) : 0;
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  if (n > 16) n = 16;  // hosts share cores between GPUs
  return n < 1 ? 1 : n;
}

template <class F>
void parallel_for(int64_t n, F f) {
  int T = nthreads();
  if (n < 65536 || T == 1) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++) {
    int64_t lo = n * t / T, hi = n * (t + 1) / T;
    th.emplace_back([=] { f(lo, hi, t); });
  }
  for (auto &x : th) x.join();
}

inline uint32_t draw_weight(Stream &r, int mode, uint32_t lo, uint32_t hi) {
  if (mode == 0) return 1;
  if (mode == 2 && (r.next() >> 63)) {
    // log-uniform on [1, 65535]: floor(exp(U ln 65536)) clamped
    double u = r.unit();
    double v = std::floor(std::exp(u * std::log(65536.0)));
    if (v < 1) v = 1;
    if (v > 65535) v = 65535;
    return (uint32_t)v;
  }
  return lo + (uint32_t)r.below((uint64_t)hi - lo + 1);
}

// Edge list -> CSR with rows sorted by (dst, w).
gb_status build_csr(int64_t V, const std::vector<uint32_t> &src,
                    const std::vector<uint32_t> &dst,
                    const std::vector<uint32_t> *w, gb_host_csr *out) {
  int64_t E = (int64_t)src.size();
  if (E >= (1ll << 31) || V >= (1ll << 31)) return GB_ERR_INVALID;
  out->V = (int32_t)V;
  out->E = E;
  out->row_ptr = (int32_t *)malloc(sizeof(int32_t) * (V + 1));
  out->col = (int32_t *)malloc(sizeof(int32_t) * (E ? E : 1));
  out->w = w ? (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1)) : nullptr;
  if (!out->row_ptr || !out->col || (w && !out->w)) return GB_ERR_NOMEM;
  std::vector<std::atomic<int32_t>> cnt(V);
  parallel_for(V, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; i++) cnt[i].store(0, std::memory_order_relaxed);
  });
  parallel_for(E, [&](int64_t lo, int64_t hi, int) {
    for (int64_t e = lo; e < hi; e++) cnt[src[e]].fetch_add(1, std::memory_order_relaxed);
  });
  out->row_ptr[0] = 0;
  for (int64_t i = 0; i < V; i++) out->row_ptr[i + 1] = out->row_ptr[i] + cnt[i].load();
  parallel_for(V, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; i++) cnt[i].store(out->row_ptr[i], std::memory_order_relaxed);
  });
  // scatter packed (dst<<32 | w) keys, then sort every row
  uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (E ? E : 1));
  if (!key) return GB_ERR_NOMEM;
  parallel_for(E, [&](int64_t lo, int64_t hi, int) {
    for (int64_t e = lo; e < hi; e++) {
      int32_t pos = cnt[src[e]].fetch_add(1, std::memory_order_relaxed);
      key[pos] = ((uint64_t)dst[e] << 32) | (w ? (*w)[e] : 0u);
    }
  });
  parallel_for(V, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; i++) {
      std::sort(key + out->row_ptr[i], key + out->row_ptr[i + 1]);
      for (int32_t p = out->row_ptr[i]; p < out->row_ptr[i + 1]; p++) {
        out->col[p] = (int32_t)(key[p] >> 32);
        if (out->w) out->w[p] = (uint32_t)key[p];
      }
    }
  });
  free(key);
  return GB_OK;
}

}  // namespace

extern "C" {

uint64_t gb_splitmix64(uint64_t x) { return mix64(x); }

void gb_host_csr_free(gb_host_csr *g) {
  if (!g) return;
  free(g->row_ptr);
  free(g->col);
  free(g->w);
  g->row_ptr = nullptr;
  g->col = nullptr;
  g->w = nullptr;
}

gb_status gb_gen_rmat(gb_host_csr *out, int32_t scale, int64_t nverts,
                      double edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t weight_mode, uint32_t wlo,
                      uint32_t whi, int32_t symmetric) {
  if (!out || scale < 1 || scale > 30 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0 || whi < wlo)
    return gb_fail(GB_ERR_INVALID, "gb_gen_rmat: bad arguments");
  const int64_t fullV = 1ll << scale;
  const int64_t V = nverts > 0 ? nverts : fullV;
  if (V > fullV) return gb_fail(GB_ERR_INVALID, "gb_gen_rmat: nverts > 2^scale");
  const int64_t M = (int64_t)std::llround(edge_factor * (double)V);
  const double two64 = 18446744073709551616.0;
  auto thr = [&](double p) -> uint64_t {
    if (p >= 1.0) return ~0ull;
    return (uint64_t)(p * two64);
  };
  const uint64_t tA = thr(a), tAB = thr(a + b), tABC = thr(a + b + c);
  std::vector<uint32_t> src(M), dst(M), wt(weight_mode ? M : 0);
  std::vector<uint8_t> keep(M);
  parallel_for(M, [&](int64_t lo, int64_t hi, int) {
    for (int64_t e = lo; e < hi; e++) {
      Stream r(seed, (uint64_t)e);
      uint64_t u, v;
      for (;;) {  // rejection keeps endpoints < V for non-power-of-two V
        u = 0;
        v = 0;
        for (int l = 0; l < scale; l++) {
          uint64_t x = r.next();
          uint64_t bu = 0, bv = 0;
          if (x < tA) {
          } else if (x < tAB) {
            bv = 1;
          } else if (x < tABC) {
            bu = 1;
          } else {
            bu = 1;
            bv = 1;
          }
          u = (u << 1) | bu;
          v = (v << 1) | bv;
        }
        if ((int64_t)u < V && (int64_t)v < V) break;
      }
      src[e] = (uint32_t)u;
      dst[e] = (uint32_t)v;
      keep[e] = u != v;
      if (weight_mode) wt[e] = draw_weight(r, weight_mode, wlo, whi);
    }
  });
  // compact (drop self-loops), optional symmetrization
  std::vector<uint32_t> s2, d2, w2;
  int64_t n = 0;
  for (int64_t e = 0; e < M; e++) n += keep[e];
  int64_t tot = symmetric ? 2 * n : n;
  s2.resize(tot);
  d2.resize(tot);
  if (weight_mode) w2.resize(tot);
  int64_t k = 0;
  for (int64_t e = 0; e < M; e++) {
    if (!keep[e]) continue;
    s2[k] = src[e];
    d2[k] = dst[e];
    if (weight_mode) w2[k] = wt[e];
    if (symmetric) {
      s2[n + k] = dst[e];
      d2[n + k] = src[e];
      if (weight_mode) w2[n + k] = wt[e];
    }
    k++;
  }
  std::vector<uint32_t>().swap(src);
  std::vector<uint32_t>().swap(dst);
  gb_status st = build_csr(V, s2, d2, weight_mode ? &w2 : nullptr, out);
  if (st != GB_OK) return gb_fail(st, "gb_gen_rmat: CSR build failed (E >= 2^31?)");
  return GB_OK;
}

gb_status gb_gen_er(gb_host_csr *out, int64_t n, double avg_degree,
                    uint64_t seed, int32_t weight_mode, uint32_t wlo,
                    uint32_t whi) {
  if (!out || n < 2 || avg_degree < 0 || whi < wlo)
    return gb_fail(GB_ERR_INVALID, "gb_gen_er: bad arguments");
  const int64_t M = (int64_t)std::llround(avg_degree * (double)n);
  std::vector<uint32_t> src(M), dst(M), wt(weight_mode ? M : 0);
  parallel_for(M, [&](int64_t lo, int64_t hi, int) {
    for (int64_t e = lo; e < hi; e++) {
      Stream r(seed ^ 0xE5E5E5E5ull, (uint64_t)e);
      uint64_t u = r.below((uint64_t)n);
      uint64_t v = r.below((uint64_t)n - 1);
      if (v >= u) v++;
      src[e] = (uint32_t)u;
      dst[e] = (uint32_t)v;
      if (weight_mode) wt[e] = draw_weight(r, weight_mode, wlo, whi);
    }
  });
  gb_status st = build_csr(n, src, dst, weight_mode ? &wt : nullptr, out);
  if (st != GB_OK) return gb_fail(st, "gb_gen_er: CSR build failed");
  return GB_OK;
}

gb_status gb_transpose(const gb_host_csr *in, gb_host_csr *out) {
  if (!in || !out) return gb_fail(GB_ERR_INVALID, "gb_transpose: null");
  const int64_t V = in->V, E = in->E;
  std::vector<uint32_t> src(E), dst(E), w(in->w ? E : 0);
  parallel_for(V, [&](int64_t lo, int64_t hi, int) {
    for (int64_t v = lo; v < hi; v++)
      for (int32_t p = in->row_ptr[v]; p < in->row_ptr[v + 1]; p++) {
        src[p] = (uint32_t)in->col[p];
        dst[p] = (uint32_t)v;
        if (in->w) w[p] = in->w[p];
      }
  });
  return build_csr(V, src, dst, in->w ? &w : nullptr, out);
}

gb_status gb_pick_sources(const gb_host_csr *g, int32_t k, uint64_t seed,
                          int32_t *out) {
  if (!g || !out || k < 0) return gb_fail(GB_ERR_INVALID, "gb_pick_sources");
  std::vector<int32_t> cand;
  for (int32_t v = 0; v < g->V; v++)
    if (g->row_ptr[v + 1] > g->row_ptr[v]) cand.push_back(v);
  if ((int64_t)cand.size() < k)
    return gb_fail(GB_ERR_INVALID, "gb_pick_sources: not enough candidates");
  Stream r(seed ^ 0x5005CE5ull, 0);
  std::vector<uint8_t> used(cand.size(), 0);
  for (int32_t i = 0; i < k; i++) {
    uint64_t j;
    do {
      j = r.below(cand.size());
    } while (used[j]);
    used[j] = 1;
    out[i] = cand[j];
  }
  return GB_OK;
}

}  // extern "C"
