/*
 * oracle.c -- CPU restatement of the reference's CSR analytics algorithms.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP path is tested against, and the CPU baseline
 * timed by bench.py.  Nothing in graphbench_amd/ links or calls this file.
 *
 * Reference citations are file:line in stplaydog/GRAPHBENCH.
 * Parity pins for every function are listed in tests/test_oracle.py (closed
 * forms, scipy/networkx cross-checks, brute force, reference golden vectors).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void orc_set_threads(int32_t n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

int32_t orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* binary heap of (key, vertex), lazy deletion                               */
typedef struct {
  double k;
  int32_t v;
} hnode;
typedef struct {
  hnode *a;
  int64_t n, cap;
} heap;

static void hpush(heap *h, double k, int32_t v) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->a = (hnode *)realloc(h->a, sizeof(hnode) * h->cap);
  }
  int64_t i = h->n++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h->a[p].k <= k) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i].k = k;
  h->a[i].v = v;
}

static hnode hpop(heap *h) {
  hnode top = h->a[0];
  hnode last = h->a[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && h->a[c + 1].k < h->a[c].k) c++;
    if (last.k <= h->a[c].k) break;
    h->a[i] = h->a[c];
    i = c;
  }
  if (h->n > 0) h->a[i] = last;
  return top;
}

/* run_djkstra, sssp/csr_sspath.c:5-36: pop the minimum tentative distance,
 * relax its out-edges with strict '<'.  Heap instead of the list pop_min. */
int orc_dijkstra_u32(int32_t V, const int32_t *rp, const int32_t *col,
                     const uint32_t *w, int32_t src, uint32_t *dist) {
  int overflow = 0;
  for (int32_t i = 0; i < V; i++) dist[i] = ORC_UINF;
  if (src < 0 || src >= V) return -1;
  heap h = {0};
  dist[src] = 0;
  hpush(&h, 0.0, src);
  while (h.n > 0) {
    hnode t = hpop(&h);
    uint32_t du = dist[t.v];
    if ((double)du != t.k) continue; /* stale entry */
    for (int32_t e = rp[t.v]; e < rp[t.v + 1]; e++) {
      uint64_t nd = (uint64_t)du + w[e];
      if (nd >= ORC_UINF) {
        overflow = 1;
        nd = ORC_UINF - 1;
      }
      int32_t v = col[e];
      if (nd < dist[v]) {
        dist[v] = (uint32_t)nd;
        hpush(&h, (double)nd, v);
      }
    }
  }
  free(h.a);
  return overflow;
}

int orc_dijkstra_f64(int32_t V, const int32_t *rp, const int32_t *col,
                     const double *w, int32_t src, double *dist) {
  for (int32_t i = 0; i < V; i++) dist[i] = INFINITY;
  if (src < 0 || src >= V) return -1;
  heap h = {0};
  dist[src] = 0.0;
  hpush(&h, 0.0, src);
  while (h.n > 0) {
    hnode t = hpop(&h);
    double du = dist[t.v];
    if (du != t.k) continue;
    for (int32_t e = rp[t.v]; e < rp[t.v + 1]; e++) {
      double nd = du + w[e];
      int32_t v = col[e];
      if (nd < dist[v]) {
        dist[v] = nd;
        hpush(&h, nd, v);
      }
    }
  }
  free(h.a);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* growable bucket array: bucket b holds vertex ids (duplicates allowed)      */
typedef struct {
  int32_t *a;
  int64_t n, cap;
} ivec;
static void ipush(ivec *v, int32_t x) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 16;
    v->a = (int32_t *)realloc(v->a, sizeof(int32_t) * v->cap);
  }
  v->a[v->n++] = x;
}
typedef struct {
  ivec *b;
  int64_t nb;
} buckets;
static ivec *bucket_at(buckets *B, int64_t i) {
  if (i >= B->nb) {
    int64_t nn = B->nb ? B->nb : 64;
    while (nn <= i) nn *= 2;
    B->b = (ivec *)realloc(B->b, sizeof(ivec) * nn);
    memset(B->b + B->nb, 0, sizeof(ivec) * (nn - B->nb));
    B->nb = nn;
  }
  return &B->b[i];
}
static void buckets_free(buckets *B) {
  for (int64_t i = 0; i < B->nb; i++) free(B->b[i].a);
  free(B->b);
}

/* run_sspath_arr, sssp/csr_sspath_array.c:149-363.
 *  - light/heavy: separate_csr (graph/csr.c:85, 118): light iff w <= threshold.
 *  - inner loop 240-300: request light edges of B[iter] (request_arr 872-939),
 *    move B[iter] into S (remember_arr 1044-1098), relax (relax_arr 625-826).
 *  - heavy phase 310-331 over S, then empty S and go to the next bucket.
 *  - bucket of a distance d is floor(d / delta) (relax_arr 672).
 * Fixes: all requests relaxed with min (A5), unbounded buckets (A8),
 * bucket iterator local (A13), settled re-insertions tracked (A6). */
int orc_delta_stepping_u32(int32_t V, const int32_t *rp, const int32_t *col,
                           const uint32_t *w, int32_t src, uint32_t delta,
                           uint32_t light_max, uint32_t *dist, int64_t *stats) {
  int overflow = 0;
  int64_t st_buckets = 0, st_phases = 0, st_relax = 0, st_impr = 0;
  if (delta == 0) delta = 1;
  for (int32_t i = 0; i < V; i++) dist[i] = ORC_UINF;
  if (src < 0 || src >= V) return -1;
  buckets B = {0};
  uint32_t *sstamp = (uint32_t *)calloc(V, sizeof(uint32_t)); /* in S of round */
  uint32_t *lstamp = (uint32_t *)calloc(V, sizeof(uint32_t)); /* dedupe in L */
  uint32_t round = 0, lround = 0;
  ivec S = {0}, L = {0};
  dist[src] = 0;
  ipush(bucket_at(&B, 0), src);
  int64_t cur = 0;
  for (;;) {
    while (cur < B.nb && B.b[cur].n == 0) cur++;
    if (cur >= B.nb) break;
    round++;
    S.n = 0;
    st_buckets++;
    for (;;) {
      while (B.b[cur].n > 0) {
        /* take the bucket (request_arr reads it, remember_arr empties it) */
        ivec *bk = bucket_at(&B, cur);
        ivec tmp = *bk;
        *bk = L;
        L = tmp;
        bk->n = 0;
        lround++;
        st_phases++;
        int64_t nl = 0;
        for (int64_t k = 0; k < L.n; k++) {
          int32_t u = L.a[k];
          if ((int64_t)(dist[u] / delta) != cur) continue; /* moved away */
          if (lstamp[u] == lround) continue;
          lstamp[u] = lround;
          L.a[nl++] = u;
          if (sstamp[u] != round) {
            sstamp[u] = round;
            ipush(&S, u);
          }
        }
        L.n = nl;
        /* requests: (u, v, w) for light edges; relax reads dist[u] at relax time */
        for (int64_t k = 0; k < L.n; k++) {
          int32_t u = L.a[k];
          for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
            if (w[e] > light_max) continue;
            uint64_t nd = (uint64_t)dist[u] + w[e];
            if (nd >= ORC_UINF) {
              overflow = 1;
              nd = ORC_UINF - 1;
            }
            int32_t v = col[e];
            st_relax++;
            if (nd < dist[v]) {
              dist[v] = (uint32_t)nd;
              st_impr++;
              ipush(bucket_at(&B, (int64_t)(nd / delta)), v);
            }
          }
        }
      }
      /* heavy phase over S (csr_sspath_array.c:310-331) */
      for (int64_t k = 0; k < S.n; k++) {
        int32_t u = S.a[k];
        for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
          if (w[e] <= light_max) continue;
          uint64_t nd = (uint64_t)dist[u] + w[e];
          if (nd >= ORC_UINF) {
            overflow = 1;
            nd = ORC_UINF - 1;
          }
          int32_t v = col[e];
          st_relax++;
          if (nd < dist[v]) {
            dist[v] = (uint32_t)nd;
            st_impr++;
            ipush(bucket_at(&B, (int64_t)(nd / delta)), v);
          }
        }
      }
      S.n = 0;
      /* a heavy edge shorter than delta can land back in this bucket */
      if (B.b[cur].n == 0) break;
      round++;
    }
  }
  if (stats) {
    stats[0] = st_buckets;
    stats[1] = st_phases;
    stats[2] = st_relax;
    stats[3] = st_impr;
  }
  buckets_free(&B);
  free(S.a);
  free(L.a);
  free(sstamp);
  free(lstamp);
  return overflow;
}

/* run_sspath, sssp/csr_sspath.c:38-122 on f64 weights: light phase on w<=delta
 * then heavy phase, bucket (lint)(d/delta) (lines 104-113). */
int orc_delta_stepping_f64(int32_t V, const int32_t *rp, const int32_t *col,
                           const double *w, int32_t src, double delta,
                           double *dist) {
  for (int32_t i = 0; i < V; i++) dist[i] = INFINITY;
  if (src < 0 || src >= V || !(delta > 0)) return -1;
  buckets B = {0};
  uint32_t *sstamp = (uint32_t *)calloc(V, sizeof(uint32_t));
  uint32_t round = 0;
  ivec S = {0}, L = {0};
  dist[src] = 0.0;
  ipush(bucket_at(&B, 0), src);
  int64_t cur = 0;
  for (;;) {
    while (cur < B.nb && B.b[cur].n == 0) cur++;
    if (cur >= B.nb) break;
    round++;
    S.n = 0;
    for (;;) {
      while (B.b[cur].n > 0) {
        ivec *bk = bucket_at(&B, cur);
        ivec tmp = *bk;
        *bk = L;
        L = tmp;
        bk->n = 0;
        for (int64_t k = 0; k < L.n; k++) {
          int32_t u = L.a[k];
          if ((int64_t)(dist[u] / delta) != cur) continue;
          if (sstamp[u] != round) {
            sstamp[u] = round;
            ipush(&S, u);
          }
          for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
            if (!(w[e] <= delta)) continue;
            double nd = dist[u] + w[e];
            int32_t v = col[e];
            if (nd < dist[v]) {
              dist[v] = nd;
              ipush(bucket_at(&B, (int64_t)(nd / delta)), v);
            }
          }
        }
      }
      for (int64_t k = 0; k < S.n; k++) {
        int32_t u = S.a[k];
        for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
          if (w[e] <= delta) continue;
          double nd = dist[u] + w[e];
          int32_t v = col[e];
          if (nd < dist[v]) {
            dist[v] = nd;
            ipush(bucket_at(&B, (int64_t)(nd / delta)), v);
          }
        }
      }
      S.n = 0;
      if (B.b[cur].n == 0) break;
      round++;
    }
  }
  buckets_free(&B);
  free(S.a);
  free(L.a);
  free(sstamp);
  return 0;
}

/* ------------------------------------------------------------------------ */
void orc_bfs_levels(int32_t V, const int32_t *rp, const int32_t *col,
                    int32_t src, int32_t *level) {
  for (int32_t i = 0; i < V; i++) level[i] = -1;
  if (src < 0 || src >= V) return;
  int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(V > 0 ? V : 1));
  int64_t h = 0, t = 0;
  q[t++] = src;
  level[src] = 0;
  while (h < t) {
    int32_t u = q[h++];
    for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
      int32_t v = col[e];
      if (level[v] < 0) {
        level[v] = level[u] + 1;
        q[t++] = v;
      }
    }
  }
  free(q);
}

/* Brandes, btwcentral/csr_bc.c:16-130 on BFS levels.
 * comp_num_spath (70-96): sigma[w] = sum of sigma over predecessors.
 * comp_delta (98-123): delta[v] += ((1+delta[w])*sigma[v])/sigma[w].
 * run_bc (57-61): bc[j] += delta[j] for j != source. */
/* One source of orc_brandes_bfs: delta[] (V entries) for source s, using the
 * caller's scratch arrays. */
static void brandes_bfs_one(int32_t V, const int32_t *rp, const int32_t *col,
                            const int32_t *rrp, const int32_t *rcol, int32_t s,
                            int32_t *level, int32_t *order, double *sigma,
                            double *delta) {
  for (int32_t i = 0; i < V; i++) {
    level[i] = -1;
    sigma[i] = 0.0;
    delta[i] = 0.0;
  }
  int64_t h = 0, t = 0;
  order[t++] = s;
  level[s] = 0;
  while (h < t) {
    int32_t u = order[h++];
    for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
      int32_t v = col[e];
      if (level[v] < 0) {
        level[v] = level[u] + 1;
        order[t++] = v;
      }
    }
  }
  sigma[s] = 1.0;
  for (int64_t k = 1; k < t; k++) {
    int32_t v = order[k];
    double acc = 0.0;
    for (int32_t e = rrp[v]; e < rrp[v + 1]; e++) {
      int32_t u = rcol[e];
      if (level[u] == level[v] - 1) acc += sigma[u];
    }
    sigma[v] = acc;
  }
  for (int64_t k = t - 1; k >= 0; k--) {
    int32_t v = order[k];
    double acc = 0.0;
    for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
      int32_t x = col[e];
      if (level[x] == level[v] + 1)
        acc += ((1.0 + delta[x]) * sigma[v]) / sigma[x];
    }
    delta[v] = acc;
  }
}

/* Brandes, btwcentral/csr_bc.c:16-130 on BFS levels.
 * comp_num_spath (70-96): sigma[w] = sum of sigma over predecessors.
 * comp_delta (98-123): delta[v] += ((1+delta[w])*sigma[v])/sigma[w].
 * run_bc (57-61): bc[j] += delta[j] for j != source.
 * Sources are solved in parallel, a batch of one per thread, and each
 * batch's delta vectors are added to bc in source order, so the result does
 * not depend on the thread count. */
void orc_brandes_bfs(int32_t V, const int32_t *rp, const int32_t *col,
                     const int32_t *rrp, const int32_t *rcol,
                     const int32_t *sources, int32_t nsrc, double *bc) {
  const size_t n = (size_t)(V ? V : 1);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  if (nt > nsrc) nt = nsrc > 0 ? nsrc : 1;
  int32_t *level = (int32_t *)malloc(sizeof(int32_t) * n * nt);
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n * nt);
  double *sigma = (double *)malloc(sizeof(double) * n * nt);
  double *delta = (double *)malloc(sizeof(double) * n * nt);
  int32_t *ok = (int32_t *)malloc(sizeof(int32_t) * (size_t)nt);
  for (int32_t i = 0; i < V; i++) bc[i] = 0.0;
  for (int32_t b0 = 0; b0 < nsrc; b0 += nt) {
    const int nb = nsrc - b0 < nt ? nsrc - b0 : nt;
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int k = 0; k < nb; k++) {
      const int32_t s = sources[b0 + k];
      ok[k] = s >= 0 && s < V;
      if (ok[k])
        brandes_bfs_one(V, rp, col, rrp, rcol, s, level + n * k, order + n * k,
                        sigma + n * k, delta + n * k);
    }
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < V; j++) {
      double acc = bc[j];
      for (int k = 0; k < nb; k++)
        if (ok[k] && j != sources[b0 + k]) acc += delta[n * k + j];
      bc[j] = acc;
    }
  }
  free(level);
  free(order);
  free(sigma);
  free(delta);
  free(ok);
}

/* Weighted Brandes: the run_bc loop (csr_bc.c:40-63) over exact distances;
 * predecessors are the tight in-edges (csr_sspath.c:101-118 with '<=' ties),
 * each edge counted once. Vertices are ordered by tight-DAG depth. */
void orc_brandes_weighted_f64(int32_t V, const int32_t *rp, const int32_t *col,
                              const double *w, const int32_t *rrp,
                              const int32_t *rcol, const double *rw,
                              const int32_t *sources, int32_t nsrc,
                              double *bc) {
  size_t n = (size_t)(V ? V : 1);
  double *dist = (double *)malloc(sizeof(double) * n);
  double *sigma = (double *)malloc(sizeof(double) * n);
  double *delta = (double *)malloc(sizeof(double) * n);
  int32_t *depth = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t *indeg = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n);
  for (int32_t i = 0; i < V; i++) bc[i] = 0.0;
  for (int32_t si = 0; si < nsrc; si++) {
    int32_t s = sources[si];
    if (s < 0 || s >= V) continue;
    orc_dijkstra_f64(V, rp, col, w, s, dist);
    /* Kahn over the tight DAG gives a depth order */
    for (int32_t v = 0; v < V; v++) {
      int32_t c = 0;
      if (dist[v] != INFINITY)
        for (int32_t e = rrp[v]; e < rrp[v + 1]; e++) {
          int32_t u = rcol[e];
          if (dist[u] != INFINITY && dist[u] + rw[e] == dist[v]) c++;
        }
      indeg[v] = c;
      depth[v] = -1;
      sigma[v] = 0.0;
      delta[v] = 0.0;
    }
    int64_t h = 0, t = 0;
    order[t++] = s;
    depth[s] = 0;
    while (h < t) {
      int32_t u = order[h++];
      for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
        int32_t v = col[e];
        if (dist[u] + w[e] == dist[v] && v != s) {
          if (--indeg[v] == 0) {
            depth[v] = depth[u] + 1;
            order[t++] = v;
          }
        }
      }
    }
    sigma[s] = 1.0;
    for (int64_t k = 1; k < t; k++) {
      int32_t v = order[k];
      double acc = 0.0;
      for (int32_t e = rrp[v]; e < rrp[v + 1]; e++) {
        int32_t u = rcol[e];
        if (dist[u] != INFINITY && dist[u] + rw[e] == dist[v]) acc += sigma[u];
      }
      sigma[v] = acc;
    }
    for (int64_t k = t - 1; k >= 0; k--) {
      int32_t v = order[k];
      double acc = 0.0;
      for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
        int32_t x = col[e];
        if (x != s && dist[v] + w[e] == dist[x])
          acc += ((1.0 + delta[x]) * sigma[v]) / sigma[x];
      }
      delta[v] = acc;
    }
    for (int32_t j = 0; j < V; j++)
      if (j != s) bc[j] += delta[j];
  }
  free(dist);
  free(sigma);
  free(delta);
  free(depth);
  free(indeg);
  free(order);
}

/* ------------------------------------------------------------------------ */
/* pagerank/pagerank_mpi.c with one rank (p=1: every in-edge is local). */
int orc_pagerank_ref(int32_t V, const int32_t *irp, const int32_t *icol,
                     const double *recip, int32_t max_iter, double *x) {
  const double rand_jump = 0.3;              /* RDM_JMP, graph.h:17 */
  const double purpose_jump = 1.0 - 0.3;     /* pagerank_mpi.c:689 */
  const double thresh = 0.000001;            /* THRESH, graph.h:18 */
  size_t n = (size_t)(V ? V : 1);
  double *diff = (double *)malloc(sizeof(double) * n);
  double *wb = (double *)malloc(sizeof(double) * n);
  int32_t i;
  for (i = 0; i < V; i++) x[i] = 1.0;                 /* init_weight 63-68 */
  for (i = 0; i < V; i++) diff[i] = x[i] * recip[i];  /* 752-754 */
  for (i = 0; i < V; i++) wb[i] = rand_jump;          /* 808-809 */
  int32_t iter = 0;
  int done = 0;
  while (!done) {
    if (max_iter > 0 && iter >= max_iter) {
      iter = -1;
      break;
    }
    iter++;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t v = 0; v < V; v++) {
      double acc = wb[v];
      for (int32_t e = irp[v]; e < irp[v + 1]; e++)
        acc += purpose_jump * diff[icol[e]]; /* 833-846 */
      wb[v] = acc;
    }
    int term = 1;
#pragma omp parallel for reduction(&& : term)
    for (int32_t v = 0; v < V; v++) {
      diff[v] = (wb[v] - x[v]) * recip[v]; /* 995-997 */
      if (fabs(diff[v]) > thresh) term = 0; /* 999-1000 */
      x[v] = wb[v];                         /* 1001 */
    }
    done = term;
  }
  free(diff);
  free(wb);
  return iter;
}

/* Sum of f(i) over [0, n) in fixed chunks of ORC_CHUNK: chunk partials in
 * parallel, then added in chunk order, so the result does not depend on the
 * thread count. */
#define ORC_CHUNK 65536

int orc_pagerank_cfg(int32_t V, const int32_t *irp, const int32_t *icol,
                     const int32_t *outdeg, double d, double tol,
                     int32_t max_iter, double *x, double *resid) {
  size_t n = (size_t)(V ? V : 1);
  double *contrib = (double *)malloc(sizeof(double) * n);
  double *xn = (double *)malloc(sizeof(double) * n);
  const int64_t nch = ((int64_t)V + ORC_CHUNK - 1) / ORC_CHUNK;
  double *part = (double *)malloc(sizeof(double) * (size_t)(nch ? nch : 1));
  const double N = (double)V;
  for (int32_t i = 0; i < V; i++) x[i] = 1.0 / N;
  int32_t it = 0;
  double r = INFINITY;
  while (it < max_iter) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nch; c++) {
      const int32_t lo = (int32_t)(c * ORC_CHUNK);
      const int32_t hi = (int32_t)((c + 1) * ORC_CHUNK < V ? (c + 1) * ORC_CHUNK : V);
      double dg = 0.0;
      for (int32_t u = lo; u < hi; u++) {
        if (outdeg[u] == 0) {
          dg += x[u];
          contrib[u] = 0.0;
        } else {
          contrib[u] = x[u] / (double)outdeg[u];
        }
      }
      part[c] = dg;
    }
    double dangling = 0.0;
    for (int64_t c = 0; c < nch; c++) dangling += part[c];
    const double base = (1.0 - d) / N + d * dangling / N;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t v = 0; v < V; v++) {
      double s = 0.0;
      for (int32_t e = irp[v]; e < irp[v + 1]; e++) s += contrib[icol[e]];
      xn[v] = base + d * s;
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nch; c++) {
      const int32_t lo = (int32_t)(c * ORC_CHUNK);
      const int32_t hi = (int32_t)((c + 1) * ORC_CHUNK < V ? (c + 1) * ORC_CHUNK : V);
      double acc = 0.0;
      for (int32_t v = lo; v < hi; v++) {
        acc += fabs(xn[v] - x[v]);
        x[v] = xn[v];
      }
      part[c] = acc;
    }
    double acc = 0.0;
    for (int64_t c = 0; c < nch; c++) acc += part[c];
    r = acc;
    it++;
    if (r < tol) break;
  }
  if (resid) *resid = r;
  free(contrib);
  free(xn);
  free(part);
  return it;
}

/* ------------------------------------------------------------------------ */
/* comp_distrib_csr, partition/csr_partition.c:5-30 (SEND form). */
void orc_comp_distrib(int64_t V, const int64_t *deg, int64_t E, int32_t p,
                      int64_t *distrib) {
  int64_t distrib_num = E / p;
  int64_t *size = (int64_t *)calloc((size_t)p + 1, sizeof(int64_t));
  int64_t j = 0;
  for (int32_t k = 0; k < p; k++) distrib[k] = 0;
  for (int64_t i = 0; i < V; i++) {
    size[j] += deg[i];
    if (size[j] > distrib_num) {
      distrib[j] = i;
      j++;
      if (j > p) j = p; /* cannot happen (sum of p parts > E); guard only */
    }
  }
  distrib[p - 1] = V - 1;
  free(size);
}

/* bin_search, utils/utils.c:222-239 (same branch structure, iterative). */
int64_t orc_bin_search(const int64_t *distrib, int64_t val, int64_t start,
                       int64_t end) {
  for (;;) {
    int64_t mid = start + (end - start) / 2;
    if (start == end) return start;
    if (start == end - 1) return val <= distrib[start] ? start : end;
    if (val <= distrib[mid] && val > distrib[mid - 1]) return mid;
    if (val > distrib[mid]) {
      start = mid;
      continue;
    }
    if (val <= distrib[mid - 1]) {
      end = mid;
      continue;
    }
    return -1;
  }
}

/* Merge path over A = row ends (rp[1..V]) and B = edge indices 0..E-1:
 * row end i is taken before edge j iff rp[i+1] <= j. */
void orc_merge_path_split(int32_t V, const int32_t *rp, int64_t items,
                          int64_t ntiles, int32_t *tile_row,
                          int32_t *tile_edge) {
  int64_t E = rp[V];
  int64_t total = (int64_t)V + E;
  for (int64_t t = 0; t <= ntiles; t++) {
    int64_t d = t * items;
    if (d > total) d = total;
    int64_t lo = d - E > 0 ? d - E : 0;
    int64_t hi = d < V ? d : V;
    while (lo < hi) {
      int64_t mid = (lo + hi) / 2;
      if ((int64_t)rp[mid + 1] <= d - mid - 1)
        lo = mid + 1;
      else
        hi = mid;
    }
    tile_row[t] = (int32_t)lo;
    tile_edge[t] = (int32_t)(d - lo);
  }
}

/* separate_csr, graph/csr.c:75-130 (degree counters zeroed: defect A1). */
void orc_separate_csr(int32_t V, const int32_t *rp, const int32_t *col,
                      const uint32_t *w, uint32_t light_max, int32_t *lrp,
                      int32_t *lcol, uint32_t *lw, int32_t *hrp, int32_t *hcol,
                      uint32_t *hw) {
  int32_t nl = 0, nh = 0;
  lrp[0] = hrp[0] = 0;
  for (int32_t v = 0; v < V; v++) {
    for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
      if (w[e] <= light_max) {
        lcol[nl] = col[e];
        lw[nl++] = w[e];
      } else {
        hcol[nh] = col[e];
        hw[nh++] = w[e];
      }
    }
    lrp[v + 1] = nl;
    hrp[v + 1] = nh;
  }
}

/* chunk_files (partition/csr_partition.c:99-108): send_id and recv_id by
 * bin_search(distrib, x, 0, num_nodes-1), send_ub[send_id][recv_id]++ when
 * they differ. */
void orc_edge_parts(int64_t E, const int32_t *src, const int32_t *dst,
                    const int64_t *distrib, int32_t p, int32_t *sp, int32_t *rp,
                    int64_t *ub) {
  for (int64_t k = 0; k < (int64_t)p * p; k++) ub[k] = 0;
  for (int64_t e = 0; e < E; e++) {
    const int64_t s = orc_bin_search(distrib, src[e], 0, p - 1);
    const int64_t r = orc_bin_search(distrib, dst[e], 0, p - 1);
    sp[e] = (int32_t)s;
    rp[e] = (int32_t)r;
    if (s != r) ub[s * p + r]++;
  }
}

int64_t orc_remove_duplicate(int64_t E, const int64_t *pairs, int64_t *out) {
  /* open-addressing set of seen (from, to) pairs; slot holds index + 1 */
  int64_t cap = 16;
  while (cap < 2 * E + 2) cap <<= 1;
  int64_t *slot = (int64_t *)calloc((size_t)cap, sizeof(int64_t));
  int64_t kept = 0;
  for (int64_t i = 0; i < E; i++) {
    const int64_t a = pairs[2 * i], b = pairs[2 * i + 1];
    uint64_t h = (uint64_t)a * 0x9E3779B97F4A7C15ull ^ ((uint64_t)b + 0x632BE59BD9B4E019ull);
    h ^= h >> 29;
    int64_t s = (int64_t)(h & (uint64_t)(cap - 1));
    int found = 0;
    while (slot[s]) {
      const int64_t j = slot[s] - 1;
      if (pairs[2 * j] == a && pairs[2 * j + 1] == b) {
        found = 1;
        break;
      }
      s = (s + 1) & (cap - 1);
    }
    if (!found) {
      slot[s] = i + 1;
      out[2 * kept] = a;
      out[2 * kept + 1] = b;
      kept++;
    }
  }
  free(slot);
  return kept;
}

void orc_separate_csr_f64(int32_t V, const int32_t *rp, const int32_t *col,
                          const double *w, double delta, int32_t *lrp, int32_t *lcol,
                          double *lw, int32_t *hrp, int32_t *hcol, double *hw) {
  int32_t nl = 0, nh = 0;
  lrp[0] = hrp[0] = 0;
  for (int32_t v = 0; v < V; v++) {
    for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
      if (w[e] <= delta) { /* csr.c:85, 118 */
        lcol[nl] = col[e];
        lw[nl++] = w[e];
      } else {
        hcol[nh] = col[e];
        hw[nh++] = w[e];
      }
    }
    lrp[v + 1] = nl;
    hrp[v + 1] = nh;
  }
}

/* ------------------------------------------------------------------------ */
double orc_mlp_loss_grad(int32_t n, int32_t din, int32_t h, int32_t dout,
                         const float *X, const float *T, const float *W1,
                         const float *b1, const float *W2, const float *b2,
                         const float *W3, const float *b3, double *gW1,
                         double *gb1, double *gW2, double *gb2, double *gW3,
                         double *gb3) {
  double *a1 = (double *)malloc(sizeof(double) * h);
  double *a2 = (double *)malloc(sizeof(double) * h);
  double *z1 = (double *)malloc(sizeof(double) * h);
  double *z2 = (double *)malloc(sizeof(double) * h);
  double *g1 = (double *)malloc(sizeof(double) * h);
  double *g2 = (double *)malloc(sizeof(double) * h);
  double *y = (double *)malloc(sizeof(double) * dout);
  memset(gW1, 0, sizeof(double) * h * din);
  memset(gb1, 0, sizeof(double) * h);
  memset(gW2, 0, sizeof(double) * h * h);
  memset(gb2, 0, sizeof(double) * h);
  memset(gW3, 0, sizeof(double) * dout * h);
  memset(gb3, 0, sizeof(double) * dout);
  double loss = 0.0;
  for (int32_t r = 0; r < n; r++) {
    const float *x = X + (size_t)r * din;
    const float *t = T + (size_t)r * dout;
    for (int32_t j = 0; j < h; j++) {
      double s = b1[j];
      for (int32_t k = 0; k < din; k++) s += (double)W1[j * din + k] * x[k];
      a1[j] = s;
      z1[j] = s > 0 ? s : 0;
    }
    for (int32_t j = 0; j < h; j++) {
      double s = b2[j];
      for (int32_t k = 0; k < h; k++) s += (double)W2[j * h + k] * z1[k];
      a2[j] = s;
      z2[j] = s > 0 ? s : 0;
    }
    for (int32_t o = 0; o < dout; o++) {
      double s = b3[o];
      for (int32_t k = 0; k < h; k++) s += (double)W3[o * h + k] * z2[k];
      y[o] = s;
      double e = s - t[o];
      loss += 0.5 * e * e;
      y[o] = e / n; /* dL/dy */
    }
    for (int32_t o = 0; o < dout; o++) {
      gb3[o] += y[o];
      for (int32_t k = 0; k < h; k++) gW3[o * h + k] += y[o] * z2[k];
    }
    for (int32_t k = 0; k < h; k++) {
      double s = 0;
      for (int32_t o = 0; o < dout; o++) s += (double)W3[o * h + k] * y[o];
      g2[k] = a2[k] > 0 ? s : 0;
    }
    for (int32_t j = 0; j < h; j++) {
      gb2[j] += g2[j];
      for (int32_t k = 0; k < h; k++) gW2[j * h + k] += g2[j] * z1[k];
    }
    for (int32_t k = 0; k < h; k++) {
      double s = 0;
      for (int32_t j = 0; j < h; j++) s += (double)W2[j * h + k] * g2[j];
      g1[k] = a1[k] > 0 ? s : 0;
    }
    for (int32_t j = 0; j < h; j++) {
      gb1[j] += g1[j];
      for (int32_t k = 0; k < din; k++) gW1[j * din + k] += g1[j] * x[k];
    }
  }
  free(a1);
  free(a2);
  free(z1);
  free(z2);
  free(g1);
  free(g2);
  free(y);
  return n > 0 ? loss / n : 0.0;
}
