/*
 * graphbench.h -- native C-ABI of the MI355X CSR analytics library
 * (libgraphbench.so).  Plain C types only: device pointers are passed as
 * pointers, streams as void* (a hipStream_t, NULL = the default stream).
 *
 * Calls share per-process device workspaces (queues, scratch arrays) and
 * cache derived structures on the graph handle: issue them from one host
 * thread at a time.  Results are ready when a call returns.
 *
 * Graph layout on the device (one graph):
 *   row_ptr[V+1]  int32   CSR offsets, row_ptr[0] = 0, E = row_ptr[V] < 2^31
 *   col[E]        int32   destination ids (source ids for the reverse CSR)
 *   w[E]          uint32  edge weights (SSSP only; may be NULL otherwise)
 * The reference's t_csr keeps end offsets vet_idx[V] (graph/csr.c:194-200);
 * row_ptr[i+1] == vet_idx[i].  The legacy t_csr entry points live in
 * gb_legacy.h and convert into this layout.
 *
 * Which reference function each call replaces:
 *   gb_sssp            run_sspath_arr   sssp/csr_sspath_array.c:149-363
 *                      run_sspath       sssp/csr_sspath.c:38-83 (distances)
 *   gb_pagerank        run_pagerank_csr_mpi local gather + update,
 *                      pagerank/pagerank_mpi.c:833-846, 995-1020 (config form)
 *   gb_pagerank_ref    the same loop with the reference constants, FP64
 *   gb_bfs / gb_bc     run_bc, btwcentral/csr_bc.c:16-130 (unweighted form)
 *   gb_bc_weighted     run_bc on weighted graphs (csr_bc.c:40-63 + run_sspath)
 *   gb_merge_path_split comp_distrib_csr partition/csr_partition.c:5-30 and
 *                      bin_search utils/utils.c:222-239 (edge-balanced split)
 *   gb_separate        separate_csr graph/csr.c:75-130 (light/heavy split)
 */
#ifndef GRAPHBENCH_H
#define GRAPHBENCH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gb_status {
  GB_OK = 0,
  GB_ERR_INVALID = 1,       /* bad argument or graph */
  GB_ERR_NOMEM = 2,         /* device allocation failed or above gb_set_mem_cap */
  GB_ERR_HIP = 3,           /* HIP runtime error (see gb_last_error) */
  GB_ERR_OVERFLOW = 4,      /* uint32 distance overflow */
  GB_ERR_NOT_CONVERGED = 5, /* iteration cap reached (result still written) */
  GB_ERR_NO_DEVICE = 6,     /* no HIP device */
  GB_ERR_UNSUPPORTED = 7
} gb_status;

#define GB_DIST_INF 0xFFFFFFFFu

const char *gb_status_string(gb_status s);
/* Text of the last error on this thread (never NULL). */
const char *gb_last_error(void);
/* Select the device and create per-device state; idempotent. */
gb_status gb_init(int32_t device);
/* Cap on bytes the library itself allocates (0 = no cap). */
void gb_set_mem_cap(uint64_t bytes);
/* Release cached workspaces. */
void gb_release_workspaces(void);
/* Library version string. */
const char *gb_version(void);

/* ---- graph handles --------------------------------------------------------
 * A handle borrows device arrays (no copy) or owns uploaded copies.  It also
 * caches derived device structures (merge-path tiles, 1/outdeg, light/heavy
 * orderings).  The reverse CSR (in-edges) is needed by PageRank and by the
 * sigma pull of weighted Brandes on directed graphs; for a symmetric graph
 * pass the forward arrays again. */
typedef struct gb_graph_s *gb_graph;
gb_status gb_graph_wrap(gb_graph *out, int32_t V, int64_t E,
                        const int32_t *d_row_ptr, const int32_t *d_col,
                        const uint32_t *d_w, const int32_t *d_rrow_ptr,
                        const int32_t *d_rcol, const uint32_t *d_rw);
gb_status gb_graph_upload(gb_graph *out, int32_t V, int64_t E,
                          const int32_t *row_ptr, const int32_t *col,
                          const uint32_t *w, const int32_t *rrow_ptr,
                          const int32_t *rcol, const uint32_t *rw);
/* f64 weights (reference t_csr edge_weight); stored alongside, used by
 * gb_sssp_f64 / gb_bc_weighted. d_w64/d_rw64 are device pointers. */
gb_status gb_graph_set_f64_weights(gb_graph g, const double *d_w64,
                                   const double *d_rw64);
gb_status gb_graph_destroy(gb_graph g);

/* ---- merge-path split -----------------------------------------------------
 * Splits the merge of row ends and edges into ntiles tiles of `items` items
 * (rows + edges).  tile_row/tile_edge have ntiles+1 entries (device). */
gb_status gb_merge_path_split(int32_t V, const int32_t *d_row_ptr, int64_t items,
                              int64_t ntiles, int32_t *d_tile_row,
                              int32_t *d_tile_edge, void *stream);
/* comp_distrib_csr on the device: p parts, distrib[k] = last vertex of part k
 * (same cut rule as the reference).  d_distrib has p entries (int64). */
gb_status gb_edge_balanced_parts(int32_t V, const int32_t *d_row_ptr,
                                 int32_t p, int64_t *d_distrib, void *stream);
/* The per-edge work of chunk_files (partition/csr_partition.c:99-108):
 * owners of both endpoints of every edge under a 1D vertex partition, and the
 * cross-part edge counts written to the <file>_send_<nodes> file.  d_distrib
 * holds each part's last vertex (comp_distrib_csr's output); an endpoint x
 * (0-based) belongs to part bin_search(distrib, x, 0, p-1) (utils/utils.c:
 * 222-239).  Writes d_send_part[e] (owner of d_src[e]) and d_recv_part[e]
 * (owner of d_dst[e]) when non-NULL, and, when d_send_ub is non-NULL,
 * overwrites its p*p counters with the number of edges whose source part s
 * differs from their target part r, at d_send_ub[s*p + r].  Parts lie in
 * [0, p-1] for any distrib, ascending or not.  1 <= p <= GB_MAX_PARTS. */
#define GB_MAX_PARTS 4096
gb_status gb_edge_parts(int64_t E, const int32_t *d_src, const int32_t *d_dst,
                        const int64_t *d_distrib, int32_t p, int32_t *d_send_part,
                        int32_t *d_recv_part, unsigned long long *d_send_ub,
                        void *stream);

/* ---- CSR construction on the device -----------------------------------------
 * Replace the host CSR builds of the reference's readers (scan_csr_idx
 * graph/csr.c:133-202 counts, read_graph_csr graph/csr.c:4-66 and
 * read_csr_bin graph/csr.c:211-299 scatter in file order).
 * gb_csr_from_edges: CSR of E device edges (d_src[e] -> d_dst[e]), rows in
 * source order, the edges of a row in input order (stable), optional weights
 * of w_bytes 4 (uint32) or 8 (double) carried along (w_bytes 0: none).
 * d_row_ptr has V+1 entries, d_col (and d_w_out) E.  Endpoints outside
 * [0, V) are rejected.  gb_csr_transpose: the reverse CSR of a CSR (rows of
 * the transpose keep the order of their source rows). */
gb_status gb_csr_from_edges(int32_t V, int64_t E, const int32_t *d_src,
                            const int32_t *d_dst, const void *d_w, int32_t w_bytes,
                            int32_t *d_row_ptr, int32_t *d_col, void *d_w_out,
                            void *stream);
gb_status gb_csr_transpose(int32_t V, int64_t E, const int32_t *d_row_ptr,
                           const int32_t *d_col, const void *d_w, int32_t w_bytes,
                           int32_t *d_trow_ptr, int32_t *d_tcol, void *d_tw,
                           void *stream);
/* Replaces the dedup of remove_duplicate (partition/csr_partition_bin_mpi.c:
 * 31-106): keeps the first occurrence of every (d_src[e], d_dst[e]) pair, in
 * input order, writing the kept pairs to d_src_out / d_dst_out (E entries of
 * room) and their number to *h_kept.  Negative ids are rejected (the
 * reference stops on -1).  Synchronises the stream. */
gb_status gb_edges_unique(int64_t E, const int32_t *d_src, const int32_t *d_dst,
                          int32_t *d_src_out, int32_t *d_dst_out, int64_t *h_kept,
                          void *stream);

/* ---- light/heavy split ----------------------------------------------------
 * Stable per-row partition: light edges (w <= light_max) first.  Writes
 * d_col_out/d_w_out (E each) and d_light_end[V] (absolute offset where the
 * heavy edges of the row begin). */
gb_status gb_separate(gb_graph g, uint32_t light_max, int32_t *d_col_out,
                      uint32_t *d_w_out, int32_t *d_light_end, void *stream);

/* ---- SSSP (delta-stepping) --------------------------------------------------*/
typedef struct gb_sssp_params {
  uint32_t delta;     /* bucket width (0 = plain rule max_w / avg_degree) */
  uint32_t light_max; /* light/heavy threshold (0 = delta) */
  int32_t max_phases; /* safety cap on phases (0 = none) */
  int32_t reserved;
} gb_sssp_params;

typedef struct gb_sssp_stats {
  int64_t buckets;      /* lockstep buckets processed */
  int64_t phases;       /* light + heavy + far-scan launches */
  int64_t relax_edges;  /* edge relaxations attempted */
  int64_t reached;      /* (source, vertex) pairs with finite distance */
  int64_t traversed;    /* out-edges of reached vertices, summed over sources */
  double ms;            /* device time of the solve */
  /* relaxation outcomes (the reference's REV_STATS, csr_sspath_array.c:738-759,
   * as totals): relaxations that lowered a distance, with the new distance in
   * the current bucket (re-expanded there) or in a later bucket */
  int64_t improved_near;
  int64_t improved_far;
  /* (vertex, edge) pairs read by expansions: the batched engine reads an
   * edge once for the mask of sources it is expanded for (relax_edges /
   * expanded_edges is the mean number of sources per edge read); the
   * per-source engine reads it once per source (equal to relax_edges) */
  int64_t expanded_edges;
} gb_sssp_stats;

/* Distances from nsrc sources, all processed together in lockstep buckets.
 * d_dist: nsrc*V uint32 (device), row-major by source; GB_DIST_INF when
 * unreached.  Returns GB_ERR_OVERFLOW if some distance exceeds 2^32-2. */
gb_status gb_sssp(gb_graph g, const gb_sssp_params *p, const int32_t *h_sources,
                  int32_t nsrc, uint32_t *d_dist, gb_sssp_stats *st,
                  void *stream);
/* Same on the f64 weights (reference semantics: d_dist f64, +inf unreached).
 * delta is a double here. */
gb_status gb_sssp_f64(gb_graph g, double delta, const int32_t *h_sources,
                      int32_t nsrc, double *d_dist, gb_sssp_stats *st,
                      void *stream);

/* Batched SSSP over many independent graphs: graph i is gs[i] with source
 * sources[i] and bucket width deltas[i] (0 = plain rule), light threshold
 * light_max[i] (0 = delta); distances to d_dist[i] (device, V_i entries each).
 * All graphs advance in one lockstep bucket loop. */
gb_status gb_sssp_multi(int32_t ngraphs, const gb_graph *gs,
                        const int32_t *sources, const uint32_t *deltas,
                        const uint32_t *light_max, uint32_t *const *d_dist,
                        gb_sssp_stats *st, void *stream);

/* ---- PageRank ----------------------------------------------------------------*/
typedef struct gb_pr_params {
  double damping;    /* 0.85 */
  double tol;        /* stop when L1 residual < tol */
  int32_t max_iter;  /* 100 */
  int32_t items_per_thread; /* merge-path items per thread (0 = autotuned) */
  int32_t use_graph; /* capture iteration chunks in a hipGraph (1) or not (0) */
  int32_t time_kernels; /* record per-launch gather times with HIP events */
  const float *d_x0; /* optional warm start (device, V; NULL = 1/N) */
  int32_t relabel;   /* gather in out-degree order: 0 auto (E >= 2^20), 1 on, -1 off;
                        built once per graph handle, results in the caller's ids */
  int32_t row_block; /* gather form: 0 autotuned (merge path, row blocks or
                        binned), 1024 / 2048 / 4096 row blocks of that many
                        edges, -1 merge path only */
  int32_t binned;    /* binned (propagation-blocking) form: 0 left to the
                        autotuner, -1 never, 1..6 that form with (hub chunk,
                        chunks, bins) of (8192, 8192, 8192), (16384, 16384,
                        8192), (8192, 8192, 4096), (16384, 8192, 8192),
                        (20480, 8192, 8192), (16384, 4096, 8192) ids; its
                        row sums are fixed-point sums, independent of the
                        order of adds */
} gb_pr_params;

typedef struct gb_pr_result {
  int32_t iterations;
  int32_t converged;
  double residual;       /* last L1 residual (FP64) */
  double dangling;       /* dangling mass of the result */
  double gather_ms;      /* summed gather-kernel time (time_kernels=1) */
  int32_t gather_launches;
  int32_t tiles;
  int32_t items_per_thread; /* merge-path shape used: tile = 256 x items_per_thread */
  int32_t hot;              /* sources served from LDS (0 = none) */
  int32_t relabeled;        /* 1 if the gather ran in out-degree order */
  int32_t row_block;        /* row-block form used: edges per block (0 = merge path) */
  int32_t binned;           /* binned form used (its gb_pr_params.binned value), 0 = not */
} gb_pr_result;

/* Configured PageRank: in-edge pull gather over the reverse CSR, FP32 ranks
 * with per-row compensated sums, FP64 residual and dangling reductions.
 * d_x: V floats (device) receives the ranks. */
gb_status gb_pagerank(gb_graph g, const gb_pr_params *p, float *d_x,
                      gb_pr_result *res, void *stream);
/* Reference-semantics PageRank (RDM_JMP 0.3, THRESH 1e-6, x0 = 1, delta
 * form, FP64, per-row sums in in-CSR order: bit-identical to the reference's
 * loop).  d_recip: V doubles (device, the reference's recip values).
 * iterations receives the count, or -1 when max_iter is hit. */
gb_status gb_pagerank_ref(gb_graph g, const double *d_recip, int32_t max_iter,
                          double *d_x, int32_t *iterations, void *stream);
/* gb_pagerank_ref without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  The reverse CSR is passed as
 * plain device pointers (d_irp: V+1 int32 row offsets, d_icol: the in-edge
 * sources, in-CSR order).  Scratch (diff, wb, flag) lives in a fixed
 * per-device array of 2 * 8 * GB_NOALLOC_PR_MAX_V + 256 bytes held until the
 * process ends, so V may not exceed GB_NOALLOC_PR_MAX_V (GB_ERR_UNSUPPORTED
 * otherwise).  Results, iteration count and status are those of
 * gb_pagerank_ref on the same graph.  All calls on one device share its
 * array: calls that overlap on one device race. */
#define GB_NOALLOC_PR_MAX_V (1 << 22)
gb_status gb_pagerank_ref_noalloc(int32_t V, const int32_t *d_irp, const int32_t *d_icol,
                                  const double *d_recip, int32_t max_iter, double *d_x,
                                  int32_t *iterations, void *stream);

/* ---- BFS / Brandes ---------------------------------------------------------*/
/* BFS levels (-1 unreached) from each source; d_level nsrc*V int32. */
gb_status gb_bfs(gb_graph g, const int32_t *h_sources, int32_t nsrc,
                 int32_t *d_level, void *stream);

typedef struct gb_bc_stats {
  int32_t max_depth;
  int32_t batches;
  double ms;
} gb_bc_stats;

/* Unweighted Brandes betweenness (sigma, delta FP64; sums not halved, as in
 * csr_bc.c:57-61).  d_bc: V doubles (device). batch = sources per launch
 * group (0 = all at once). */
gb_status gb_bc(gb_graph g, const int32_t *h_sources, int32_t nsrc,
                int32_t batch, double *d_bc, gb_bc_stats *st, void *stream);
/* Weighted Brandes over the f64 weights (gb_graph_set_f64_weights). */
gb_status gb_bc_weighted(gb_graph g, const int32_t *h_sources, int32_t nsrc,
                         int32_t batch, double *d_bc, gb_bc_stats *st,
                         void *stream);

/* ---- learned delta and PageRank warm start --------------------------------------
 * A 2-hidden-layer ReLU MLP (width 64, FP32) trained and evaluated by the
 * library's own kernels.  Parameters are one float array (device):
 * W1[64][din] b1[64] W2[64][64] b2[64] W3[dout][64] b3[dout]; din <= 16,
 * dout <= 4. */
int32_t gb_mlp_param_count(int32_t din, int32_t dout);
/* Y[n][dout] = MLP(X[n][din]) */
gb_status gb_mlp_forward(const float *d_params, int32_t din, int32_t dout,
                         const float *d_X, int64_t n, float *d_Y, void *stream);
/* The same inference with the 64 x 64 hidden layer on FP32 MFMA
 * (v_mfma_f32_32x32x2f32: exact FP32 products, FP32 sums; layers 1 and 3 on
 * the vector units): for many rows, e.g. one per vertex.  Results agree with
 * gb_mlp_forward to FP32 summation order.  The PageRank warm start uses it.
 * No counterpart in the reference (the learned models are this library's). */
gb_status gb_mlp_forward_mfma(const float *d_params, int32_t din, int32_t dout,
                              const float *d_X, int64_t n, float *d_Y, void *stream);
/* The same inference with the 64 x 64 hidden layer on bf16 MFMA (FP32 sums,
 * layers 1 and 3 in FP32): for many rows, e.g. one per vertex.  Results agree
 * with gb_mlp_forward to bf16 rounding of the hidden activations and W2.
 * No counterpart in the reference (the learned models are this library's). */
gb_status gb_mlp_forward_bf16(const float *d_params, int32_t din, int32_t dout,
                              const float *d_X, int64_t n, float *d_Y, void *stream);
/* Mean squared-error loss (0.5*||y-t||^2 averaged over rows) and its gradient
 * (same layout as the parameters), summed in a fixed order. */
gb_status gb_mlp_loss_grad(const float *d_params, int32_t din, int32_t dout,
                           const float *d_X, const float *d_T, int64_t n,
                           float *d_grad, double *loss, void *stream);
/* One Adam step (step counts from 1). */
gb_status gb_mlp_adam(float *d_params, const float *d_grad, float *d_m, float *d_v,
                      int32_t total, float lr, float beta1, float beta2, float eps,
                      int32_t step, void *stream);
/* Inputs of the delta model (6 floats, host): log2 V, log2 avg degree,
 * log2 max weight, log2 mean weight, log2(max out-degree / avg degree),
 * log2 nsrc. */
gb_status gb_graph_features(gb_graph g, int32_t nsrc, float *h_feat, void *stream);
/* The same for n graphs at once (h_feat: n rows of 6), one launch and one
 * copy back; also caches each graph's largest weight.  1 <= n <= 65535 (n = 0
 * is a no-op). */
gb_status gb_graph_features_multi(int32_t n, const gb_graph *graphs, int32_t nsrc, float *h_feat,
                                  void *stream);
/* PageRank warm start from the warm-start model (din 3, dout 1): h_norm holds
 * the feature means[3], stds[3] and the target mean, std.  d_x0: V floats,
 * normalised to sum 1 (pass it as gb_pr_params.d_x0). */
gb_status gb_pr_warm_start(gb_graph g, const float *d_params, const float *h_norm,
                           float *d_x0, void *stream);

/* ---- synthetic graphs (host) -------------------------------------------------
 * Deterministic, counter-based generators used by tests and benchmarks.
 * Rows are sorted by (destination, weight).  Output arrays are host memory
 * allocated by the library (free with gb_host_csr_free). */
typedef struct gb_host_csr {
  int32_t V;
  int64_t E;
  int32_t *row_ptr; /* V+1 */
  int32_t *col;     /* E */
  uint32_t *w;      /* E (NULL if unweighted) */
} gb_host_csr;

/* R-MAT (a,b,c; d = 1-a-b-c), 2^scale vertices, edge_factor*2^scale draws,
 * self-loops dropped, duplicates kept.  weight_mode: 0 none, 1 uniform
 * [wlo,whi], 2 50/50 mix of uniform [wlo,whi] and log-uniform [1,65535].
 * symmetric != 0 adds the reverse of every edge. nverts (0 = 2^scale) allows
 * non-power-of-two V (endpoints >= nverts are rejected and redrawn). */
gb_status gb_gen_rmat(gb_host_csr *out, int32_t scale, int64_t nverts,
                      double edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t weight_mode, uint32_t wlo,
                      uint32_t whi, int32_t symmetric);
/* Erdos-Renyi G(n, m = n*avg_degree), u != v. */
gb_status gb_gen_er(gb_host_csr *out, int64_t n, double avg_degree,
                    uint64_t seed, int32_t weight_mode, uint32_t wlo,
                    uint32_t whi);
/* Reverse (transpose) CSR, rows sorted by (source, weight). */
gb_status gb_transpose(const gb_host_csr *in, gb_host_csr *out);
/* k distinct sources uniform among vertices with out-degree > 0. */
gb_status gb_pick_sources(const gb_host_csr *g, int32_t k, uint64_t seed,
                          int32_t *out);
void gb_host_csr_free(gb_host_csr *g);
/* SplitMix64 finaliser used by the generators (exposed for tests). */
uint64_t gb_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif
#endif
