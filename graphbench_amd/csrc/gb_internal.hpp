// gb_internal.hpp -- shared host-side internals of libgraphbench.so.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

#include "graphbench.h"

// Records the error text for gb_last_error() and returns s.
gb_status gb_fail(gb_status s, const char *fmt, ...);

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>

#include <vector>

#define GB_HIP(call)                                                        \
  do {                                                                      \
    hipError_t _e = (call);                                                 \
    if (_e != hipSuccess)                                                   \
      return gb_fail(GB_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                     hipGetErrorString(_e));                                \
  } while (0)

#define GB_TRY(call)                 \
  do {                               \
    gb_status _s = (call);           \
    if (_s != GB_OK) return _s;      \
  } while (0)

// Device allocation through the library (honours gb_set_mem_cap).
gb_status gb_dmalloc(void **p, size_t bytes);
void gb_dfree(void *p);

// A growable device buffer owned by a graph handle or a workspace.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  gb_status ensure(size_t n) {
    if (n <= bytes) return GB_OK;
    gb_dfree(p);
    p = nullptr;
    bytes = 0;
    size_t want = n < 256 ? 256 : n;
    gb_status s = gb_dmalloc(&p, want);
    if (s == GB_OK) bytes = want;
    return s;
  }
  void release() {
    gb_dfree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
  ~DevBuf() { release(); }
};

// A fresh id per graph handle, never reused (keys of per-handle caches that
// outlive the handle, such as an instantiated PageRank hipGraph).
uint64_t gb_next_graph_uid();

struct gb_graph_s {
  uint64_t uid = gb_next_graph_uid();
  int32_t V = 0;
  int64_t E = 0;
  const int32_t *rp = nullptr;
  const int32_t *col = nullptr;
  const uint32_t *w = nullptr;
  const int32_t *rrp = nullptr;
  const int32_t *rcol = nullptr;
  const uint32_t *rw = nullptr;
  const double *w64 = nullptr;
  const double *rw64 = nullptr;
  bool owns = false;
  DevBuf own_rp, own_col, own_w, own_rrp, own_rcol, own_rw;
  // cached derived data
  uint32_t max_w = 0;  // 0 = not computed
  int32_t max_outdeg = -1;  // -1 = not computed
  bool max_w_valid = false;
  DevBuf inv_outdeg;  // float[V]
  bool inv_outdeg_valid = false;
  // PageRank gather shape {items per thread, hot-set size, row-block edges,
  // binned shape} picked by the autotuner, per (fixed items per thread or
  // none, relabelled or not, row blocks among the candidates or not, binned
  // among them or not); merge-path form when the last two are 0; all 0 = not
  // tuned yet
  int32_t pr_tuned[40][4] = {};
  // batched SSSP shapes picked by the autotuner, per (group width, delta)
  struct MsTuned {
    int32_t W;
    uint64_t delta;
    int8_t scan, per_cu, cshift;
  };
  std::vector<MsTuned> ms_tuned;
  // batched BFS/Brandes shapes picked by the autotuner, per number of
  // 64-source groups in a launch
  struct BcTuned {
    int32_t groups, per_cu, alpha;
  };
  std::vector<BcTuned> bc_tuned;
  int32_t pr_tile_items = 0;
  bool pr_tile_relabeled = false;
  int64_t pr_ntiles = 0;
  DevBuf pr_tile_row, pr_tile_edge;
  // PageRank row blocks (row-block form), per edges-per-block and layout
  int32_t pr_rb_nb = 0;
  bool pr_rb_relabeled = false;
  int64_t pr_rb_nblocks = 0;
  int32_t pr_rb_nlong = 0;
  DevBuf pr_rb, pr_rb_long;
  // PageRank binned form (gb_pagerank.hip): per (layout, chunk and bin
  // widths) the edge streams and work lists, built once per handle
  struct PrBinned {
    bool valid = false, relabeled = false;
    int32_t hub = 0, csl = 0, bsl = 0;  // hub chunk width, log2 chunk / bin widths
    int32_t piece = 0;                  // entries per accumulation piece at most
    int64_t padded = 0;                   // stream entries (segments padded to 4)
    int64_t ntasks = 0, nunits = 0, npieces = 0, nmulti = 0;
    DevBuf src16, dst16, vals, tasks, units, pieces, multi, sum64;
  } pr_pb;
  // PageRank relabelling: vertices renumbered by out-degree (descending), the
  // in-CSR and 1/outdeg in the new numbering (gb_graph_pr_relabel)
  bool rl_valid = false;
  DevBuf rl_perm, rl_rank, rl_rrp, rl_rcol, rl_inv;
  uint32_t split_light_max = 0;
  bool split_valid = false;
  DevBuf split_col, split_w, split_le;
  bool split64_valid = false;
  double split64_delta = 0;
  DevBuf split64_col, split64_w, split64_le;
};

hipStream_t gb_stream(void *s);
// Builds (once per handle) the out-degree relabelling used by the PageRank
// gather; needs the reverse CSR.  Sets g->rl_*.
gb_status gb_graph_pr_relabel(gb_graph g, hipStream_t st);
// Out-degrees (int32[V]) from the forward CSR, or counted from the reverse.
gb_status gb_graph_outdeg(gb_graph g, int32_t *d_deg, hipStream_t st);
int gb_num_cus();
// Largest out-degree (cached; needs the forward CSR or the reverse).
gb_status gb_graph_max_outdeg(gb_graph g, int32_t *out, hipStream_t st);
// Max weight of the graph's uint32 weights (cached).
gb_status gb_graph_max_w(gb_graph g, uint32_t *out, hipStream_t st);
// Light-first ordering of the forward CSR for a threshold (cached per handle).
// sync false: the split stays queued on st (later splits on st only; see
// split_rows)
gb_status gb_graph_split(gb_graph g, uint32_t light_max, hipStream_t st, bool sync = true);
gb_status gb_graph_split64(gb_graph g, double light_max, hipStream_t st);
// Reference-form PageRank with a hook called at the top of every iteration
// with the device wb array (before that iteration's gather).
typedef void (*gb_pr_ref_hook)(int32_t iter, const double *d_wb, void *ctx);
gb_status gb_pagerank_ref_impl(gb_graph g, const double *d_recip, int32_t max_iter, double *d_x,
                               int32_t *iterations, hipStream_t st, gb_pr_ref_hook hook,
                               void *ctx);
// The same loop on caller-provided scratch: d_diff and d_wb (V doubles each)
// and d_flag (one int); allocates nothing.
gb_status gb_pagerank_ref_core(int32_t V, const int32_t *d_irp, const int32_t *d_icol,
                               const double *d_recip, int32_t max_iter, double *d_x,
                               int32_t *iterations, hipStream_t st, double *d_diff, double *d_wb,
                               int *d_flag, gb_pr_ref_hook hook, void *ctx);
// Looks up, with the device current, the per-device address of the
// no-allocation scratch array (gb_noalloc.hip); called by gb_init.
gb_status gb_noalloc_setup(int32_t device);
// comp_distrib_csr on the device: distrib (p, int64) and, when non-null, the
// per-part degree sums of two degree prefix arrays (send and recv).
gb_status gb_edge_balanced_parts_ex(int32_t V, const int32_t *d_cut_rp, const int32_t *d_send_rp,
                                    const int32_t *d_recv_rp, int32_t p, int64_t *d_distrib,
                                    int64_t *d_send_size, int64_t *d_recv_size, hipStream_t st);
// init_node_id on the device: part of every edge target.
gb_status gb_node_ids(int64_t E, const int32_t *d_col, const int64_t *d_distrib, int32_t p,
                      int64_t *d_out, hipStream_t st);
// per-module workspace release (gb_release_workspaces)
void gb_pr_release();
void gb_sssp_release();
void gb_msssp_release();
// Batched delta-stepping for nsrc >= 2 sources of one graph (u32 weights),
// distances to d_dist rows as gb_sssp; accumulates phases, buckets,
// relaxations and the overflow flag.
gb_status gb_msssp_run(gb_graph g, uint32_t delta, const int32_t *h_sources, int32_t nsrc,
                       uint32_t *d_dist, int max_phases, int64_t *phases, int64_t *buckets,
                       unsigned long long *relax, int64_t *imp_near, int64_t *imp_far,
                       int64_t *expanded, bool *ovf, hipStream_t st);
gb_status gb_msssp_count(gb_graph g, int32_t nsrc, const uint32_t *d_dist, int64_t *reached,
                         int64_t *traversed, hipStream_t st);
// Adds the per-wave statistics slots (nwaves * GB_WSTAT, gb_device.hpp) to
// the counters: cnt[idx[k]] += sum over waves of slot k (k < 4; idx3 < 0
// leaves slot 3 out).
gb_status gb_wstat_fold(int64_t nwaves, const unsigned long long *slots, unsigned long long *cnt,
                        int idx0, int idx1, int idx2, hipStream_t st, int idx3 = -1);
void gb_bc_release();
void gb_msbc_release();
void gb_mlp_release();
void gb_learned_release();

#endif
