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

struct gb_graph_s {
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
  bool max_w_valid = false;
  DevBuf inv_outdeg;  // float[V]
  bool inv_outdeg_valid = false;
  int32_t pr_tile_items = 0;
  int64_t pr_ntiles = 0;
  DevBuf pr_tile_row, pr_tile_edge;
  uint32_t split_light_max = 0;
  bool split_valid = false;
  DevBuf split_col, split_w, split_le;
  bool split64_valid = false;
  double split64_delta = 0;
  DevBuf split64_col, split64_w, split64_le;
};

hipStream_t gb_stream(void *s);
int gb_num_cus();
// Max weight of the graph's uint32 weights (cached).
gb_status gb_graph_max_w(gb_graph g, uint32_t *out, hipStream_t st);
// Light-first ordering of the forward CSR for a threshold (cached per handle).
gb_status gb_graph_split(gb_graph g, uint32_t light_max, hipStream_t st);
gb_status gb_graph_split64(gb_graph g, double light_max, hipStream_t st);
// per-module workspace release (gb_release_workspaces)
void gb_pr_release();
void gb_sssp_release();
void gb_bc_release();

#endif
