// gb_pagerank.hip -- PageRank pull gather for gfx950.
//
// Replaces the local gather and update/termination of run_pagerank_csr_mpi
// (pagerank/pagerank_mpi.c:833-846 and 995-1020) on one GPU.
//
// Configured form (gb_pagerank): x_k[v] = base_k + d * sum_{u->v} x_{k-1}[u]/outdeg[u],
// base_k = (1-d)/N + d * dangling_{k-1}/N.  One iteration = two launches:
//   pr_gather  merge-path tiles over the in-CSR (rows + edges split evenly),
//              edge values gathered into LDS with coalesced col reads, each
//              thread consumes IPT items with a compensated FP32 sum, a block
//              segmented scan joins rows split between threads, the epilogue
//              writes x_k, contrib_k = x_k/outdeg and FP64 residual/dangling
//              tile partials.  A row that began in an earlier tile is left to
//   pr_fixup   which adds the earlier tiles' tail partials in tile order and
//              sums the tile partials per block; the last block to finish
//              reduces the block sums in a fixed order, sets the next base
//              and the stop flag (L1 residual < tol or max_iter).
// Every reduction has a fixed order, so results are run-to-run identical.
// Kernels read the stop flag first, so chunks of iterations are captured in
// one hipGraph and replayed until the flag is set.
// Large graphs run in out-degree order (gb_relabel.hip): the hot sources'
// contributions then share cache lines.  The gather shape (items per thread,
// an optional LDS copy of the hottest contributions) is autotuned per graph.
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gb_device.hpp"
#include "gb_internal.hpp"

namespace {

constexpr int PR_BLOCK = 256;

struct PrState {
  double residual;
  double dangling;
  double tol;
  float base;
  float damping;
  int32_t iter;
  int32_t done;
  int32_t max_iter;
  int32_t V;
  int32_t fx_shift;  // binned form: sums in fixed point, value * 2^fx_shift
};

__global__ void inv_outdeg_kernel(int32_t V, const int32_t *__restrict__ rp,
                                  float *__restrict__ inv) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  int32_t d = rp[v + 1] - rp[v];
  inv[v] = d > 0 ? 1.0f / (float)d : 0.0f;
}

// Out-degrees from the in-CSR when the forward CSR is absent.
__global__ void count_outdeg_kernel(int64_t E, const int32_t *__restrict__ icol,
                                    int32_t *__restrict__ deg) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) atomicAdd(&deg[icol[e]], 1);
}
__global__ void inv_from_deg_kernel(int32_t V, const int32_t *__restrict__ deg,
                                    float *__restrict__ inv) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) inv[v] = deg[v] > 0 ? 1.0f / (float)deg[v] : 0.0f;
}

__device__ __forceinline__ double block_sum_f64(double v, double *s_red) {
  v = gb_wave_sum_f64(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if (gb_lane() == 0) s_red[w] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int k = 0; k < nw; k++) t += s_red[k];
  return t;
}

// perm (optional): gather numbering -> caller's numbering, for x0
__global__ __launch_bounds__(PR_BLOCK) void pr_init_kernel(
    int32_t V, const float *__restrict__ x0, const int32_t *__restrict__ perm,
    const float *__restrict__ inv_od, float *__restrict__ x, float *__restrict__ contrib,
    double *__restrict__ part_dang, double *__restrict__ part_mass) {
  __shared__ double s_red[PR_BLOCK / 64];
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dang = 0.0, mass = 0.0;
  if (v < V) {
    float xv = x0 ? x0[perm ? perm[v] : v] : (float)(1.0 / (double)V);
    float io = inv_od[v];
    x[v] = xv;
    contrib[v] = xv * io;
    if (io == 0.0f) dang = (double)xv;
    mass = fabs((double)xv);
  }
  double t = block_sum_f64(dang, s_red);
  double m = block_sum_f64(mass, s_red);
  if (threadIdx.x == 0) {
    part_dang[blockIdx.x] = t;
    part_mass[blockIdx.x] = m;
  }
}

// Sets the first base and the binned form's fixed-point shift.  The L1 mass
// M = sum |x| bounds every contribution and every row sum, and after an
// iteration M' <= (1 - d) + d * M <= max(1, M), so with 2^e >= M (e >= 0)
// every contribution, row sum and partial sum stays at most 2^50 in fixed
// point (value * 2^fx_shift) for the whole solve: fx_shift = 50 - e.  Each
// contribution is rounded to 2^-51 of the mass bound, whatever the start's
// scale.
__global__ void pr_init_finalize_kernel(int32_t nparts, const double *__restrict__ part_dang,
                                        const double *__restrict__ part_mass,
                                        PrState *__restrict__ st) {
  __shared__ double s_red[1024 / 64];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    a += part_dang[i];
    b += part_mass[i];
  }
  double dang = block_sum_f64(a, s_red);
  double mass = block_sum_f64(b, s_red);
  if (threadIdx.x == 0) {
    const double N = (double)st->V;
    const double d = (double)st->damping;
    int e = 0;
    const double mb = mass * (1.0 + 1.0 / 1024.0);  // FP32 rounding of later iterations
    while (e < 60 && ldexp(1.0, e) < mb) ++e;
    st->fx_shift = 50 - e;
    st->dangling = dang;
    st->base = (float)((1.0 - d) / N + d * dang / N);
    st->iter = 0;
    st->residual = INFINITY;
    st->done = st->max_iter <= 0 ? 1 : 0;
  }
}

// ranks back to the caller's numbering: x[old] = x_gather[rank[old]]
__global__ void pr_unpermute_kernel(int32_t V, const int32_t *__restrict__ rank,
                                    const float *__restrict__ xg, float *__restrict__ x) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) x[v] = xg[rank[v]];
}

struct RowOut {
  float base, damping;
  const float *__restrict__ x_old;
  const float *__restrict__ inv_od;
  float *__restrict__ x_new;
  float *__restrict__ contrib_new;
  // row v's new rank from its gathered sum s, given x_old[v] and inv_od[v]
  __device__ __forceinline__ void put(int32_t v, float s, float xo, float io, double &res,
                                      double &dang) const {
    const float xn = base + damping * s;
    res += fabs((double)xn - (double)xo);
    x_new[v] = xn;
    contrib_new[v] = xn * io;
    if (io == 0.0f) dang += (double)xn;
  }
  __device__ __forceinline__ void operator()(int32_t v, F2 tot, double &res, double &dang) const {
    put(v, f2_value(tot), x_old[v], inv_od[v], res, dang);
  }
};

// s_hot[k] = contrib[k] for k < nh (nh >= 1) by the BT threads of a block:
// every read issued before the first store, so the copy costs one round trip.
template <int HOT, int BT>
__device__ __forceinline__ void fill_hot(float *s_hot, const float *__restrict__ contrib, int nh) {
  constexpr int PER = (HOT + BT - 1) / BT;
  float v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = (int)threadIdx.x + j * BT;
    v[j] = contrib[k < nh ? k : nh - 1];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = (int)threadIdx.x + j * BT;
    if (k < nh) s_hot[k] = v[j];
  }
}

// HOT > 0: the contributions of vertices [0, HOT) -- the highest out-degrees
// once the graph is relabelled -- are copied into LDS once per block, and the
// block walks tiles blockIdx.x, blockIdx.x + gridDim.x, ...; gathers of hot
// sources are served from LDS instead of the L2.
template <int IPT, int HOT>
__global__ __launch_bounds__(PR_BLOCK) void pr_gather_kernel(
    const int32_t *__restrict__ irp, const int32_t *__restrict__ icol,
    const int32_t *__restrict__ tile_row, const int32_t *__restrict__ tile_edge,
    const float *__restrict__ contrib, RowOut out, F2 *__restrict__ tail,
    F2 *__restrict__ head, double *__restrict__ part_res, double *__restrict__ part_dang,
    const PrState *__restrict__ st, int64_t ntiles) {
  constexpr int TILE = PR_BLOCK * IPT;
  __shared__ int32_t s_end[TILE];
  __shared__ float s_val[TILE];
  __shared__ F2 s_scan[PR_BLOCK];
  __shared__ int32_t s_flag[PR_BLOCK];
  __shared__ double s_red[PR_BLOCK / 64];
  __shared__ float s_hot[HOT > 0 ? HOT : 1];
  if (st->done) return;
  out.base = st->base;
  const int tid = threadIdx.x;
  if constexpr (HOT > 0) {
    fill_hot<HOT, PR_BLOCK>(s_hot, contrib, (int)min((int64_t)HOT, (int64_t)st->V));
    __syncthreads();
  }
  for (int64_t tt = blockIdx.x; tt < ntiles; tt += gridDim.x) {
  const int64_t t = tt;
  const int32_t r0 = tile_row[t], e0 = tile_edge[t];
  const int32_t r1 = tile_row[t + 1], e1 = tile_edge[t + 1];
  const int nrows = r1 - r0, nedges = e1 - e0;
  const bool cont = irp[r0] < e0;  // row r0 began in an earlier tile
  // Stage the tile: IPT independent column loads per thread (streamed past the
  // caches), then IPT independent gathers, so many requests are in flight.
  // The loads are unconditional, on indices clamped into the tile (items past
  // its end re-read its last edge or row): a load under a per-item condition
  // is compiled to a branch that waits for it, one round trip per item.
  // nedges and nrows are uniform over the block.
  if (nedges > 0) {
    const int32_t *icol_t = icol + e0;
    int32_t cidx[IPT];
    float val[IPT];
#pragma unroll
    for (int m = 0; m < IPT; ++m) {
      const int k = tid + m * PR_BLOCK;
      cidx[m] = __builtin_nontemporal_load(icol_t + (k < nedges ? k : nedges - 1));
    }
#pragma unroll
    for (int m = 0; m < IPT; ++m) {
      const uint32_t c = (uint32_t)cidx[m];
      if (HOT > 0) {
        const float g = contrib[c < (uint32_t)HOT ? 0u : c];  // hot lanes: a dummy read
        val[m] = c < (uint32_t)HOT ? s_hot[c] : g;
      } else {
        val[m] = contrib[c];
      }
    }
#pragma unroll
    for (int m = 0; m < IPT; ++m) {
      const int k = tid + m * PR_BLOCK;
      if (k < nedges) s_val[k] = val[m];
    }
  }
  if (nrows > 0) {
    const int32_t *irp_t = irp + r0 + 1;
    int32_t ends[IPT];
#pragma unroll
    for (int m = 0; m < IPT; ++m) {
      const int k = tid + m * PR_BLOCK;
      ends[m] = irp_t[k < nrows ? k : nrows - 1];
    }
#pragma unroll
    for (int m = 0; m < IPT; ++m) {
      const int k = tid + m * PR_BLOCK;
      if (k < nrows) s_end[k] = ends[m] - e0;
    }
  }
  __syncthreads();

  // this thread's start on the merge path, searched in LDS
  const int total = nrows + nedges;
  int dg = tid * IPT;
  if (dg > total) dg = total;
  int lo = dg - nedges > 0 ? dg - nedges : 0;
  int hi = dg < nrows ? dg : nrows;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (s_end[mid] <= dg - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  int i = lo, j = dg - lo;
  F2 acc = {0.0f, 0.0f}, first = {0.0f, 0.0f};
  int first_row = -1;
  double res = 0.0, dang = 0.0;
  // rows closed after the first one, at the steps set in done_mask: they are
  // the rows first_row + 1, first_row + 2, ... in step order.  Their sums go
  // to LDS after the scan and are written out by a row-parallel pass.
  uint32_t done_mask = 0;
  float done_sum[IPT];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    done_sum[k] = 0.0f;
    if (i < nrows && s_end[i] <= j) {
      if (first_row < 0) {
        first_row = i;
        first = acc;
      } else {
        done_mask |= 1u << k;
        done_sum[k] = f2_value(acc);
      }
      acc.s = 0.0f;
      acc.c = 0.0f;
      ++i;
    } else if (j < nedges) {
      f2_add(acc, s_val[j]);
      ++j;
    }
  }
  // segmented inclusive scan of the open-row partials over the threads
  s_scan[tid] = acc;
  s_flag[tid] = first_row >= 0;
  __syncthreads();
  for (int off = 1; off < PR_BLOCK; off <<= 1) {
    F2 vb = s_scan[tid];
    int fb = s_flag[tid];
    F2 va;
    int fa = 0;
    const bool has = tid >= off;
    if (has) {
      va = s_scan[tid - off];
      fa = s_flag[tid - off];
    }
    __syncthreads();
    if (has) {
      s_scan[tid] = fb ? vb : f2_merge(va, vb);
      s_flag[tid] = fa | fb;
    }
    __syncthreads();
  }
  // every row of the tile closes here; its sum goes to s_val[row] (the edge
  // values are no longer read: the scan's barriers came after the merge)
  if (first_row >= 0) {
    F2 cin = tid > 0 ? s_scan[tid - 1] : F2{0.0f, 0.0f};
    F2 tot = f2_merge(cin, first);
    if (first_row == 0 && cont)
      head[t] = tot;  // row r0 began in an earlier tile: pr_fixup finishes it
    else
      s_val[first_row] = f2_value(tot);
    int ir = first_row + 1;
#pragma unroll
    for (int k = 0; k < IPT; ++k)
      if ((done_mask >> k) & 1u) s_val[ir++] = done_sum[k];
  }
  if (tid == PR_BLOCK - 1) tail[t] = s_scan[PR_BLOCK - 1];
  __syncthreads();
  // row-parallel output: coalesced reads of x_old / inv_od, coalesced writes
  for (int k = (cont ? 1 : 0) + tid; k < nrows; k += PR_BLOCK)
    out.put(r0 + k, s_val[k], out.x_old[r0 + k], out.inv_od[r0 + k], res, dang);
  double rs = block_sum_f64(res, s_red);
  double ds = block_sum_f64(dang, s_red);
  if (tid == 0) {
    part_res[t] = rs;
    part_dang[t] = ds;
  }
  __syncthreads();  // the next tile reuses the LDS staging
  }
}

// Finishes rows that began in an earlier tile and sums, per block, the gather
// partials of its tiles with its own; the last block to finish reduces the
// block sums (fixed order) and sets the next base and the stop flag.
__global__ __launch_bounds__(PR_BLOCK) void pr_fixup_kernel(
    int64_t ntiles, const int32_t *__restrict__ irp, const int32_t *__restrict__ tile_row,
    const int32_t *__restrict__ tile_edge, const F2 *__restrict__ tail,
    const F2 *__restrict__ head, RowOut out, const double *__restrict__ p_res,
    const double *__restrict__ p_dang, double *__restrict__ fix_res,
    double *__restrict__ fix_dang, unsigned int *__restrict__ done_blocks,
    PrState *__restrict__ st) {
  __shared__ double s_red[PR_BLOCK / 64];
  __shared__ int s_last;
  if (st->done) return;
  out.base = st->base;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0, dang = 0.0;
  if (t < ntiles) {
    // this tile's gather partials join the fix-up block's sum, so the last
    // block reduces one value per fix-up block instead of one per tile
    res = p_res[t];
    dang = p_dang[t];
    const int32_t r0 = tile_row[t], e0 = tile_edge[t], r1 = tile_row[t + 1];
    if (r1 > r0) {
      const int32_t rs = irp[r0];
      if (rs < e0) {
        F2 acc = head[t];
        for (int64_t k = t - 1; k >= 0; --k) {
          acc = f2_merge(tail[k], acc);
          if (tile_edge[k] <= rs) break;
        }
        out(r0, acc, res, dang);
      }
    }
  }
  double rsum = block_sum_f64(res, s_red);
  double dsum = block_sum_f64(dang, s_red);
  if (threadIdx.x == 0) {
    fix_res[blockIdx.x] = rsum;
    fix_dang[blockIdx.x] = dsum;
    __threadfence();
    s_last = atomicAdd(done_blocks, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  const int64_t n2 = gridDim.x;
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < n2; i += blockDim.x) {
    a += __builtin_nontemporal_load(fix_res + i);
    b += __builtin_nontemporal_load(fix_dang + i);
  }
  const double rtot = block_sum_f64(a, s_red);
  const double dtot = block_sum_f64(b, s_red);
  if (threadIdx.x == 0) {
    const double N = (double)st->V;
    const double d = (double)st->damping;
    st->iter += 1;
    st->residual = rtot;
    st->dangling = dtot;
    st->base = (float)((1.0 - d) / N + d * dtot / N);
    st->done = (rtot < st->tol || st->iter >= st->max_iter) ? 1 : 0;
    *done_blocks = 0;  // ready for the next iteration
  }
}

// ---- reference-semantics form (FP64, bit-identical to the C loop) ---------

__global__ void pr_ref_init_kernel(int32_t V, const double *__restrict__ recip,
                                   double *__restrict__ x, double *__restrict__ diff,
                                   double *__restrict__ wb) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  x[v] = 1.0;                          // init_weight, pagerank_mpi.c:63-68
  diff[v] = __dmul_rn(1.0, recip[v]);  // 752-754
  wb[v] = 0.3;                         // RDM_JMP, 808-809
}

__global__ void pr_ref_gather_kernel(int32_t V, const int32_t *__restrict__ irp,
                                     const int32_t *__restrict__ icol,
                                     const double *__restrict__ diff, double *__restrict__ wb,
                                     double pj) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  double acc = wb[v];
  for (int32_t e = irp[v]; e < irp[v + 1]; e++)  // 833-846, in-CSR order
    acc = __dadd_rn(acc, __dmul_rn(pj, diff[icol[e]]));
  wb[v] = acc;
}

__global__ void pr_ref_flag_clear_kernel(int *__restrict__ not_done) { *not_done = 0; }

__global__ void pr_ref_update_kernel(int32_t V, const double *__restrict__ recip,
                                     double *__restrict__ x, double *__restrict__ diff,
                                     const double *__restrict__ wb, int *__restrict__ not_done) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool big = false;
  if (v < V) {
    double dv = __dmul_rn(__dsub_rn(wb[v], x[v]), recip[v]);  // 995-997
    diff[v] = dv;
    big = fabs(dv) > 0.000001;  // THRESH; NaN compares false (converged)
    x[v] = wb[v];
  }
  if (__ballot(big) != 0ull && gb_lane() == 0) atomicOr(not_done, 1);
}

// ---- row-block form (no merge path) -----------------------------------------
// The rows are cut, once per graph layout, into blocks of consecutive rows
// holding at most NB edges and RB_MAXR rows; a row with more than NB edges is
// cut into chunks of NB edges, one block each.  A row block stages its edges'
// contributions in LDS (coalesced col reads, one batch of gathers) and sums
// every row there: one thread per row, or a power-of-two group of lanes per
// row when the block has few rows (lanes stride the row, then a fixed
// butterfly merges their partial sums).  No merge-path search, no segmented
// scan, no cross-tile fix-up for rows that fit a block.  Chunks of long rows
// leave one partial each; pr_rb_fixup_kernel joins them in chunk order.
// Residual and dangling partials are per wave.
constexpr int RB_MAXR = 1024;  // rows per block at most
constexpr int RB_RPT = RB_MAXR / PR_BLOCK;

struct RowBlock {
  int32_t r0, nrows, e0, e1;  // nrows == 0: one chunk [e0, e1) of long row r0
};
struct LongRow {
  int32_t row, b0, nb;  // chunks are blocks b0 .. b0 + nb - 1
};

// Row block bi, summed by PR_BLOCK threads (tid = thread among them) with the
// NB floats at s_val.  All reads (column indices, gathers, the rows' pointers,
// old ranks and inverse out-degrees) are issued before the one barrier, so
// they are in flight together; every thread of the launch block reaches that
// barrier exactly once per call, also when the row block is inactive (active
// false: past the last row block).  HOT > 0: contributions of sources
// [0, HOT) come from s_hot.  w_res / w_dang: the wave's sums of |x_new - x_old|
// and of the dangling mass over its rows (every lane; 0 for a chunk of a long
// row, whose row the fix-up finishes, and for an inactive row block).
template <int NB, int HOT>
__device__ __forceinline__ void rowblock_body(
    const int32_t *__restrict__ irp, const int32_t *__restrict__ icol,
    const RowBlock *__restrict__ blocks, const float *__restrict__ contrib, const RowOut &out,
    F2 *__restrict__ chunk_sum, int64_t bi, bool active, int tid, float *s_val,
    const float *s_hot, double &w_res, double &w_dang) {
  constexpr int EPT = NB / PR_BLOCK;  // staged edges per thread
  const int lane = gb_lane(), wave = tid / GB_WAVE;
  RowBlock b = {0, 0, 0, 0};
  if (active) b = blocks[bi];
  const int nnz = b.e1 - b.e0;
  const int nrows = b.nrows;
  const bool chunk = active && nrows == 0;
  // edge values: unconditional reads on clamped indices, one batch each
  // (hot lanes read a dummy address)
  float val[EPT];
  if (nnz > 0) {
    const int32_t *col_b = icol + b.e0;
    int32_t c[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const int k = tid + m * PR_BLOCK;
      c[m] = __builtin_nontemporal_load(col_b + (k < nnz ? k : nnz - 1));
    }
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const uint32_t cu = (uint32_t)c[m];
      if (HOT > 0) {
        const float g = contrib[cu < (uint32_t)HOT ? 0u : cu];
        val[m] = cu < (uint32_t)HOT ? s_hot[cu] : g;
      } else {
        val[m] = contrib[cu];
      }
    }
  }
  // lanes per row: the largest power of two with nrows * tpr <= 256, at most 64
  int tpr = 1;
  while (tpr < GB_WAVE && nrows * tpr * 2 <= PR_BLOCK) tpr <<= 1;
  // this thread's rows: tid + m * PR_BLOCK (tpr == 1) or tid / tpr; their
  // edge ranges [lo, hi) and old values, read on clamped rows
  int32_t lo[RB_RPT], hi[RB_RPT];
  float xo[RB_RPT], io[RB_RPT];
  if (nrows > 0) {
#pragma unroll
    for (int m = 0; m < RB_RPT; ++m) {
      const int r = tpr == 1 ? tid + m * PR_BLOCK : tid / tpr;
      const int rc = r < nrows ? r : nrows - 1;
      lo[m] = irp[b.r0 + rc];
      hi[m] = irp[b.r0 + rc + 1];
      xo[m] = out.x_old[b.r0 + rc];
      io[m] = out.inv_od[b.r0 + rc];
    }
  }
  if (chunk) {  // a chunk of a long row: one compensated sum for the block
    F2 acc = {0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < EPT; ++m) f2_add(acc, tid + m * PR_BLOCK < nnz ? val[m] : 0.0f);
#pragma unroll
    for (int d = 1; d < GB_WAVE; d <<= 1) {
      F2 o;
      o.s = __shfl_xor(acc.s, d, GB_WAVE);
      o.c = __shfl_xor(acc.c, d, GB_WAVE);
      acc = (lane & d) ? f2_merge(o, acc) : f2_merge(acc, o);
    }
    if (lane == 0) reinterpret_cast<F2 *>(s_val)[wave] = acc;  // s_val is free here
  } else if (nrows > 0) {
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const int k = tid + m * PR_BLOCK;
      if (k < nnz) s_val[k] = val[m];
    }
  }
  __syncthreads();
  w_res = 0.0;
  w_dang = 0.0;
  if (chunk) {
    if (tid == 0) {
      const F2 *s_red = reinterpret_cast<const F2 *>(s_val);
      F2 tot = s_red[0];
      for (int w = 1; w < PR_BLOCK / GB_WAVE; w++) tot = f2_merge(tot, s_red[w]);
      chunk_sum[bi] = tot;
    }
    return;
  }
  if (nrows == 0) return;  // inactive
  double res = 0.0, dang = 0.0;
  if (tpr == 1) {
#pragma unroll
    for (int m = 0; m < RB_RPT; ++m) {
      const int r = tid + m * PR_BLOCK;
      F2 acc = {0.0f, 0.0f};
      if (r < nrows)
        for (int k = lo[m] - b.e0; k < hi[m] - b.e0; ++k) f2_add(acc, s_val[k]);
      if (r < nrows) out.put(b.r0 + r, f2_value(acc), xo[m], io[m], res, dang);
    }
  } else {
    const int r = tid / tpr, q = tid % tpr;
    F2 acc = {0.0f, 0.0f};
    if (r < nrows)
      for (int k = lo[0] - b.e0 + q; k < hi[0] - b.e0; k += tpr) f2_add(acc, s_val[k]);
    // fixed butterfly over the row's lanes (aligned groups inside one wave)
    for (int d = 1; d < tpr; d <<= 1) {
      F2 o;
      o.s = __shfl_xor(acc.s, d, GB_WAVE);
      o.c = __shfl_xor(acc.c, d, GB_WAVE);
      acc = (q & d) ? f2_merge(o, acc) : f2_merge(acc, o);
    }
    if (r < nrows && q == 0) out.put(b.r0 + r, f2_value(acc), xo[0], io[0], res, dang);
  }
  w_res = gb_wave_sum_f64(res);
  w_dang = gb_wave_sum_f64(dang);
}

template <int NB>
__global__ __launch_bounds__(PR_BLOCK) void pr_rowblock_kernel(
    const int32_t *__restrict__ irp, const int32_t *__restrict__ icol,
    const RowBlock *__restrict__ blocks, const float *__restrict__ contrib, RowOut out,
    F2 *__restrict__ chunk_sum, double *__restrict__ part_res, double *__restrict__ part_dang,
    const PrState *__restrict__ st) {
  __shared__ float s_val[NB];
  if (st->done) return;
  out.base = st->base;
  const int64_t bi = blockIdx.x;
  double r, d;
  rowblock_body<NB, 0>(irp, icol, blocks, contrib, out, chunk_sum, bi, true, threadIdx.x, s_val,
                       nullptr, r, d);
  if (gb_lane() == 0) {
    part_res[bi * (PR_BLOCK / GB_WAVE) + threadIdx.x / GB_WAVE] = r;
    part_dang[bi * (PR_BLOCK / GB_WAVE) + threadIdx.x / GB_WAVE] = d;
  }
}

// Row blocks per launch block of the hot-set variant: the count that keeps the
// most waves resident per CU (160 KiB of LDS per CU shared by blocks of
// HOT + SUBS * NB floats; 24 waves per CU counted at most), the fewest on ties.
constexpr int rb_subs(int nb, int hot) {
  int best = 1, best_waves = 0;
  for (int s = 1; s <= 4; ++s) {
    const int per_cu = 160 * 1024 / (4 * (hot + s * nb));
    int waves = per_cu * s * (PR_BLOCK / GB_WAVE);
    if (waves > 24) waves = 24;
    if (waves > best_waves) {
      best_waves = waves;
      best = s;
    }
  }
  return best;
}

// Persistent variant with an LDS copy of the hottest HOT contributions (the
// highest out-degrees once relabelled), shared by SUBS row blocks summed side
// by side (PR_BLOCK threads each): launch block j walks row blocks
// SUBS * (j + k * gridDim.x) + sub, k = 0, 1, ...  Each wave sums its
// residual and dangling partials over its walk (in walk order) and writes
// them once: gridDim.x * SUBS * 4 partials.
template <int NB, int HOT>
__global__ __launch_bounds__(PR_BLOCK * rb_subs(NB, HOT)) void pr_rowblock_hot_kernel(
    const int32_t *__restrict__ irp, const int32_t *__restrict__ icol,
    const RowBlock *__restrict__ blocks, const float *__restrict__ contrib, RowOut out,
    F2 *__restrict__ chunk_sum, double *__restrict__ part_res, double *__restrict__ part_dang,
    const PrState *__restrict__ st, int64_t nblocks) {
  constexpr int SUBS = rb_subs(NB, HOT);
  __shared__ float s_hot[HOT];
  __shared__ float s_val[SUBS][NB];
  if (st->done) return;
  out.base = st->base;
  fill_hot<HOT, PR_BLOCK * SUBS>(s_hot, contrib, (int)min((int64_t)HOT, (int64_t)st->V));
  __syncthreads();
  // wave-uniform sub-block index (every lane of a wave has the same one)
  const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x / PR_BLOCK);
  const int tid = (int)threadIdx.x - sub * PR_BLOCK;
  double acc_res = 0.0, acc_dang = 0.0;
  for (int64_t b0 = (int64_t)blockIdx.x * SUBS; b0 < nblocks; b0 += (int64_t)gridDim.x * SUBS) {
    const int64_t bi = b0 + sub;
    double r, d;
    rowblock_body<NB, HOT>(irp, icol, blocks, contrib, out, chunk_sum, bi, bi < nblocks, tid,
                           s_val[sub], s_hot, r, d);
    acc_res += r;
    acc_dang += d;
    __syncthreads();  // the next row blocks reuse the LDS staging
  }
  if (gb_lane() == 0) {
    const int64_t gw = (int64_t)blockIdx.x * (PR_BLOCK * SUBS / GB_WAVE) + threadIdx.x / GB_WAVE;
    part_res[gw] = acc_res;
    part_dang[gw] = acc_dang;
  }
}

// A long row's sum from its chunk partials, merged in chunk order; the loads
// go out four at a time.
__device__ __forceinline__ F2 join_chunks(const F2 *__restrict__ chunk_sum, const LongRow &L) {
  F2 acc = chunk_sum[L.b0];
  int32_t k = 1;
  for (; k + 4 <= L.nb; k += 4) {
    const F2 c0 = chunk_sum[L.b0 + k], c1 = chunk_sum[L.b0 + k + 1];
    const F2 c2 = chunk_sum[L.b0 + k + 2], c3 = chunk_sum[L.b0 + k + 3];
    acc = f2_merge(f2_merge(f2_merge(f2_merge(acc, c0), c1), c2), c3);
  }
  for (; k < L.nb; k++) acc = f2_merge(acc, chunk_sum[L.b0 + k]);
  return acc;
}

// The end of an iteration from its residual and dangling totals: the next
// base and the stop flag.
__device__ __forceinline__ void pr_finish_iteration(PrState *st, double rtot, double dtot) {
  const double N = (double)st->V;
  const double d = (double)st->damping;
  st->iter += 1;
  st->residual = rtot;
  st->dangling = dtot;
  st->base = (float)((1.0 - d) / N + d * dtot / N);
  st->done = (rtot < st->tol || st->iter >= st->max_iter) ? 1 : 0;
}

// The fix-up in one block, when the long rows and the partials are few (no
// cross-block reduction, so no fence): long rows from their chunk sums (chunk
// order), then the partials, each thread's share read in batches of 8 on
// clamped indices, summed in a fixed order.
constexpr int RB_FIX1_THREADS = 1024;
constexpr int64_t RB_FIX1_MAX_PARTS = 16384;
constexpr int32_t RB_FIX1_MAX_LONG = 4 * RB_FIX1_THREADS;
__global__ __launch_bounds__(RB_FIX1_THREADS) void pr_rb_fixup_one_kernel(
    int64_t nparts, int32_t nlong, const LongRow *__restrict__ longs,
    const F2 *__restrict__ chunk_sum, RowOut out, const double *__restrict__ p_res,
    const double *__restrict__ p_dang, PrState *__restrict__ st) {
  __shared__ double s_red[RB_FIX1_THREADS / 64];
  if (st->done) return;
  out.base = st->base;
  double res = 0.0, dang = 0.0;
  for (int32_t t = threadIdx.x; t < nlong; t += RB_FIX1_THREADS) {
    const LongRow L = longs[t];
    out(L.row, join_chunks(chunk_sum, L), res, dang);
  }
  constexpr int U = 8;
  for (int64_t b = threadIdx.x; b < nparts; b += (int64_t)RB_FIX1_THREADS * U) {
    double a[U], c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + (int64_t)u * RB_FIX1_THREADS;
      const int64_t j = i < nparts ? i : nparts - 1;
      a[u] = p_res[j];
      c[u] = p_dang[j];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = b + (int64_t)u * RB_FIX1_THREADS < nparts;
      res += in ? a[u] : 0.0;
      dang += in ? c[u] : 0.0;
    }
  }
  const double rtot = block_sum_f64(res, s_red);
  const double dtot = block_sum_f64(dang, s_red);
  if (threadIdx.x == 0) pr_finish_iteration(st, rtot, dtot);
}

// Long rows from their chunk sums (chunk order), the gather's per-wave
// partials summed per block; the last block reduces the block sums in a fixed
// order and sets the next base and the stop flag (as pr_fixup_kernel).
__global__ __launch_bounds__(PR_BLOCK) void pr_rb_fixup_kernel(
    int64_t nparts, int32_t nlong, const LongRow *__restrict__ longs,
    const F2 *__restrict__ chunk_sum, RowOut out, const double *__restrict__ p_res,
    const double *__restrict__ p_dang, double *__restrict__ fix_res,
    double *__restrict__ fix_dang, unsigned int *__restrict__ done_blocks,
    PrState *__restrict__ st) {
  __shared__ double s_red[PR_BLOCK / 64];
  __shared__ int s_last;
  if (st->done) return;
  out.base = st->base;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0, dang = 0.0;
  if (t < nparts) {
    res = p_res[t];
    dang = p_dang[t];
  }
  if (t < nlong) {
    const LongRow L = longs[t];
    out(L.row, join_chunks(chunk_sum, L), res, dang);
  }
  double rsum = block_sum_f64(res, s_red);
  double dsum = block_sum_f64(dang, s_red);
  if (threadIdx.x == 0) {
    fix_res[blockIdx.x] = rsum;
    fix_dang[blockIdx.x] = dsum;
    __threadfence();
    s_last = atomicAdd(done_blocks, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  const int64_t n2 = gridDim.x;
  double a = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < n2; i += blockDim.x) {
    a += __builtin_nontemporal_load(fix_res + i);
    c += __builtin_nontemporal_load(fix_dang + i);
  }
  const double rtot = block_sum_f64(a, s_red);
  const double dtot = block_sum_f64(c, s_red);
  if (threadIdx.x == 0) {
    pr_finish_iteration(st, rtot, dtot);
    *done_blocks = 0;  // ready for the next iteration
  }
}

// ---- binned form -----------------------------------------------------------
//
// The pull gather reads contrib[u] once per in-edge at a random address: one
// L2 request per edge, and the L2's request rate bounds it.  The binned form
// (propagation blocking) streams instead.  Sources are cut into a hub chunk
// (chunk 0: the first HUB ids, the highest out-degrees once relabelled)
// and chunks of 2^CSL ids after it, destinations into bins of 2^BSL ids; the
// edges of each (chunk, bin) pair form a segment, padded to a multiple of 4
// entries.
//   pb_scatter  one block per unit (a run of one chunk's segments): the
//               chunk's contributions are copied into LDS, then each wave
//               task (<= 256 entries of one segment) reads 4 u16 local source
//               ids per lane and writes the 4 contributions to the value
//               stream, where segments lie in (bin, chunk) order.
//   pb_accum    one block per piece (a run of one bin's stream): 4 values and
//               4 u16 local destination ids per lane, added into 2^BSL int64
//               LDS accumulators.  A bin held by one piece finishes its rows
//               in the block; a bin split over several pieces adds its sums
//               to sum64 with integer atomics and
//   pb_fixup    finishes those rows, sums the per-wave residual and dangling
//               partials and sets the next base and stop flag (as the other
//               forms' fix-ups do).
// Sums are fixed point (value * 2^fx_shift, rounded to nearest, int64), so
// the order of the LDS and global atomics changes nothing: results are
// run-to-run identical, and each row sum is exact up to one rounding per
// edge of 2^-51 of the mass bound (2^-51 for ranks summing to 1) and the
// final rounding to FP32.
constexpr int PB_TASK = 256;  // stream entries per wave task, 4 per lane
constexpr int PB_SCATTER_BLOCK = 512;
constexpr int PB_ACCUM_BLOCK = 1024;
constexpr int PB_ACCUM_WAVES = PB_ACCUM_BLOCK / GB_WAVE;
constexpr uint16_t PB_PAD = 0xFFFF;  // padding entry of the source stream
constexpr int PB_ALIGN = 4;          // segments padded to a multiple of this many entries

struct PbTask {
  int32_t in0, out0, n;  // n entries (multiple of 4) at src16[in0..], vals[out0..]
};
struct PbUnit {
  int32_t src0, t0, t1;  // chunk's first source, tasks [t0, t1)
};
// Stream entries [e0, e1) of bin; multi: the bin is split over pieces.  The
// bin's first segment (the hub chunk's edges) ends at hub_e1 and is read from
// src16 at e + hub_off: its contributions come from an LDS copy, the scatter
// skips the hub chunk.
struct PbPiece {
  int32_t bin, e0, e1, multi, hub_e1, hub_off;
};
typedef uint32_t pb_u32x2 __attribute__((ext_vector_type(2)));  // 4 u16 ids
typedef float pb_f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ pb_u32x2 pb_ids4(const uint16_t *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const pb_u32x2 *>(p));
}
__device__ __forceinline__ pb_f32x4 pb_vals4(const float *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const pb_f32x4 *>(p));
}

// round(v * scale) as int64 (nearest, ties to even), scale a power of two
// with |v| * scale < 2^51: the product is exact in FP64, and adding 1.5 * 2^52
// leaves the rounded integer in the low mantissa bits.
__device__ __forceinline__ long long pb_fixed(float v, double scale) {
  const double y = (double)v * scale + 6755399441055744.0;
  return __double_as_longlong(y) - 0x4338000000000000LL;
}

__device__ __forceinline__ float pb_value(unsigned long long a, int S) {
  return (float)ldexp((double)(long long)a, -S);
}

// contributions of 4 local source ids (padding ids, >= CS, give 0)
template <int CS>
__device__ __forceinline__ pb_f32x4 pb_lookup4(const float *s_c, pb_u32x2 s) {
  const uint32_t i0 = s.x & 0xFFFFu, i1 = s.x >> 16, i2 = s.y & 0xFFFFu, i3 = s.y >> 16;
  pb_f32x4 o;
  o.x = i0 < (uint32_t)CS ? s_c[i0] : 0.0f;
  o.y = i1 < (uint32_t)CS ? s_c[i1] : 0.0f;
  o.z = i2 < (uint32_t)CS ? s_c[i2] : 0.0f;
  o.w = i3 < (uint32_t)CS ? s_c[i3] : 0.0f;
  return o;
}

// The end of a binned iteration: its nparts residual / dangling partials (one
// per piece, then one per split-bin row block) summed in a fixed order by one
// block of PB_SCATTER_BLOCK threads, then the next base and the stop flag.
// Thread t sums partials t, t + NT, t + 2 NT, ... in that order whatever U
// (the loads in flight per batch), so every U gives the same totals.
template <int U>
__device__ __forceinline__ void pb_fix_block(int64_t nparts, const double *__restrict__ p_res,
                                             const double *__restrict__ p_dang, PrState *st,
                                             double *s_red) {
  constexpr int NT = PB_SCATTER_BLOCK;
  double res = 0.0, dang = 0.0;
  for (int64_t b = threadIdx.x; b < nparts; b += (int64_t)NT * U) {
    double a[U], c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + (int64_t)u * NT;
      const int64_t j = i < nparts ? i : nparts - 1;
      a[u] = p_res[j];
      c[u] = p_dang[j];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = b + (int64_t)u * NT < nparts;
      res += in ? a[u] : 0.0;
      dang += in ? c[u] : 0.0;
    }
  }
  const double rtot = block_sum_f64(res, s_red);
  const double dtot = block_sum_f64(dang, s_red);
  if (threadIdx.x == 0) pr_finish_iteration(st, rtot, dtot);
}

// pb_fix_block as a launch of its own (the last iteration of a chunk, or
// when there is nothing to scatter).
__global__ __launch_bounds__(PB_SCATTER_BLOCK) void pb_fix_kernel(
    int64_t nparts, const double *__restrict__ p_res, const double *__restrict__ p_dang,
    PrState *st) {
  __shared__ double s_red[PB_SCATTER_BLOCK / GB_WAVE];
  if (st->done) return;
  pb_fix_block<8>(nparts, p_res, p_dang, st, s_red);
}

// fold = 1: block 0 ends the previous iteration (pb_fix_block on its
// partials) while the other blocks scatter.  They read the stop flag as it
// was before that iteration or as block 0 sets it: either way, this
// iteration's accumulation runs only if the flag stays clear, so a scatter
// done after the last iteration is only wasted work, and at most one is.
template <int CSL>
__global__ __launch_bounds__(PB_SCATTER_BLOCK) void pb_scatter_kernel(
    const PbUnit *__restrict__ units, const PbTask *__restrict__ tasks,
    const uint16_t *__restrict__ src16, const float *__restrict__ contrib,
    float *__restrict__ vals, int fold, int64_t nparts, const double *__restrict__ p_res,
    const double *__restrict__ p_dang, PrState *st) {
  constexpr int CS = 1 << CSL;
  constexpr int PER = CS / PB_SCATTER_BLOCK;
  constexpr int WAVES = PB_SCATTER_BLOCK / GB_WAVE;
  __shared__ float s_c[CS];
  if (fold && blockIdx.x == 0) {
    __shared__ double s_red[WAVES];
    if (st->done) return;
    pb_fix_block<2>(nparts, p_res, p_dang, st, s_red);  // few registers: the scatter's occupancy
    return;
  }
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const PbUnit u = units[blockIdx.x - fold];
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / GB_WAVE);
  const int lane = gb_lane(), l4 = lane * 4;
  // waves take R consecutive tasks in turn (wave w: tasks t0 + w * R + k *
  // WAVES * R), so a unit's tasks spread over all its waves; a round's R
  // descriptors come one per lane, loaded during the round before, and its R
  // id loads are in flight together.  The first round's descriptors and ids
  // load together with the chunk's contributions.
  constexpr int R = 16;
  constexpr int32_t STRIDE = WAVES * R;
  int32_t base = u.t0 + wave * R;
  PbTask my = {0, 0, 0};
  if (lane < R && base + lane < u.t1) my = tasks[base + lane];
  // the chunk's contributions (the last chunk may be short: clamped reads,
  // entries past it are never indexed)
  const int32_t nsrc = min(CS, st->V - u.src0);
  float c[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = (int)threadIdx.x + j * PB_SCATTER_BLOCK;
    c[j] = contrib[u.src0 + (k < nsrc ? k : nsrc - 1)];
  }
  int32_t in0[R], out0[R], n[R];
  pb_u32x2 ids[R];
  auto decode_and_load = [&](int nt) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      n[k] = k < nt ? __builtin_amdgcn_readlane(my.n, k) : 0;
      in0[k] = __builtin_amdgcn_readlane(my.in0, k);
      out0[k] = __builtin_amdgcn_readlane(my.out0, k);
    }
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (l4 < n[k]) ids[k] = pb_ids4(src16 + in0[k] + l4);
  };
  if (base < u.t1) decode_and_load(min(R, u.t1 - base));
#pragma unroll
  for (int j = 0; j < PER; ++j) s_c[threadIdx.x + j * PB_SCATTER_BLOCK] = c[j];
  __syncthreads();
  for (bool first = true; base < u.t1; base += STRIDE, first = false) {
    PbTask nx = {0, 0, 0};
    if (lane < R && base + STRIDE + lane < u.t1) nx = tasks[base + STRIDE + lane];
    if (!first) decode_and_load(min(R, u.t1 - base));
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (l4 < n[k])
        *reinterpret_cast<pb_f32x4 *>(vals + out0[k] + l4) = pb_lookup4<CS>(s_c, ids[k]);
    my = nx;
  }
}

template <int BSL>
__device__ __forceinline__ void pb_add4(unsigned long long *s_acc, pb_f32x4 v, pb_u32x2 d,
                                        double scale) {
  constexpr uint32_t M = (1u << BSL) - 1;
  const long long f0 = pb_fixed(v.x, scale), f1 = pb_fixed(v.y, scale);
  const long long f2 = pb_fixed(v.z, scale), f3 = pb_fixed(v.w, scale);
  if (f0) atomicAdd(&s_acc[d.x & M], (unsigned long long)f0);
  if (f1) atomicAdd(&s_acc[(d.x >> 16) & M], (unsigned long long)f1);
  if (f2) atomicAdd(&s_acc[d.y & M], (unsigned long long)f2);
  if (f3) atomicAdd(&s_acc[(d.y >> 16) & M], (unsigned long long)f3);
}

template <int HUB, int BSL>
__global__ __launch_bounds__(PB_ACCUM_BLOCK) void pb_accum_kernel(
    const PbPiece *__restrict__ pieces, const float *__restrict__ vals,
    const uint16_t *__restrict__ src16, const uint16_t *__restrict__ dst16,
    const float *__restrict__ contrib, RowOut out, unsigned long long *__restrict__ sum64,
    double *__restrict__ part_res, double *__restrict__ part_dang,
    const PrState *__restrict__ st) {
  constexpr int BS = 1 << BSL;
  constexpr int CS = HUB;  // hub chunk width (a multiple of the block size)
  static_assert(HUB % PB_ACCUM_BLOCK == 0, "hub chunk width");
  constexpr int STEP = PB_ACCUM_BLOCK * 4;
  constexpr int U = 4;  // batches of loads in flight per thread
  __shared__ unsigned long long s_acc[BS];
  __shared__ float s_hot[CS];
  if (st->done) return;
  out.base = st->base;
  const int S = st->fx_shift;
  const double scale = ldexp(1.0, S);
  const PbPiece pc = pieces[blockIdx.x];
  const int32_t he = min(pc.e1, pc.hub_e1);  // chunk-0 entries of the piece: [e0, he)
  // the hub copy's loads go out first, the accumulators are cleared while
  // they are in flight
  constexpr int PER = CS / PB_ACCUM_BLOCK;
  const bool hub = pc.e0 < he;
  float c[PER];
  if (hub) {
    const int32_t n = min(CS, st->V);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int k = (int)threadIdx.x + j * PB_ACCUM_BLOCK;
      c[j] = contrib[k < n ? k : n - 1];
    }
  }
  for (int k = threadIdx.x; k < BS; k += PB_ACCUM_BLOCK) s_acc[k] = 0ull;
  if (hub) {
#pragma unroll
    for (int j = 0; j < PER; ++j) s_hot[threadIdx.x + j * PB_ACCUM_BLOCK] = c[j];
  }
  __syncthreads();
  // e0, e1, hub_e1 and every e are multiples of 4
  int32_t e = pc.e0 + (int32_t)threadIdx.x * 4;
  for (; e + (U - 1) * STEP < he; e += U * STEP) {
    pb_u32x2 q[U], d[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      q[k] = pb_ids4(src16 + pc.hub_off + e + k * STEP);
      d[k] = pb_ids4(dst16 + e + k * STEP);
    }
#pragma unroll
    for (int k = 0; k < U; ++k) pb_add4<BSL>(s_acc, pb_lookup4<CS>(s_hot, q[k]), d[k], scale);
  }
  for (; e < he; e += STEP)
    pb_add4<BSL>(s_acc, pb_lookup4<CS>(s_hot, pb_ids4(src16 + pc.hub_off + e)),
                 pb_ids4(dst16 + e), scale);
  e = max(pc.e0, pc.hub_e1) + (int32_t)threadIdx.x * 4;
  for (; e + (U - 1) * STEP < pc.e1; e += U * STEP) {
    pb_f32x4 v[U];
    pb_u32x2 d[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      v[k] = pb_vals4(vals + e + k * STEP);
      d[k] = pb_ids4(dst16 + e + k * STEP);
    }
#pragma unroll
    for (int k = 0; k < U; ++k) pb_add4<BSL>(s_acc, v[k], d[k], scale);
  }
  for (; e < pc.e1; e += STEP)
    pb_add4<BSL>(s_acc, pb_vals4(vals + e), pb_ids4(dst16 + e), scale);
  __syncthreads();
  const int32_t v0 = pc.bin << BSL;
  const int32_t nv = min(BS, st->V - v0);
  double res = 0.0, dang = 0.0;
  if (pc.multi) {
    for (int k = threadIdx.x; k < nv; k += PB_ACCUM_BLOCK) {
      const unsigned long long a = s_acc[k];
      if (a) atomicAdd(&sum64[v0 + k], a);
    }
  } else {
    // every row's old rank and 1/outdeg loads in flight together (rows
    // clamped into the bin; nv >= 1)
    constexpr int RPT = BS / PB_ACCUM_BLOCK;
    float xo[RPT], io[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int k = (int)threadIdx.x + j * PB_ACCUM_BLOCK;
      const int kc = k < nv ? k : nv - 1;
      xo[j] = out.x_old[v0 + kc];
      io[j] = out.inv_od[v0 + kc];
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int k = (int)threadIdx.x + j * PB_ACCUM_BLOCK;
      if (k < nv) out.put(v0 + k, pb_value(s_acc[k], S), xo[j], io[j], res, dang);
    }
  }
  __shared__ double s_red[PB_ACCUM_WAVES];
  res = block_sum_f64(res, s_red);
  dang = block_sum_f64(dang, s_red);
  if (threadIdx.x == 0) {
    part_res[blockIdx.x] = res;
    part_dang[blockIdx.x] = dang;
  }
}

// Rows of split bins from sum64 (cleared for the next iteration), one
// residual / dangling partial per block.  pb_fix_block (in the next
// iteration's scatter launch, or in pb_fix_kernel) then sums these and
// pb_accum's partials in a fixed order and sets the next base and the stop
// flag.
__global__ __launch_bounds__(PR_BLOCK) void pb_rows_kernel(
    int64_t nrows, int bsl, const int32_t *__restrict__ multi,
    unsigned long long *__restrict__ sum64, RowOut out, double *__restrict__ part_res,
    double *__restrict__ part_dang, const PrState *__restrict__ st) {
  __shared__ double s_red[PR_BLOCK / 64];
  if (st->done) return;
  out.base = st->base;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0, dang = 0.0;
  if (t < nrows) {
    const int32_t v = (multi[t >> bsl] << bsl) + (int32_t)(t & ((1 << bsl) - 1));
    if (v < st->V) {
      const unsigned long long a = sum64[v];
      sum64[v] = 0ull;
      out.put(v, pb_value(a, st->fx_shift), out.x_old[v], out.inv_od[v], res, dang);
    }
  }
  res = block_sum_f64(res, s_red);
  dang = block_sum_f64(dang, s_red);
  if (threadIdx.x == 0) {
    part_res[blockIdx.x] = res;
    part_dang[blockIdx.x] = dang;
  }
}

// -- binned layout build (once per graph handle and shape) --

// Per in-CSR edge u -> v (row v): key (bin of v, u), value v.  Rows are found
// by binary search on the row offsets.
__global__ void pb_keys_kernel(int32_t V, int64_t E, const int32_t *__restrict__ irp,
                               const int32_t *__restrict__ icol, int bsl, int vbits,
                               uint64_t *__restrict__ keys, int32_t *__restrict__ vs) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int32_t lo = 0, hi = V;  // irp[lo] <= e < irp[hi]
  while (hi - lo > 1) {
    const int32_t mid = lo + (hi - lo) / 2;
    if (irp[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  keys[e] = ((uint64_t)(lo >> bsl) << vbits) | (uint32_t)icol[e];
  vs[e] = lo;
}

// Chunk of source u: 0 for the hub chunk [0, hub), then chunks of 2^csl.
__host__ __device__ __forceinline__ int32_t pb_chunk(int32_t u, int hub, int csl) {
  return u < hub ? 0 : 1 + ((u - hub) >> csl);
}
__host__ __device__ __forceinline__ int32_t pb_chunk_base(int32_t c, int hub, int csl) {
  return c == 0 ? 0 : hub + ((c - 1) << csl);
}

// First sorted index of every non-empty (bin, chunk) segment.
__global__ void pb_first_kernel(int64_t E, const uint64_t *__restrict__ keys, int vbits, int hub,
                                int csl, int64_t nchunks, int32_t *__restrict__ first) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  const uint64_t mask = (1ull << vbits) - 1;
  const uint64_t key = keys[k];
  const int64_t seg = (int64_t)(key >> vbits) * nchunks + pb_chunk((int32_t)(key & mask), hub, csl);
  bool head = k == 0;
  if (!head) {
    const uint64_t p = keys[k - 1];
    head = (int64_t)(p >> vbits) * nchunks + pb_chunk((int32_t)(p & mask), hub, csl) != seg;
  }
  if (head) first[seg] = (int32_t)k;
}

// Sorted edge k of segment (b, c) goes to src16[k + d1[seg]] (chunk-major
// order) and dst16[k + d2[seg]] (bin-major order) as local ids.
__global__ void pb_place_kernel(int64_t E, const uint64_t *__restrict__ keys,
                                const int32_t *__restrict__ vs, int vbits, int hub, int csl,
                                int bsl,
                                int64_t nchunks, const int32_t *__restrict__ d1,
                                const int32_t *__restrict__ d2, uint16_t *__restrict__ src16,
                                uint16_t *__restrict__ dst16) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  const uint64_t key = keys[k];
  const int32_t b = (int32_t)(key >> vbits);
  const int32_t u = (int32_t)(key & ((1ull << vbits) - 1));
  const int32_t c = pb_chunk(u, hub, csl);
  const int64_t seg = (int64_t)b * nchunks + c;
  src16[k + d1[seg]] = (uint16_t)(u - pb_chunk_base(c, hub, csl));
  dst16[k + d2[seg]] = (uint16_t)(vs[k] - (b << bsl));
}

// Row blocks of an in-CSR (host side; irp copied from the device).
void build_row_blocks(const std::vector<int32_t> &rp, int nb, std::vector<RowBlock> &blocks,
                      std::vector<LongRow> &longs) {
  blocks.clear();
  longs.clear();
  const int32_t V = (int32_t)rp.size() - 1;
  int32_t r = 0;
  while (r < V) {
    const int32_t deg = rp[r + 1] - rp[r];
    if (deg > nb) {
      const int32_t n = (deg + nb - 1) / nb;
      longs.push_back(LongRow{r, (int32_t)blocks.size(), n});
      for (int32_t k = 0; k < n; k++)
        blocks.push_back(RowBlock{r, 0, rp[r] + k * nb, std::min(rp[r + 1], rp[r] + (k + 1) * nb)});
      r++;
      continue;
    }
    int32_t r1 = r + 1;
    while (r1 < V && r1 - r < RB_MAXR && rp[r1 + 1] - rp[r] <= nb) r1++;
    blocks.push_back(RowBlock{r, r1 - r, rp[r], rp[r1]});
    r = r1;
  }
}

struct PrWork {
  DevBuf xa, xb, ca, cb, tail, head, pres, pdang, fres, fdang, state, counter;
  // the instantiated hipGraph of CH iterations and what its kernels were
  // captured with (graph id, form, sizes, every device pointer): replayed by
  // later calls with the same key instead of captured again
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> exec_key;
  // gather timing events (time_kernels), created once and reused: creating
  // two per iteration on every call held up the first launches
  std::vector<hipEvent_t> events;
  hipError_t ensure_events(size_t n) {
    while (events.size() < n) {
      hipEvent_t e = nullptr;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      events.push_back(e);
    }
    return hipSuccess;
  }
  void release() {
    for (DevBuf *b :
         {&xa, &xb, &ca, &cb, &tail, &head, &pres, &pdang, &fres, &fdang, &state, &counter})
      b->release();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    events.clear();
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    exec_key.clear();
  }
};
// Heap singleton, never destroyed: device memory must not be freed from a
// static destructor, which can run after the HIP runtime has shut down.
PrWork &g_prw_ref() {
  static PrWork *w = new PrWork();
  return *w;
}
#define g_prw (g_prw_ref())

bool valid_ipt(int ipt) { return ipt == 4 || ipt == 8 || ipt == 12 || ipt == 16; }

// Blocks of a persistent (HOT > 0) gather: as many as are resident at once.
template <int IPT, int HOT>
int64_t gather_grid(int64_t ntiles) {
  if (HOT == 0) return ntiles;
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pr_gather_kernel<IPT, HOT>, PR_BLOCK,
                                                     0) != hipSuccess ||
        nb < 1)
      nb = 1;
    per_cu = nb;
  }
  return std::min<int64_t>(ntiles, (int64_t)gb_num_cus() * per_cu);
}

template <int IPT, int HOT>
void launch_gather(int64_t ntiles, hipStream_t st, const int32_t *irp, const int32_t *icol,
                   const int32_t *trow, const int32_t *tedge, const float *contrib, RowOut out,
                   F2 *tail, F2 *head, double *pres, double *pdang, const PrState *state) {
  const int64_t grid = gather_grid<IPT, HOT>(ntiles);
  hipLaunchKernelGGL((pr_gather_kernel<IPT, HOT>), dim3((unsigned)std::max<int64_t>(grid, 1)),
                     dim3(PR_BLOCK), 0, st, irp, icol, trow, tedge, contrib, out, tail, head, pres,
                     pdang, state, ntiles);
}

// gather variant by (items per thread, hot-set size)
template <int IPT>
void launch_gather_hot(int hot, int64_t ntiles, hipStream_t st, const int32_t *irp,
                       const int32_t *icol, const int32_t *trow, const int32_t *tedge,
                       const float *contrib, RowOut out, F2 *tail, F2 *head, double *pres,
                       double *pdang, const PrState *state) {
  switch (hot) {
    case 4096:
      launch_gather<IPT, 4096>(ntiles, st, irp, icol, trow, tedge, contrib, out, tail, head, pres,
                               pdang, state);
      break;
    case 8192:
      launch_gather<IPT, 8192>(ntiles, st, irp, icol, trow, tedge, contrib, out, tail, head, pres,
                               pdang, state);
      break;
    default:
      launch_gather<IPT, 0>(ntiles, st, irp, icol, trow, tedge, contrib, out, tail, head, pres,
                            pdang, state);
  }
}

// row-block launches: one block per row block, or the persistent hot-set
// variant on as many blocks as are resident at once
struct RbArgs {
  const int32_t *irp, *icol;
  const RowBlock *blocks;
  const float *contrib;
  RowOut out;
  F2 *chunk_sum;
  double *pres, *pdang;
  const PrState *state;
  int64_t nblocks;
};

// Returns the number of residual/dangling partials the launch writes.
template <int NB, int HOT>
int64_t launch_rb(const RbArgs &a, hipStream_t st) {
  if constexpr (HOT == 0) {
    hipLaunchKernelGGL((pr_rowblock_kernel<NB>), dim3((unsigned)std::max<int64_t>(a.nblocks, 1)),
                       dim3(PR_BLOCK), 0, st, a.irp, a.icol, a.blocks, a.contrib, a.out,
                       a.chunk_sum, a.pres, a.pdang, a.state);
    return std::max<int64_t>(a.nblocks, 1) * (PR_BLOCK / GB_WAVE);
  } else {
    constexpr int SUBS = rb_subs(NB, HOT);
    static int per_cu = 0;
    if (per_cu == 0) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pr_rowblock_hot_kernel<NB, HOT>,
                                                       PR_BLOCK * SUBS, 0) != hipSuccess ||
          nb < 1)
        nb = 1;
      per_cu = nb;
    }
    const int64_t need = (a.nblocks + SUBS - 1) / SUBS;
    const int64_t grid =
        std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)gb_num_cus() * per_cu));
    // grid * SUBS <= nblocks + SUBS - 1: at most 4 * nblocks + 12 partials
    hipLaunchKernelGGL((pr_rowblock_hot_kernel<NB, HOT>), dim3((unsigned)grid),
                       dim3(PR_BLOCK * SUBS), 0,
                       st, a.irp, a.icol, a.blocks, a.contrib, a.out, a.chunk_sum, a.pres, a.pdang,
                       a.state, a.nblocks);
    return grid * SUBS * (PR_BLOCK / GB_WAVE);
  }
}

template <int NB>
int64_t launch_rb_hot(int hot, const RbArgs &a, hipStream_t st) {
  switch (hot) {
    case 4096:
      return launch_rb<NB, 4096>(a, st);
    case 8192:
      return launch_rb<NB, 8192>(a, st);
    default:
      return launch_rb<NB, 0>(a, st);
  }
}

}  // namespace

void gb_pr_release() { g_prw.release(); }

static gb_status ensure_inv_outdeg(gb_graph g, hipStream_t st) {
  if (g->inv_outdeg_valid) return GB_OK;
  GB_TRY(g->inv_outdeg.ensure(sizeof(float) * (g->V + 1)));
  int blocks = (int)((g->V + 255) / 256);
  if (g->V == 0) return GB_OK;
  if (g->rp) {
    hipLaunchKernelGGL(inv_outdeg_kernel, dim3(blocks), dim3(256), 0, st, g->V, g->rp,
                       g->inv_outdeg.as<float>());
  } else {
    DevBuf deg;
    GB_TRY(deg.ensure(sizeof(int32_t) * g->V));
    GB_HIP(hipMemsetAsync(deg.p, 0, sizeof(int32_t) * g->V, st));
    int eb = (int)((g->E + 255) / 256);
    if (eb > 0)
      hipLaunchKernelGGL(count_outdeg_kernel, dim3(eb), dim3(256), 0, st, g->E, g->rcol,
                         deg.as<int32_t>());
    hipLaunchKernelGGL(inv_from_deg_kernel, dim3(blocks), dim3(256), 0, st, g->V,
                       deg.as<int32_t>(), g->inv_outdeg.as<float>());
    GB_HIP(hipStreamSynchronize(st));
  }
  GB_HIP(hipGetLastError());
  g->inv_outdeg_valid = true;
  return GB_OK;
}

// Binned layout of the gather's in-CSR (irp, icol) for a hub chunk of hub
// sources, chunks of 2^csl sources after it and bins of 2^bsl destinations,
// cached on the handle.  Edges are
// sorted by (bin, source) with a stable radix sort (in-CSR order, so by
// destination, among equal keys); each (bin, chunk) segment is padded to a
// multiple of 4 entries.  Host side: segment offsets in both orders, wave
// tasks (<= 256 entries of one segment), units (runs of one chunk's tasks of
// about `unit` entries: one scatter block each) and pieces (runs of a bin's
// stream of at most `piece` entries: one accumulation block each; largest
// first, so that the block scheduler hands out the long ones first).
static gb_status pb_build(gb_graph g, bool relabel, const int32_t *irp, const int32_t *icol,
                          int hub, int csl, int bsl, int piece, hipStream_t st) {
  gb_graph_s::PrBinned &L = g->pr_pb;
  if (L.valid && L.relabeled == relabel && L.hub == hub && L.csl == csl && L.bsl == bsl &&
      L.piece == piece)
    return GB_OK;
  L.valid = false;
  const int32_t V = g->V;
  const int64_t E = g->E;
  const int64_t nchunks =
      1 + (V > hub ? (((int64_t)V - hub + (1 << csl) - 1) >> csl) : 0);
  const int64_t nbins = ((int64_t)V + (1 << bsl) - 1) >> bsl;
  const int64_t nseg = nchunks * nbins;
  if (nseg > ((int64_t)1 << 26) ||
      E + (PB_ALIGN - 1) * std::min<int64_t>(nseg, E) > INT32_MAX - PB_ALIGN)
    return gb_fail(GB_ERR_UNSUPPORTED, "gb_pagerank: graph too large for the binned form");
  int vbits = 1;
  while (((int64_t)1 << vbits) < V) vbits++;
  int bbits = 1;
  while (((int64_t)1 << bbits) < nbins) bbits++;
  // sorted (bin, source) keys and destinations
  std::vector<int32_t> first((size_t)nseg, -1);
  DevBuf keys, keys2, vs, vs2, dfirst, tmp;
  if (E > 0) {
    GB_TRY(keys.ensure(sizeof(uint64_t) * E));
    GB_TRY(keys2.ensure(sizeof(uint64_t) * E));
    GB_TRY(vs.ensure(sizeof(int32_t) * E));
    GB_TRY(vs2.ensure(sizeof(int32_t) * E));
    const unsigned eb = (unsigned)((E + 255) / 256);
    hipLaunchKernelGGL(pb_keys_kernel, dim3(eb), dim3(256), 0, st, V, E, irp, icol, bsl, vbits,
                       keys.as<uint64_t>(), vs.as<int32_t>());
    size_t tb = 0;
    GB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint64_t>(),
                                              keys2.as<uint64_t>(), vs.as<int32_t>(),
                                              vs2.as<int32_t>(), (int)E, 0, vbits + bbits, st));
    GB_TRY(tmp.ensure(tb));
    GB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.as<uint64_t>(),
                                              keys2.as<uint64_t>(), vs.as<int32_t>(),
                                              vs2.as<int32_t>(), (int)E, 0, vbits + bbits, st));
    GB_HIP(hipStreamSynchronize(st));
    tmp.release();
    keys.release();
    vs.release();
    GB_TRY(dfirst.ensure(sizeof(int32_t) * nseg));
    GB_HIP(hipMemsetAsync(dfirst.p, 0xFF, sizeof(int32_t) * nseg, st));
    hipLaunchKernelGGL(pb_first_kernel, dim3(eb), dim3(256), 0, st, E, keys2.as<uint64_t>(),
                       vbits, hub, csl, nchunks, dfirst.as<int32_t>());
    GB_HIP(hipMemcpyAsync(first.data(), dfirst.p, sizeof(int32_t) * nseg, hipMemcpyDeviceToHost,
                          st));
    GB_HIP(hipStreamSynchronize(st));
  }
  // segment sizes (bin-major = sorted order), padded sizes and offsets
  std::vector<int32_t> cnt((size_t)nseg, 0);
  {
    int64_t next = E;
    for (int64_t i = nseg - 1; i >= 0; --i)
      if (first[i] >= 0) {
        cnt[i] = (int32_t)(next - first[i]);
        next = first[i];
      }
  }
  auto pad4 = [](int64_t n) { return (n + 3) & ~(int64_t)3; };
  auto padseg = [](int64_t n) { return (n + PB_ALIGN - 1) & ~(int64_t)(PB_ALIGN - 1); };
  std::vector<int32_t> d1((size_t)nseg), d2((size_t)nseg);
  std::vector<int32_t> out_start((size_t)nseg);
  std::vector<int32_t> bin_start((size_t)nbins + 1);
  int64_t P = 0;
  for (int64_t b = 0; b < nbins; ++b) {
    bin_start[b] = (int32_t)P;
    for (int64_t c = 0; c < nchunks; ++c) {
      const int64_t i = b * nchunks + c;
      out_start[i] = (int32_t)P;
      d2[i] = (int32_t)(P - (first[i] >= 0 ? first[i] : 0));
      P += padseg(cnt[i]);
    }
  }
  bin_start[nbins] = (int32_t)P;
  std::vector<PbTask> tasks;
  std::vector<PbUnit> units;
  const int64_t unit = 65536;
  std::vector<int32_t> hub_in((size_t)nbins);  // chunk-major offset of segment (0, b)
  {
    int64_t in = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
      int32_t ut0 = (int32_t)tasks.size();
      int64_t ue = 0;
      for (int64_t b = 0; b < nbins; ++b) {
        const int64_t i = b * nchunks + c;
        d1[i] = (int32_t)(in - (first[i] >= 0 ? first[i] : 0));
        const int64_t n = padseg(cnt[i]);
        if (c == 0) {  // read by pb_accum from src16, no scatter
          hub_in[b] = (int32_t)in;
          in += n;
          continue;
        }
        for (int64_t j = 0; j < n; j += PB_TASK) {
          tasks.push_back(PbTask{(int32_t)(in + j), (int32_t)(out_start[i] + j),
                                 (int32_t)std::min<int64_t>(PB_TASK, n - j)});
          ue += std::min<int64_t>(PB_TASK, n - j);
          if (ue >= unit) {
            units.push_back(PbUnit{pb_chunk_base((int32_t)c, hub, csl), ut0, (int32_t)tasks.size()});
            ut0 = (int32_t)tasks.size();
            ue = 0;
          }
        }
        in += n;
      }
      if ((int32_t)tasks.size() > ut0)
        units.push_back(PbUnit{pb_chunk_base((int32_t)c, hub, csl), ut0, (int32_t)tasks.size()});
    }
  }
  std::vector<PbPiece> pieces;
  std::vector<int32_t> multi;
  for (int64_t b = 0; b < nbins; ++b) {
    const int64_t s = bin_start[b], n = bin_start[b + 1] - s;
    const int64_t k = std::max<int64_t>(1, (n + piece - 1) / piece);
    if (k > 1) multi.push_back((int32_t)b);
    const int32_t hub_e1 = (int32_t)(s + padseg(cnt[b * nchunks]));
    for (int64_t j = 0; j < k; ++j)
      pieces.push_back(PbPiece{(int32_t)b, (int32_t)(s + pad4(n * j / k)),
                               (int32_t)(j + 1 == k ? s + n : s + pad4(n * (j + 1) / k)),
                               k > 1 ? 1 : 0, hub_e1, (int32_t)(hub_in[b] - s)});
  }
  std::stable_sort(pieces.begin(), pieces.end(), [](const PbPiece &x, const PbPiece &y) {
    return x.e1 - x.e0 > y.e1 - y.e0;
  });
  // device copies
  GB_TRY(L.src16.ensure(sizeof(uint16_t) * (P + 4)));
  GB_TRY(L.dst16.ensure(sizeof(uint16_t) * (P + 4)));
  GB_TRY(L.vals.ensure(sizeof(float) * (P + 4)));
  GB_TRY(L.tasks.ensure(sizeof(PbTask) * (tasks.size() + 1)));
  GB_TRY(L.units.ensure(sizeof(PbUnit) * (units.size() + 1)));
  GB_TRY(L.pieces.ensure(sizeof(PbPiece) * (pieces.size() + 1)));
  GB_TRY(L.multi.ensure(sizeof(int32_t) * (multi.size() + 1)));
  GB_TRY(L.sum64.ensure(sizeof(unsigned long long) * V));
  GB_HIP(hipMemsetAsync(L.src16.p, 0xFF, sizeof(uint16_t) * (P + 4), st));
  GB_HIP(hipMemsetAsync(L.dst16.p, 0, sizeof(uint16_t) * (P + 4), st));
  GB_HIP(hipMemsetAsync(L.sum64.p, 0, sizeof(unsigned long long) * V, st));
  if (E > 0) {
    DevBuf dd1, dd2;
    GB_TRY(dd1.ensure(sizeof(int32_t) * nseg));
    GB_TRY(dd2.ensure(sizeof(int32_t) * nseg));
    GB_HIP(hipMemcpyAsync(dd1.p, d1.data(), sizeof(int32_t) * nseg, hipMemcpyHostToDevice, st));
    GB_HIP(hipMemcpyAsync(dd2.p, d2.data(), sizeof(int32_t) * nseg, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(pb_place_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, st, E,
                       keys2.as<uint64_t>(), vs2.as<int32_t>(), vbits, hub, csl, bsl, nchunks,
                       dd1.as<int32_t>(), dd2.as<int32_t>(), L.src16.as<uint16_t>(),
                       L.dst16.as<uint16_t>());
    GB_HIP(hipStreamSynchronize(st));  // dd1 / dd2 go out of scope
  }
  if (!tasks.empty())
    GB_HIP(hipMemcpyAsync(L.tasks.p, tasks.data(), sizeof(PbTask) * tasks.size(),
                          hipMemcpyHostToDevice, st));
  if (!units.empty())
    GB_HIP(hipMemcpyAsync(L.units.p, units.data(), sizeof(PbUnit) * units.size(),
                          hipMemcpyHostToDevice, st));
  GB_HIP(hipMemcpyAsync(L.pieces.p, pieces.data(), sizeof(PbPiece) * pieces.size(),
                        hipMemcpyHostToDevice, st));
  if (!multi.empty())
    GB_HIP(hipMemcpyAsync(L.multi.p, multi.data(), sizeof(int32_t) * multi.size(),
                          hipMemcpyHostToDevice, st));
  GB_HIP(hipStreamSynchronize(st));  // host vectors go out of scope
  GB_HIP(hipGetLastError());
  L.valid = true;
  L.relabeled = relabel;
  L.hub = hub;
  L.csl = csl;
  L.bsl = bsl;
  L.piece = piece;
  L.padded = P;
  L.ntasks = (int64_t)tasks.size();
  L.nunits = (int64_t)units.size();
  L.npieces = (int64_t)pieces.size();
  L.nmulti = (int64_t)multi.size();
  return GB_OK;
}

// The gather's form and shape: merge path with ipt items per thread (tile =
// 256 x ipt items) and an LDS hot set of `hot` contributions, row blocks of
// rb edges (rb > 0; ipt and hot unused), or the binned form (pb > 0: chunk
// hub, chunk and bin widths PB_SHAPES[pb]; ipt, hot and rb unused).
struct PrShape {
  int ipt, hot, rb, pb;
};
struct PbShape {
  int hub, csl, bsl;  // hub chunk width, log2 of the chunk and bin widths
};
constexpr PbShape PB_SHAPES[] = {{0, 0, 0},         {8192, 13, 13},  {16384, 14, 13},
                                 {8192, 13, 12},    {16384, 13, 13}, {20480, 13, 13},
                                 {16384, 12, 13}};
constexpr int PB_NSHAPES = 7;
// stream entries per accumulation block at most (GB_PR_PB_PIECE, multiple of 4)
static int pb_piece() {
  const char *e = getenv("GB_PR_PB_PIECE");
  const int v = e ? atoi(e) : 0;
  return v >= 4096 && v <= (1 << 26) ? v & ~3 : 262144;
}
static bool valid_rb(int rb) { return rb == 1024 || rb == 2048 || rb == 4096; }

static gb_status pagerank_run(gb_graph g, const gb_pr_params *p, float *d_x, gb_pr_result *res,
                              void *stream, PrShape sh, double *iter_ms = nullptr);

static bool use_relabel(gb_graph g, const gb_pr_params *p) {
  return p->relabel > 0 || (p->relabel == 0 && g->E >= ((int64_t)1 << 20));
}

// Gather form and shape for a graph.  gb_pr_params.row_block > 0 or GB_PR_RB
// fix the row-block form; items_per_thread (or GB_PR_IPT) and GB_PR_HOT fix
// merge-path shapes.  Otherwise the autotuner runs once per graph handle and
// layout: three timed iterations for each candidate, keeping the fastest mean
// gather.  Merge-path candidates: small tiles cost staging and scan work per
// item, large ones occupancy and rows that span threads; a hot set (relabelled
// graphs only) takes the most-read sources off the L2 but its LDS lowers
// occupancy.  Row-block candidates (unless row_block is -1) skip the merge-path
// search and scan but leave lanes idle in blocks of uneven rows.  The binned
// candidate (with the row-block ones, unless binned is -1) streams every edge
// twice but issues no random reads.  Candidates are compared by the mean time
// of a whole iteration (gather and fix-up).
static gb_status pick_shape(gb_graph g, const gb_pr_params *p, float *d_x, void *stream,
                            PrShape *out) {
  const bool rl = use_relabel(g, p);
  const char *ei = getenv("GB_PR_IPT");
  const char *eh = getenv("GB_PR_HOT");
  const char *er = getenv("GB_PR_RB");
  const char *ep = getenv("GB_PR_PB");
  const bool free_form = p->row_block == 0 && !p->items_per_thread && p->binned == 0;
  int fixed_pb = p->binned > 0 ? p->binned : 0;
  if (!fixed_pb && free_form && ep && atoi(ep) > 0 && atoi(ep) < PB_NSHAPES) fixed_pb = atoi(ep);
  if (fixed_pb) {
    *out = PrShape{0, 0, 0, fixed_pb};
    return GB_OK;
  }
  int fixed_rb = p->row_block > 0 ? p->row_block : 0;
  if (!fixed_rb && p->row_block == 0 && !p->items_per_thread && er && valid_rb(atoi(er)))
    fixed_rb = atoi(er);
  if (fixed_rb) {
    const int h = eh ? atoi(eh) : 0;
    *out = PrShape{0, rl && (h == 4096 || h == 8192) ? h : 0, fixed_rb, 0};
    return GB_OK;
  }
  // fixed parts of the shape (0 / -1 = left to the tuner)
  int fixed_ipt = p->items_per_thread;
  if (!fixed_ipt && ei && valid_ipt(atoi(ei))) fixed_ipt = atoi(ei);
  int fixed_hot = -1;
  if (eh && (atoi(eh) == 0 || atoi(eh) == 4096 || atoi(eh) == 8192)) fixed_hot = atoi(eh);
  if (!rl) fixed_hot = 0;
  if (fixed_ipt && fixed_hot >= 0) {
    *out = PrShape{fixed_ipt, fixed_hot, 0, 0};
    return GB_OK;
  }
  const bool try_rb = !fixed_ipt && fixed_hot <= 0 && p->row_block == 0;
  const bool try_pb = try_rb && p->binned == 0;
  // cached per (fixed ipt or none, row blocks allowed, layout, binned allowed)
  const int key = (fixed_ipt ? fixed_ipt / 4 : 0) * 2 + (rl ? 1 : 0) + (try_rb ? 0 : 10) +
                  (try_pb ? 0 : 20);
  if (g->pr_tuned[key][0] || g->pr_tuned[key][2] || g->pr_tuned[key][3]) {
    *out = PrShape{g->pr_tuned[key][0], g->pr_tuned[key][1], g->pr_tuned[key][2],
                   g->pr_tuned[key][3]};
    return GB_OK;
  }
  std::vector<PrShape> cands;
  for (int ci : fixed_ipt ? std::vector<int>{fixed_ipt} : std::vector<int>{4, 8, 12, 16})
    for (int ch : fixed_hot >= 0 ? std::vector<int>{fixed_hot} : std::vector<int>{0, 4096, 8192})
      cands.push_back(PrShape{ci, ch, 0, 0});
  if (try_rb)
    for (int rb : {1024, 2048, 4096})
      for (int ch : rl ? std::vector<int>{0, 4096, 8192} : std::vector<int>{0})
        cands.push_back(PrShape{0, ch, rb, 0});
  if (try_pb) cands.push_back(PrShape{0, 0, 0, 4});
  PrShape best = cands[0];
  double best_ms = 1e300;
  for (const PrShape &c : cands) {
    gb_pr_params q = *p;
    q.max_iter = 3;
    q.tol = 0.0;
    q.use_graph = 0;
    q.time_kernels = 1;
    q.d_x0 = nullptr;
    gb_pr_result r{};
    double it_ms = 0.0;
    const gb_status s = pagerank_run(g, &q, d_x, &r, stream, c, &it_ms);
    if (s == GB_ERR_UNSUPPORTED && c.pb) continue;  // graph too large for the binned layout
    GB_TRY(s);
    const double ms = r.gather_launches ? it_ms / r.gather_launches : 1e300;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  g->pr_tuned[key][0] = best.ipt;
  g->pr_tuned[key][1] = best.hot;
  g->pr_tuned[key][2] = best.rb;
  g->pr_tuned[key][3] = best.pb;
  *out = best;
  return GB_OK;
}

extern "C" gb_status gb_pagerank(gb_graph g, const gb_pr_params *p, float *d_x,
                                 gb_pr_result *res, void *stream) {
  if (!g || !p || !d_x) return gb_fail(GB_ERR_INVALID, "gb_pagerank: null argument");
  if (p->items_per_thread != 0 && !valid_ipt(p->items_per_thread))
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: items_per_thread must be 0, 4, 8, 12 or 16");
  if (!g->rrp || (!g->rcol && g->E > 0))
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: graph has no reverse CSR");
  if (g->V <= 0) return gb_fail(GB_ERR_INVALID, "gb_pagerank: empty graph");
  if (!(p->damping >= 0.0 && p->damping <= 1.0) || p->max_iter < 0)
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: bad damping/max_iter");
  if (p->row_block < -1 || (p->row_block > 0 && !valid_rb(p->row_block)))
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: row_block must be -1, 0, 1024, 2048 or 4096");
  if (p->row_block > 0 && p->items_per_thread != 0)
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: row_block and items_per_thread both set");
  if (p->binned < -1 || p->binned >= PB_NSHAPES)
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: binned must be -1 or 0 to 6");
  if (p->binned > 0 && (p->row_block != 0 || p->items_per_thread != 0))
    return gb_fail(GB_ERR_INVALID, "gb_pagerank: binned set with row_block or items_per_thread");
  PrShape sh{8, 0, 0, 0};
  GB_TRY(ensure_inv_outdeg(g, gb_stream(stream)));
  GB_TRY(pick_shape(g, p, d_x, stream, &sh));
  return pagerank_run(g, p, d_x, res, stream, sh);
}

static gb_status pagerank_run(gb_graph g, const gb_pr_params *p, float *d_x, gb_pr_result *res,
                              void *stream, PrShape sh, double *iter_ms) {
  const int ipt = sh.rb || sh.pb ? 8 : sh.ipt, rb = sh.rb, pb = sh.pb;
  int hot = sh.hot;
  hipStream_t st = gb_stream(stream);
  hipStream_t own = nullptr;
  if (!st && p->use_graph && !p->time_kernels) {
    // the null stream cannot be captured: use a blocking stream (it orders
    // itself against the null stream) for this call
    GB_HIP(hipStreamCreate(&own));
    st = own;
  }
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } guard{own};
  const int32_t V = g->V;
  const int64_t tile = (int64_t)PR_BLOCK * ipt;
  GB_TRY(ensure_inv_outdeg(g, st));
  const bool relabel = use_relabel(g, p);
  if (relabel) GB_TRY(gb_graph_pr_relabel(g, st));
  if (!relabel) hot = 0;
  // the in-CSR and 1/outdeg the gather runs on
  const int32_t *irp = relabel ? g->rl_rrp.as<int32_t>() : g->rrp;
  const int32_t *icol = relabel ? g->rl_rcol.as<int32_t>() : g->rcol;
  const float *inv = relabel ? g->rl_inv.as<float>() : g->inv_outdeg.as<float>();
  if (pb)
    GB_TRY(pb_build(g, relabel, irp, icol, PB_SHAPES[pb].hub, PB_SHAPES[pb].csl,
                    PB_SHAPES[pb].bsl, pb_piece(), st));
  if (!rb && !pb && (g->pr_tile_items != tile || g->pr_tile_relabeled != relabel)) {
    const int64_t ntiles = ((int64_t)V + g->E + tile - 1) / tile;
    GB_TRY(g->pr_tile_row.ensure(sizeof(int32_t) * (ntiles + 1)));
    GB_TRY(g->pr_tile_edge.ensure(sizeof(int32_t) * (ntiles + 1)));
    GB_TRY(gb_merge_path_split(V, irp, tile, ntiles, g->pr_tile_row.as<int32_t>(),
                               g->pr_tile_edge.as<int32_t>(), st));
    g->pr_tile_items = (int32_t)tile;
    g->pr_tile_relabeled = relabel;
    g->pr_ntiles = ntiles;
  }
  if (rb && (g->pr_rb_nb != rb || g->pr_rb_relabeled != relabel)) {
    std::vector<int32_t> hrp((size_t)V + 1);
    GB_HIP(hipMemcpyAsync(hrp.data(), irp, sizeof(int32_t) * (V + 1), hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    std::vector<RowBlock> blocks;
    std::vector<LongRow> longs;
    build_row_blocks(hrp, rb, blocks, longs);
    GB_TRY(g->pr_rb.ensure(sizeof(RowBlock) * blocks.size()));
    GB_TRY(g->pr_rb_long.ensure(sizeof(LongRow) * (longs.size() + 1)));
    GB_HIP(hipMemcpyAsync(g->pr_rb.p, blocks.data(), sizeof(RowBlock) * blocks.size(),
                          hipMemcpyHostToDevice, st));
    if (!longs.empty())
      GB_HIP(hipMemcpyAsync(g->pr_rb_long.p, longs.data(), sizeof(LongRow) * longs.size(),
                            hipMemcpyHostToDevice, st));
    GB_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    g->pr_rb_nb = rb;
    g->pr_rb_relabeled = relabel;
    g->pr_rb_nblocks = (int64_t)blocks.size();
    g->pr_rb_nlong = (int32_t)longs.size();
  }
  // merge path: tiles and their partials; row blocks: blocks, 4 partials each;
  // binned: pieces, 8 partials each, and the rows of split bins
  const gb_graph_s::PrBinned &PB = g->pr_pb;
  const int64_t ntiles = pb ? PB.npieces : rb ? g->pr_rb_nblocks : g->pr_ntiles;
  const int64_t pb_rows = pb ? PB.nmulti << PB.bsl : 0;
  const int64_t pb_row_blocks = (pb_rows + PR_BLOCK - 1) / PR_BLOCK;
  const int64_t nparts = pb   ? ntiles + pb_row_blocks
                         : rb ? ntiles * (PR_BLOCK / GB_WAVE)
                              : ntiles;
  const int64_t nfix =
      (std::max<int64_t>(nparts, rb ? g->pr_rb_nlong : 0) + PR_BLOCK - 1) / PR_BLOCK;
  const int64_t vblocks = ((int64_t)V + PR_BLOCK - 1) / PR_BLOCK;
  PrWork &w = g_prw;
  GB_TRY(w.xa.ensure(sizeof(float) * V));
  GB_TRY(w.xb.ensure(sizeof(float) * V));
  GB_TRY(w.ca.ensure(sizeof(float) * V));
  GB_TRY(w.cb.ensure(sizeof(float) * V));
  GB_TRY(w.tail.ensure(sizeof(F2) * ntiles));
  GB_TRY(w.head.ensure(sizeof(F2) * ntiles));
  // + 16: the persistent row-block form writes up to 4 * nblocks + 12 partials
  GB_TRY(w.pres.ensure(sizeof(double) * std::max(nparts + 16, vblocks)));
  GB_TRY(w.pdang.ensure(sizeof(double) * std::max(nparts + 16, vblocks)));
  GB_TRY(w.fres.ensure(sizeof(double) * (nfix + 1)));
  GB_TRY(w.fdang.ensure(sizeof(double) * (nfix + 1)));
  GB_TRY(w.state.ensure(sizeof(PrState)));
  GB_TRY(w.counter.ensure(sizeof(unsigned int)));
  GB_HIP(hipMemsetAsync(w.counter.p, 0, sizeof(unsigned int), st));
  PrState hs{};
  hs.tol = p->tol;
  hs.damping = (float)p->damping;
  hs.max_iter = p->max_iter;
  hs.V = V;
  GB_HIP(hipMemcpyAsync(w.state.p, &hs, sizeof(hs), hipMemcpyHostToDevice, st));
  PrState *state = w.state.as<PrState>();
  float *xa = w.xa.as<float>(), *xb = w.xb.as<float>();
  float *ca = w.ca.as<float>(), *cb = w.cb.as<float>();
  hipLaunchKernelGGL(pr_init_kernel, dim3((unsigned)vblocks), dim3(PR_BLOCK), 0, st, V, p->d_x0,
                     relabel ? (const int32_t *)g->rl_perm.as<int32_t>() : nullptr, inv, xa, ca,
                     w.pdang.as<double>(), w.pres.as<double>());
  hipLaunchKernelGGL(pr_init_finalize_kernel, dim3(1), dim3(1024), 0, st, (int32_t)vblocks,
                     w.pdang.as<double>(), (const double *)w.pres.as<double>(), state);
  GB_HIP(hipGetLastError());

  const int32_t *trow = g->pr_tile_row.as<int32_t>();
  const int32_t *tedge = g->pr_tile_edge.as<int32_t>();
  // One iteration; ev0..ev1 (optional) bracket the gather launches.  Binned
  // form: fold, the previous iteration's end runs in this one's scatter
  // launch; defer, this iteration's end is left to the next one's.
  const bool can_fold = pb && PB.nunits > 0;
  auto one_iter = [&](const float *x_old, const float *c_old, float *x_new, float *c_new,
                      hipEvent_t ev0, hipEvent_t ev1, bool fold, bool defer) {
    RowOut out{0.0f, (float)p->damping, x_old, inv, x_new, c_new};
    if (ev0) (void)hipEventRecord(ev0, st);
    if (pb) {
      if (PB.nunits > 0) {
        auto scatter = PB.csl == 14   ? pb_scatter_kernel<14>
                       : PB.csl == 12 ? pb_scatter_kernel<12>
                                      : pb_scatter_kernel<13>;
        const int f = fold ? 1 : 0;
        hipLaunchKernelGGL(scatter, dim3((unsigned)PB.nunits + f), dim3(PB_SCATTER_BLOCK), 0,
                           st, PB.units.as<PbUnit>(), PB.tasks.as<PbTask>(),
                           PB.src16.as<uint16_t>(), c_old, PB.vals.as<float>(), f, nparts,
                           (const double *)w.pres.as<double>(),
                           (const double *)w.pdang.as<double>(), state);
      }
      auto accum = PB.hub == 20480  ? pb_accum_kernel<20480, 13>
                   : PB.hub == 16384 ? pb_accum_kernel<16384, 13>
                   : PB.bsl == 12    ? pb_accum_kernel<8192, 12>
                                     : pb_accum_kernel<8192, 13>;
      hipLaunchKernelGGL(accum, dim3((unsigned)PB.npieces), dim3(PB_ACCUM_BLOCK), 0, st,
                         PB.pieces.as<PbPiece>(), (const float *)PB.vals.as<float>(),
                         (const uint16_t *)PB.src16.as<uint16_t>(), PB.dst16.as<uint16_t>(),
                         c_old, out, PB.sum64.as<unsigned long long>(), w.pres.as<double>(),
                         w.pdang.as<double>(), state);
      if (ev1) (void)hipEventRecord(ev1, st);
      if (pb_rows > 0)
        hipLaunchKernelGGL(pb_rows_kernel, dim3((unsigned)pb_row_blocks), dim3(PR_BLOCK), 0, st,
                           pb_rows, PB.bsl, (const int32_t *)PB.multi.as<int32_t>(),
                           PB.sum64.as<unsigned long long>(), out, w.pres.as<double>() + ntiles,
                           w.pdang.as<double>() + ntiles, state);
      // partials: one per piece, then one per row block, summed in that order
      if (!defer)
        hipLaunchKernelGGL(pb_fix_kernel, dim3(1), dim3(PB_SCATTER_BLOCK), 0, st, nparts,
                           (const double *)w.pres.as<double>(),
                           (const double *)w.pdang.as<double>(), state);
      return;
    }
    if (rb) {
      const RbArgs a{irp, icol, g->pr_rb.as<RowBlock>(), c_old, out, w.tail.as<F2>(),
                     w.pres.as<double>(), w.pdang.as<double>(), state, ntiles};
      const int64_t np = rb == 1024   ? launch_rb_hot<1024>(hot, a, st)
                         : rb == 2048 ? launch_rb_hot<2048>(hot, a, st)
                                      : launch_rb_hot<4096>(hot, a, st);
      if (ev1) (void)hipEventRecord(ev1, st);
      if (np <= RB_FIX1_MAX_PARTS && g->pr_rb_nlong <= RB_FIX1_MAX_LONG) {
        hipLaunchKernelGGL(pr_rb_fixup_one_kernel, dim3(1), dim3(RB_FIX1_THREADS), 0, st, np,
                           g->pr_rb_nlong, g->pr_rb_long.as<LongRow>(),
                           (const F2 *)w.tail.as<F2>(), out, w.pres.as<double>(),
                           w.pdang.as<double>(), state);
        return;
      }
      const int64_t nf = (std::max<int64_t>(np, g->pr_rb_nlong) + PR_BLOCK - 1) / PR_BLOCK;
      hipLaunchKernelGGL(pr_rb_fixup_kernel, dim3((unsigned)nf), dim3(PR_BLOCK), 0, st, np,
                         g->pr_rb_nlong, g->pr_rb_long.as<LongRow>(), (const F2 *)w.tail.as<F2>(),
                         out, w.pres.as<double>(), w.pdang.as<double>(), w.fres.as<double>(),
                         w.fdang.as<double>(), w.counter.as<unsigned int>(), state);
      return;
    }
    switch (ipt) {
      case 4:
        launch_gather_hot<4>(hot, ntiles, st, irp, icol, trow, tedge, c_old, out, w.tail.as<F2>(),
                             w.head.as<F2>(), w.pres.as<double>(), w.pdang.as<double>(), state);
        break;
      case 12:
        launch_gather_hot<12>(hot, ntiles, st, irp, icol, trow, tedge, c_old, out,
                              w.tail.as<F2>(), w.head.as<F2>(), w.pres.as<double>(),
                              w.pdang.as<double>(), state);
        break;
      case 16:
        launch_gather_hot<16>(hot, ntiles, st, irp, icol, trow, tedge, c_old, out,
                              w.tail.as<F2>(), w.head.as<F2>(), w.pres.as<double>(),
                              w.pdang.as<double>(), state);
        break;
      default:
        launch_gather_hot<8>(hot, ntiles, st, irp, icol, trow, tedge, c_old, out, w.tail.as<F2>(),
                             w.head.as<F2>(), w.pres.as<double>(), w.pdang.as<double>(), state);
    }
    if (ev1) (void)hipEventRecord(ev1, st);
    hipLaunchKernelGGL(pr_fixup_kernel, dim3((unsigned)nfix), dim3(PR_BLOCK), 0, st, ntiles,
                       irp, trow, tedge, w.tail.as<F2>(), w.head.as<F2>(), out,
                       w.pres.as<double>(), w.pdang.as<double>(), w.fres.as<double>(),
                       w.fdang.as<double>(), w.counter.as<unsigned int>(), state);
  };

  PrState host_state{};
  const int CH = 10;  // iterations per chunk (even: buffer parity restarts)
  double gather_ms = 0.0;
  int gather_launches = 0;
  if (p->use_graph && !p->time_kernels) {
    // everything the captured kernels take as arguments
    const gb_graph_s::PrBinned &L = g->pr_pb;
    auto u = [](const void *q) { return (uint64_t)(uintptr_t)q; };
    float dmp = (float)p->damping;
    uint32_t dbits;
    std::memcpy(&dbits, &dmp, sizeof(dbits));
    const std::vector<uint64_t> key = {
        g->uid, (uint64_t)ipt, (uint64_t)hot, (uint64_t)rb, (uint64_t)pb, relabel ? 1u : 0u,
        dbits, (uint64_t)ntiles, (uint64_t)nparts, (uint64_t)nfix, (uint64_t)g->pr_rb_nlong,
        u(irp), u(icol), u(inv), u(trow), u(tedge), u(g->pr_rb.p), u(g->pr_rb_long.p), u(xa),
        u(xb), u(ca), u(cb), u(w.tail.p), u(w.head.p), u(w.pres.p), u(w.pdang.p), u(w.fres.p),
        u(w.fdang.p), u(w.counter.p), u(state), u(L.src16.p), u(L.dst16.p), u(L.vals.p),
        u(L.tasks.p), u(L.units.p), u(L.pieces.p), u(L.multi.p), u(L.sum64.p),
        (uint64_t)L.nunits, (uint64_t)L.npieces, (uint64_t)L.nmulti, (uint64_t)L.hub,
        (uint64_t)L.csl, (uint64_t)L.bsl};
    if (!w.exec || w.exec_key != key) {
      if (w.exec) (void)hipGraphExecDestroy(w.exec);
      w.exec = nullptr;
      w.exec_key.clear();
      hipGraph_t graph = nullptr;
      GB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < CH; k++) {
        const bool fold = can_fold && k > 0, defer = can_fold && k + 1 < CH;
        if (k % 2 == 0)
          one_iter(xa, ca, xb, cb, nullptr, nullptr, fold, defer);
        else
          one_iter(xb, cb, xa, ca, nullptr, nullptr, fold, defer);
      }
      GB_HIP(hipStreamEndCapture(st, &graph));
      const hipError_t ie = hipGraphInstantiate(&w.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      GB_HIP(ie);
      w.exec_key = key;
    }
    for (int done_it = 0; done_it < p->max_iter; done_it += CH) {
      GB_HIP(hipGraphLaunch(w.exec, st));
      GB_HIP(hipMemcpyAsync(&host_state, state, sizeof(PrState), hipMemcpyDeviceToHost, st));
      GB_HIP(hipStreamSynchronize(st));
      if (host_state.done) break;
    }
  } else {
    // eager launches; with time_kernels every gather is bracketed by its own
    // event pair, a third event ends the iteration when the autotuner asks
    // for iteration times (every event record holds up the next launch by a
    // few microseconds), and the stop flag is read every CH iterations only
    if (p->time_kernels) GB_HIP(w.ensure_events(3 * (size_t)p->max_iter));
    const std::vector<hipEvent_t> &evs = w.events;
    for (int k = 0; k < p->max_iter; k++) {
      const float *xo = (k % 2 == 0) ? xa : xb;
      const float *co = (k % 2 == 0) ? ca : cb;
      float *xn = (k % 2 == 0) ? xb : xa;
      float *cn = (k % 2 == 0) ? cb : ca;
      // the stop flag is read after iterations that end a chunk: their end
      // runs on its own, every other iteration's in the next scatter launch
      const bool read = (k + 1) % CH == 0 || k + 1 == p->max_iter;
      one_iter(xo, co, xn, cn, p->time_kernels ? evs[3 * k] : nullptr,
               p->time_kernels ? evs[3 * k + 1] : nullptr, can_fold && k % CH != 0,
               can_fold && !read);
      if (p->time_kernels && iter_ms) (void)hipEventRecord(evs[3 * k + 2], st);
      if (read) {
        GB_HIP(hipMemcpyAsync(&host_state, state, sizeof(PrState), hipMemcpyDeviceToHost, st));
        GB_HIP(hipStreamSynchronize(st));
        if (host_state.done) break;
      }
    }
    if (p->time_kernels) {
      GB_HIP(hipStreamSynchronize(st));
      // launches after the stop flag exit at once: count only working ones
      for (int k = 0; k < host_state.iter && k < p->max_iter; k++) {
        float ms = 0.0f, ims = 0.0f;
        GB_HIP(hipEventElapsedTime(&ms, evs[3 * k], evs[3 * k + 1]));
        gather_ms += ms;
        if (iter_ms) {
          GB_HIP(hipEventElapsedTime(&ims, evs[3 * k], evs[3 * k + 2]));
          *iter_ms += ims;
        }
        gather_launches++;
      }
    }
  }
  GB_HIP(hipGetLastError());
  GB_HIP(hipMemcpyAsync(&host_state, state, sizeof(PrState), hipMemcpyDeviceToHost, st));
  GB_HIP(hipStreamSynchronize(st));
  const float *xf = host_state.iter % 2 == 1 ? xb : xa;
  if (relabel)
    hipLaunchKernelGGL(pr_unpermute_kernel, dim3((unsigned)vblocks), dim3(PR_BLOCK), 0, st, V,
                       g->rl_rank.as<int32_t>(), xf, d_x);
  else
    GB_HIP(hipMemcpyAsync(d_x, xf, sizeof(float) * V, hipMemcpyDeviceToDevice, st));
  GB_HIP(hipGetLastError());
  GB_HIP(hipStreamSynchronize(st));
  const bool conv = host_state.residual < p->tol;
  if (res) {
    res->iterations = host_state.iter;
    res->converged = conv ? 1 : 0;
    res->residual = host_state.residual;
    res->dangling = host_state.dangling;
    res->gather_ms = gather_ms;
    res->gather_launches = gather_launches;
    res->tiles = (int32_t)ntiles;
    res->items_per_thread = rb || pb ? 0 : ipt;
    res->hot = hot;
    res->relabeled = relabel ? 1 : 0;
    res->row_block = rb;
    res->binned = pb;
  }
  return GB_OK;
}

gb_status gb_pagerank_ref_core(int32_t V, const int32_t *d_irp, const int32_t *d_icol,
                               const double *d_recip, int32_t max_iter, double *d_x,
                               int32_t *iterations, hipStream_t st, double *d_diff, double *d_wb,
                               int *d_flag, gb_pr_ref_hook hook, void *ctx) {
  const int blocks = (V + 255) / 256;
  hipLaunchKernelGGL(pr_ref_init_kernel, dim3(blocks), dim3(256), 0, st, V, d_recip, d_x, d_diff,
                     d_wb);
  const double pj = 1.0 - 0.3;  // purpose_jump, pagerank_mpi.c:689
  int32_t it = 0;
  for (;;) {
    if (max_iter > 0 && it >= max_iter) {
      it = -1;
      break;
    }
    it++;
    if (hook) {
      GB_HIP(hipStreamSynchronize(st));
      hook(it, d_wb, ctx);  // pagerank_mpi.c:823-830 prints wb here
    }
    // a kernel, not a memset, clears the flag
    hipLaunchKernelGGL(pr_ref_flag_clear_kernel, dim3(1), dim3(1), 0, st, d_flag);
    hipLaunchKernelGGL(pr_ref_gather_kernel, dim3(blocks), dim3(256), 0, st, V, d_irp, d_icol,
                       (const double *)d_diff, d_wb, pj);
    hipLaunchKernelGGL(pr_ref_update_kernel, dim3(blocks), dim3(256), 0, st, V, d_recip, d_x,
                       d_diff, (const double *)d_wb, d_flag);
    int h = 0;
    GB_HIP(hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    GB_HIP(hipStreamSynchronize(st));
    if (!h) break;
  }
  GB_HIP(hipGetLastError());
  if (iterations) *iterations = it;
  return it < 0 ? GB_ERR_NOT_CONVERGED : GB_OK;
}

gb_status gb_pagerank_ref_impl(gb_graph g, const double *d_recip, int32_t max_iter, double *d_x,
                               int32_t *iterations, hipStream_t st, gb_pr_ref_hook hook,
                               void *ctx) {
  if (!g || !d_recip || !d_x) return gb_fail(GB_ERR_INVALID, "gb_pagerank_ref: null argument");
  if (!g->rrp || (!g->rcol && g->E > 0))
    return gb_fail(GB_ERR_INVALID, "gb_pagerank_ref: no reverse CSR");
  const int32_t V = g->V;
  if (V <= 0) return gb_fail(GB_ERR_INVALID, "gb_pagerank_ref: empty graph");
  DevBuf diff, wb, flag;
  GB_TRY(diff.ensure(sizeof(double) * V));
  GB_TRY(wb.ensure(sizeof(double) * V));
  GB_TRY(flag.ensure(sizeof(int)));
  return gb_pagerank_ref_core(V, g->rrp, g->rcol, d_recip, max_iter, d_x, iterations, st,
                              diff.as<double>(), wb.as<double>(), flag.as<int>(), hook, ctx);
}

extern "C" gb_status gb_pagerank_ref(gb_graph g, const double *d_recip, int32_t max_iter,
                                     double *d_x, int32_t *iterations, void *stream) {
  return gb_pagerank_ref_impl(g, d_recip, max_iter, d_x, iterations, gb_stream(stream), nullptr,
                              nullptr);
}
