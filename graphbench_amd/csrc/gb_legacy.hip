// gb_legacy.hip -- the reference's own C entry points over the GPU library.
//
// run_sspath_arr / run_sspath / run_pagerank_csr_mpi / run_bc and the
// readers keep the reference's signatures, t_csr layout, file formats,
// printed output and exit(123) error handling (include/gb_legacy.h).  The
// t_csr entry points convert t_csr (int64 end offsets, AoS vertex/edge info)
// into the device layout once, call the native gb_* path and write results
// back into the caller's t_csr; the file entry points (run_pagerank_csr_mpi,
// run_bc) parse the file into an edge stream and build their CSRs on the
// device.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gb_internal.hpp"
#include "gb_legacy.h"

static_assert(sizeof(t_vi) == 32, "t_vi layout (graph.h:95-102)");
static_assert(sizeof(t_e) == 16, "t_e layout (graph.h:105-110)");
static_assert(sizeof(t_csr) == 72, "t_csr layout (graph.h:113-125)");
static_assert(sizeof(t_list) == 32, "t_list layout (list.h:3-9)");

// ---- the reference's globals (main/main.c:8-22); weak so that a caller's
// own definitions win --------------------------------------------------------
extern "C" {
__attribute__((weak)) sint num_threads = 1;
__attribute__((weak)) sint num_nodes = 1;
__attribute__((weak)) lint g_v_num = 1;
__attribute__((weak)) char *par_extend = (char *)"parex/";
__attribute__((weak)) char *data_extend = (char *)"datex/";
__attribute__((weak)) char *idx_extend = (char *)"idxex/";
__attribute__((weak)) char *base_dir = (char *)"./";
__attribute__((weak)) char *nodes_extension = nullptr;
__attribute__((weak)) int bin = 0;
__attribute__((weak)) int buck_size = 10;
__attribute__((weak)) sint use_partition = 0;
__attribute__((weak)) int buff_size = 10000;
__attribute__((weak)) lint prefetch_dis = 10;
}

namespace {

// ERROR_PRINT (graph.h:74): message, then exit(123)
[[noreturn]] void fatal(const char *what) {
  printf("Error: %s\n", what);
  const char *e = gb_last_error();
  if (e && *e) printf("  %s\n", e);
  fflush(stdout);
  exit(123);
}

void check(gb_status s, const char *what) {
  if (s != GB_OK) fatal(what);
}

void hipcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    gb_fail(GB_ERR_HIP, "%s", hipGetErrorString(e));
    fatal(what);
  }
}

void *amalloc(size_t n) {  // _mm_malloc(n, 64) equivalent
  void *p = nullptr;
  if (posix_memalign(&p, 64, n ? n : 64) != 0) fatal("host allocation failed");
  return p;
}

// Buffered reader with the reference's fscanf semantics (csr.c:26-35,
// 156-176): next() is "%99s" (a whitespace-delimited token, split after 99
// characters), num() is "%lld" (optional sign and decimal digits, stopping
// at the first other character, which is left for the next read).  One
// 1 MiB read at a time instead of a libc call per field.
struct Tok {
  FILE *f;
  std::vector<char> buf = std::vector<char>(1 << 20);
  size_t pos = 0, len = 0;
  int peek() {
    if (pos == len) {
      len = fread(buf.data(), 1, buf.size(), f);
      pos = 0;
      if (len == 0) return EOF;
    }
    return (unsigned char)buf[pos];
  }
  void skip_space() {
    int c;
    while ((c = peek()) != EOF && isspace(c)) pos++;
  }
  bool next(char *out, size_t n) {
    skip_space();
    size_t k = 0;
    int c;
    while (k < 99 && k + 1 < n && (c = peek()) != EOF && !isspace(c)) {
      out[k++] = (char)c;
      pos++;
    }
    out[k] = '\0';
    return k > 0;
  }
  bool num(long long *v) {
    skip_space();
    int c = peek();
    bool neg = false;
    if (c == '-' || c == '+') {
      neg = c == '-';
      pos++;
      c = peek();
    }
    if (c == EOF || !isdigit(c)) return false;
    unsigned long long x = 0;
    int digits = 0;
    while ((c = peek()) != EOF && isdigit(c)) {
      if (++digits > 18) fatal("integer field out of range");
      x = x * 10 + (unsigned)(c - '0');
      pos++;
    }
    *v = neg ? -(long long)x : (long long)x;
    return true;
  }
};

FILE *open_in(const char *dir, const char *file, const char *mode) {
  char buf[4096];
  snprintf(buf, sizeof(buf), "%s%s", dir, file);
  FILE *f = fopen(buf, mode);
  if (!f) {
    printf("the file %s you input does not exist!\n", buf);
    fatal("cannot open graph file");
  }
  return f;
}

// Device copy of a t_csr (forward and optional reverse), f64 and u32 weights.
struct DevGraph {
  gb_graph g = nullptr;
  DevBuf rp, col, w, w64, rrp, rcol, rw64;
  int32_t V = 0;
  int64_t E = 0;
  bool integral = false;
  ~DevGraph() {
    if (g) gb_graph_destroy(g);
  }
};

template <class T>
void upload(DevBuf &b, const std::vector<T> &v) {
  check(b.ensure(sizeof(T) * (v.size() + 1)), "device allocation");
  if (!v.empty())
    hipcheck(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "upload");
}

// t_csr -> row_ptr/col (int32), weights f64 (and u32 when all are integers in
// [0, 2^32)); optional reverse CSR with weights, transposed on the device.
void to_device(const t_csr *gs, DevGraph &dg, bool reverse, bool weights) {
  if (!gs || gs->v_size <= 0 || gs->v_size >= (1ll << 31) || gs->e_size < 0 ||
      gs->e_size >= (1ll << 31))
    fatal("graph size outside the int32 device layout");
  const int32_t V = (int32_t)gs->v_size;
  const int64_t E = gs->e_size;
  std::vector<int32_t> rp(V + 1), col(E);
  rp[0] = 0;
  for (int32_t i = 0; i < V; i++) {
    lint end = gs->vet_idx[i];
    if (end < rp[i] || end > E) fatal("vet_idx is not a non-decreasing end-offset array");
    rp[i + 1] = (int32_t)end;
  }
  if (rp[V] != E) fatal("vet_idx[V-1] != e_size");
  for (int64_t e = 0; e < E; e++) {
    lint c = gs->edge_idx[e];
    if (c < 0 || c >= V) fatal("edge_idx out of range");
    col[e] = (int32_t)c;
  }
  std::vector<double> w64;
  std::vector<uint32_t> w32;
  bool integral = true;
  if (weights) {
    w64.resize(E);
    for (int64_t e = 0; e < E; e++) {
      double x = gs->edge_info[e].edge_weight;
      if (!(x >= 0.0) || std::isinf(x)) fatal("edge weights must be finite and non-negative");
      w64[e] = x;
      if (integral && !(x == std::floor(x) && x < 4294967296.0)) integral = false;
    }
    if (integral) {
      w32.resize(E);
      for (int64_t e = 0; e < E; e++) w32[e] = (uint32_t)w64[e];
    }
  }
  dg.V = V;
  dg.E = E;
  dg.integral = weights && integral;
  upload(dg.rp, rp);
  upload(dg.col, col);
  if (dg.integral) upload(dg.w, w32);
  if (weights) upload(dg.w64, w64);
  if (reverse) {  // built on the device (gb_csr_transpose)
    check(dg.rrp.ensure(sizeof(int32_t) * (V + 1)), "device allocation");
    check(dg.rcol.ensure(sizeof(int32_t) * (E + 1)), "device allocation");
    if (weights) check(dg.rw64.ensure(sizeof(double) * (E + 1)), "device allocation");
    check(gb_csr_transpose(V, E, dg.rp.as<int32_t>(), dg.col.as<int32_t>(),
                           weights ? (const void *)dg.w64.as<double>() : nullptr, weights ? 8 : 0,
                           dg.rrp.as<int32_t>(), dg.rcol.as<int32_t>(),
                           weights ? (void *)dg.rw64.as<double>() : nullptr, nullptr),
          "reverse CSR");
  }
  check(gb_graph_wrap(&dg.g, V, E, dg.rp.as<int32_t>(), dg.col.as<int32_t>(),
                      dg.integral ? dg.w.as<uint32_t>() : nullptr,
                      reverse ? dg.rrp.as<int32_t>() : nullptr,
                      reverse ? dg.rcol.as<int32_t>() : nullptr, nullptr),
        "graph handle");
  if (weights)
    check(gb_graph_set_f64_weights(dg.g, dg.w64.as<double>(),
                                   reverse ? dg.rw64.as<double>() : nullptr),
          "f64 weights");
}

void sssp_into(t_csr *gs, lint source, double delta, DevGraph &dg, std::vector<double> &dist) {
  if (source < 0 || source >= gs->v_size) fatal("source out of range");
  const int32_t src = (int32_t)source;
  const int32_t V = dg.V;
  dist.assign(V, INFINITY);
  if (dg.integral) {
    DevBuf d;
    check(d.ensure(sizeof(uint32_t) * V), "device allocation");
    gb_sssp_params p{};
    double fl = std::floor(delta);
    p.delta = (uint32_t)(fl >= 1.0 ? (fl < 4294967295.0 ? fl : 4294967295.0) : 1.0);
    p.light_max = p.delta;
    check(gb_sssp(dg.g, &p, &src, 1, d.as<uint32_t>(), nullptr, nullptr), "sssp");
    std::vector<uint32_t> h(V);
    hipcheck(hipMemcpy(h.data(), d.p, sizeof(uint32_t) * V, hipMemcpyDeviceToHost), "download");
    for (int32_t v = 0; v < V; v++) dist[v] = h[v] == GB_DIST_INF ? INFINITY : (double)h[v];
  } else {
    DevBuf d;
    check(d.ensure(sizeof(double) * V), "device allocation");
    check(gb_sssp_f64(dg.g, delta > 0 ? delta : 1.0, &src, 1, d.as<double>(), nullptr, nullptr),
          "sssp");
    hipcheck(hipMemcpy(dist.data(), d.p, sizeof(double) * V, hipMemcpyDeviceToHost), "download");
  }
}

// tight in-edges of the shortest-path DAG: d[u] + w == d[v] (csr_sspath.c:101)
__global__ void tight_edges_kernel(int32_t V, const int32_t *__restrict__ rp,
                                   const int32_t *__restrict__ col, const double *__restrict__ w,
                                   const double *__restrict__ dist, uint8_t *__restrict__ tight) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < V;
       u += (int64_t)gridDim.x * blockDim.x) {
    const double du = dist[u];
    for (int32_t e = rp[u]; e < rp[u + 1]; e++)
      tight[e] = (du != INFINITY && __dadd_rn(du, w[e]) == dist[col[e]]) ? 1 : 0;
  }
}

void add_list_one(p_list list, lint val) {  // list.c:11-26 (prepend)
  p_list elem = (p_list)malloc(sizeof(t_list));
  if (!elem) fatal("host allocation failed");
  elem->int_val = val;
  elem->int_val1 = -1;
  elem->double_val = -1;
  elem->next = list->next;
  list->next = elem;
}

void pr_print_hook(int32_t iter, const double *d_wb, void *ctx) {
  const std::vector<double> &recip = *(const std::vector<double> *)ctx;
  const int n = recip.size() < 10 ? (int)recip.size() : 10;
  double wb[10];
  if (hipMemcpy(wb, d_wb, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    fatal("download");
  printf("ITER %d ", iter);  // pagerank_mpi.c:823-830
  for (int i = 0; i < n; i++) printf("%f|%f ", wb[i], recip[i]);
  printf("\n");
  fflush(stdout);
}

// ---- graph files straight to a device CSR ------------------------------------
// run_pagerank_csr_mpi and run_bc read a graph file the way scan_csr_idx +
// read_graph_csr (text) or read_csr_bin (binary) do, but keep only the edge
// stream: the host parses the file once into per-edge (row, column) int32
// pairs in file order (text) or uploads the file's int64 pairs as they are
// (binary), and the CSR is built on the device by gb_csr_from_edges.  Its
// stable sort keeps each row's edges in file order, which is the order the
// reference's scatter (csr.c:45-56, 271-291) leaves them in.
enum { EW_NONE, EW_TEXT, EW_BIN };

struct FileGraph {
  lint v_num = 0, e_num = 0, offset = 0, g_v_size = 0;
  DevBuf row, col;  // int32 per edge in file order; -1 marks an endpoint out of range
  int weights = EW_NONE;  // which per-slot weight rule read_graph_csr / read_csr_bin use
};

// binary pairs (from, to) -> (row, column) for `type` (csr.c:250-291)
__global__ void bin_pairs_kernel(int64_t E, const int64_t *__restrict__ pairs, int64_t offset,
                                 int recv, int64_t V, int32_t *__restrict__ row,
                                 int32_t *__restrict__ col) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t from = pairs[2 * i], to = pairs[2 * i + 1];
    // OFFSET = 1 (graph.h:28-32): the other endpoint is 1-based
    const int64_t r = recv ? to - offset : from - offset;
    const int64_t c = recv ? from - 1 : to - 1;
    row[i] = (r >= 0 && r < V) ? (int32_t)r : -1;
    col[i] = (c >= 0 && c < V) ? (int32_t)c : -1;
  }
}

// Edge weights by CSR slot, as the readers assign them: slot i gets the i-th
// edge of the file's weight, v_to / v_num (csr.c:42, text), or 0.1 / 0.2 by
// the parity of i (csr.c:293-298, binary).  Text SEND files: v_to = col + 1.
__global__ void slot_weights_kernel(int64_t E, const int32_t *__restrict__ file_col, int text,
                                    double v_num, double *__restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = text ? (double)(file_col[i] + 1) / v_num : ((i & 1) ? 0.2 : 0.1);
}

int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

void load_text_edges(FileGraph &fg, const char *file, const char *dir, sint type) {
  FILE *f = open_in(dir, file, "r");
  Tok t{f};
  char str[128];
  long long v_num = 0, e_num = 0, offset = 0, g_v_size = 0, u, v, w;
  bool header = false;
  std::vector<int32_t> row, col;
  size_t read = 0;
  while (t.next(str, sizeof(str))) {
    if (str[0] == 'p' && str[1] == '\0') {
      char kind[128];
      if (!t.next(kind, sizeof(kind)) || !t.num(&v_num) || !t.num(&e_num) || !t.num(&offset) ||
          !t.num(&g_v_size))
        fatal("bad 'p' line");
      if (v_num <= 0 || v_num >= (1ll << 31) || e_num < 0 || e_num >= (1ll << 31))
        fatal("graph size outside the int32 device layout");
      row.resize(e_num);
      col.resize(e_num);
      read = 0;
      header = true;
    } else if (str[0] == 'a' && str[1] == '\0') {
      if (!header) fatal("'a' line before the 'p' line");
      if ((long long)read >= e_num) break;  // e_num == 0: no edge is read
      if (!t.num(&u) || !t.num(&v) || !t.num(&w)) fatal("bad 'a' line");
      // csr.c:37-48 / 165-176: RECV keys on the target, SEND on the source
      const long long r = type == GB_LEGACY_RECV ? v - offset : u - offset;
      const long long c = type == GB_LEGACY_RECV ? u - 1 : v - 1;
      if (r < 0 || r >= v_num) fatal("edge endpoint out of range");
      if (c < 0 || c >= v_num) fatal("edge endpoint out of range");
      row[read] = (int32_t)r;
      col[read] = (int32_t)c;
      if ((long long)++read == e_num) break;
    }
  }
  fclose(f);
  if (!header) fatal("no 'p' line");
  if ((long long)read != e_num) fatal("the file holds fewer edges than its 'p' line counts");
  fg.v_num = v_num;
  fg.e_num = e_num;
  fg.offset = offset;
  fg.g_v_size = g_v_size;
  fg.weights = EW_TEXT;
  upload(fg.row, row);
  upload(fg.col, col);
}

void load_bin_edges(FileGraph &fg, const char *file, const char *dir, sint type) {
  FILE *f = open_in(dir, file, "rb");
  lint hdr[4];
  if (fread(hdr, sizeof(lint), 4, f) != 4) fatal("short binary header");
  const lint v_num = hdr[0], e_num = hdr[1];
  if (v_num <= 0 || v_num >= (1ll << 31) || e_num < 0 || e_num >= (1ll << 31))
    fatal("graph size outside the int32 device layout");
  std::vector<lint> pairs(2 * (size_t)e_num);
  if ((lint)fread(pairs.data(), sizeof(lint), 2 * e_num, f) != 2 * e_num)
    fatal("short edge list");
  fclose(f);
  fg.v_num = v_num;
  fg.e_num = e_num;
  fg.offset = hdr[2];
  fg.g_v_size = hdr[3];
  fg.weights = EW_BIN;
  DevBuf raw;
  upload(raw, pairs);
  check(fg.row.ensure(sizeof(int32_t) * (e_num + 1)), "device allocation");
  check(fg.col.ensure(sizeof(int32_t) * (e_num + 1)), "device allocation");
  if (e_num > 0) {
    hipLaunchKernelGGL(bin_pairs_kernel, dim3(grid_for(e_num)), dim3(256), 0, 0, (int64_t)e_num,
                       raw.as<int64_t>(), (int64_t)fg.offset, type == GB_LEGACY_RECV ? 1 : 0,
                       (int64_t)v_num, fg.row.as<int32_t>(), fg.col.as<int32_t>());
    hipcheck(hipGetLastError(), "binary edges");
  }
  hipcheck(hipDeviceSynchronize(), "binary edges");
}

void load_edges(FileGraph &fg, const char *file, const char *dir, sint type) {
  if (bin)
    load_bin_edges(fg, file, dir, type);
  else
    load_text_edges(fg, file, dir, type);
}

// CSR rows keyed by fg.row (end-exclusive row_ptr of V + 1 entries).
void build_rows(const FileGraph &fg, DevBuf &rp, DevBuf &col) {
  const int32_t V = (int32_t)fg.v_num;
  const int64_t E = fg.e_num;
  check(rp.ensure(sizeof(int32_t) * (V + 1)), "device allocation");
  check(col.ensure(sizeof(int32_t) * (E + 1)), "device allocation");
  check(gb_csr_from_edges(V, E, fg.row.as<int32_t>(), fg.col.as<int32_t>(), nullptr, 0,
                          rp.as<int32_t>(), col.as<int32_t>(), nullptr, nullptr),
        "CSR build");
}

// Forward CSR with the readers' per-slot f64 weights and its transpose.
void build_send_graph(const FileGraph &fg, DevGraph &dg) {
  const int32_t V = (int32_t)fg.v_num;
  const int64_t E = fg.e_num;
  build_rows(fg, dg.rp, dg.col);
  check(dg.w64.ensure(sizeof(double) * (E + 1)), "device allocation");
  if (E > 0) {
    hipLaunchKernelGGL(slot_weights_kernel, dim3(grid_for(E)), dim3(256), 0, 0, E,
                       fg.col.as<int32_t>(), fg.weights == EW_TEXT ? 1 : 0, (double)fg.v_num,
                       dg.w64.as<double>());
    hipcheck(hipGetLastError(), "edge weights");
  }
  check(dg.rrp.ensure(sizeof(int32_t) * (V + 1)), "device allocation");
  check(dg.rcol.ensure(sizeof(int32_t) * (E + 1)), "device allocation");
  check(dg.rw64.ensure(sizeof(double) * (E + 1)), "device allocation");
  check(gb_csr_transpose(V, E, dg.rp.as<int32_t>(), dg.col.as<int32_t>(), dg.w64.as<double>(), 8,
                         dg.rrp.as<int32_t>(), dg.rcol.as<int32_t>(), dg.rw64.as<double>(),
                         nullptr),
        "reverse CSR");
  check(gb_graph_wrap(&dg.g, V, E, dg.rp.as<int32_t>(), dg.col.as<int32_t>(), nullptr,
                      dg.rrp.as<int32_t>(), dg.rcol.as<int32_t>(), nullptr),
        "graph handle");
  check(gb_graph_set_f64_weights(dg.g, dg.w64.as<double>(), dg.rw64.as<double>()), "f64 weights");
  dg.V = V;
  dg.E = E;
}

}  // namespace

extern "C" {

// ---- readers (graph/csr.c) --------------------------------------------------
// The library's own entry points call the *_impl functions (internal
// linkage), so a caller that also links the reference's graph/csr.o, whose
// exported readers then take precedence, does not change what they run.

static void scan_csr_idx_impl(t_csr *graph, char *file, char *dir, sint type) {
  FILE *f = open_in(dir, file, "r");
  char str[128];
  Tok t{f};
  long long v_num = 0, e_num = 0, offset = 0, g_v_size = 0, u, v, w;
  long long read = 0;
  bool header = false;
  while (t.next(str, sizeof(str))) {
    if (str[0] == 'p' && str[1] == '\0') {
      char kind[128];
      if (!t.next(kind, sizeof(kind)) || !t.num(&v_num) || !t.num(&e_num) || !t.num(&offset) ||
          !t.num(&g_v_size))
        fatal("bad 'p' line");
      graph->offset = offset;
      graph->v_size = v_num;
      graph->e_size = e_num;
      graph->global_v_size = g_v_size;
      graph->vet_idx = (lint *)amalloc(sizeof(lint) * v_num);
      graph->vet_info = (t_vi *)amalloc(sizeof(t_vi) * v_num);
      for (long long i = 0; i < v_num; i++) {
        graph->vet_info[i].weight = 1.0;
        graph->vet_info[i].wb = 0;
        graph->vet_info[i].recip = 0;
        graph->vet_info[i].vet_deg = 0;
      }
      header = true;
    } else if (str[0] == 'a' && str[1] == '\0') {
      if (!header) fatal("'a' line before the 'p' line");
      if (read >= e_num) break;  // e_num == 0: no edge is counted
      if (!t.num(&u) || !t.num(&v) || !t.num(&w)) fatal("bad 'a' line");
      if (type == GB_LEGACY_RECV) {
        u -= 1;
        v -= offset;
        if (v < 0 || v >= v_num) fatal("edge target out of range");
        graph->vet_info[v].vet_deg++;
      } else {
        u -= offset;
        v -= 1;
        if (u < 0 || u >= v_num) fatal("edge source out of range");
        graph->vet_info[u].vet_deg++;
      }
      if (++read == e_num) break;
    }
  }
  fclose(f);
  if (!header) fatal("no 'p' line");
  lint sum = 0;
  for (long long i = 0; i < v_num; i++) {  // csr.c:194-200
    sum += graph->vet_info[i].vet_deg;
    graph->vet_info[i].recip = 1.0 / (double)graph->vet_info[i].vet_deg;
    graph->vet_idx[i] = sum;
  }
}

static void read_graph_csr_impl(t_csr *graph, char *file, char *dir, sint type) {
  FILE *f = open_in(dir, file, "r");
  char str[128];
  Tok t{f};
  graph->edge_idx = (lint *)amalloc(sizeof(lint) * graph->e_size);
  graph->edge_info = (t_e *)amalloc(sizeof(t_e) * graph->e_size);
  long long v_num = 0, e_num = 0, offset = 0, g_v_size = 0, u, v, w;
  long long read = 0;
  std::vector<lint> pos;
  while (t.next(str, sizeof(str))) {
    if (str[0] == 'p' && str[1] == '\0') {
      char kind[128];
      if (!t.next(kind, sizeof(kind)) || !t.num(&v_num) || !t.num(&e_num) || !t.num(&offset) ||
          !t.num(&g_v_size))
        fatal("bad 'p' line");
      pos.assign(v_num, 0);
      for (long long i = 1; i < v_num; i++) pos[i] = graph->vet_idx[i - 1];
    } else if (str[0] == 'a' && str[1] == '\0') {
      if (pos.empty()) fatal("'a' line before the 'p' line");
      if (read >= graph->e_size) break;  // the edges scan_csr_idx counted
      if (!t.num(&u) || !t.num(&v) || !t.num(&w)) fatal("bad 'a' line");
      // csr.c:42: the weight of the read-th edge is v_to / v_num (file order)
      graph->edge_info[read].edge_weight = (double)v / (double)v_num;
      // csr.c:43-56: RECV keys on the target, SEND on the source
      const long long key = type == GB_LEGACY_RECV ? v - offset : u - offset;
      const long long other = type == GB_LEGACY_RECV ? u - 1 : v - 1;
      if (key < 0 || key >= v_num || key >= graph->v_size || pos[key] >= graph->e_size)
        fatal("edge does not match the scanned row counts");
      graph->edge_idx[pos[key]++] = other;
      if (++read == e_num) break;
    }
  }
  fclose(f);
}

static void read_csr_bin_impl(t_csr *graph, char *file, char *dir, sint type, int myrank) {
  (void)myrank;
  FILE *f = open_in(dir, file, "rb");
  lint hdr[4];
  if (fread(hdr, sizeof(lint), 4, f) != 4) fatal("short binary header");
  const lint v_num = hdr[0], e_num = hdr[1], offset = hdr[2];
  std::vector<lint> edges(2 * e_num);
  if ((lint)fread(edges.data(), sizeof(lint), 2 * e_num, f) != 2 * e_num) fatal("short edge list");
  fclose(f);
  const lint OFFSET = 1;  // graph.h:28-32 without USE_GRAPH500
  graph->offset = offset;
  graph->v_size = v_num;
  graph->e_size = e_num;
  graph->global_v_size = hdr[3];
  graph->vet_idx = (lint *)amalloc(sizeof(lint) * v_num);
  graph->vet_info = (t_vi *)amalloc(sizeof(t_vi) * v_num);
  graph->edge_idx = (lint *)amalloc(sizeof(lint) * e_num);
  graph->edge_info = (t_e *)amalloc(sizeof(t_e) * e_num);
  for (lint i = 0; i < v_num; i++) {
    graph->vet_info[i].weight = INFINITY;
    graph->vet_info[i].wb = 0;
    graph->vet_info[i].recip = 0;
    graph->vet_info[i].vet_deg = 0;
  }
  std::vector<lint> cnt(v_num, 0);
  for (lint i = 0; i < e_num; i++) {
    lint key = type == GB_LEGACY_RECV ? edges[2 * i + 1] - offset : edges[2 * i] - offset;
    if (key < 0 || key >= v_num) fatal("binary edge out of range");
    cnt[key]++;
  }
  lint sum = 0;
  for (lint i = 0; i < v_num; i++) {
    graph->vet_idx[i] = cnt[i] + sum;
    sum += cnt[i];
  }
  std::vector<lint> pos(v_num, 0);
  for (lint i = 1; i < v_num; i++) pos[i] = graph->vet_idx[i - 1];
  for (lint i = 0; i < e_num; i++) {
    lint from = edges[2 * i], to = edges[2 * i + 1];
    if (type == GB_LEGACY_RECV)
      graph->edge_idx[pos[to - offset]++] = from - OFFSET;
    else
      graph->edge_idx[pos[from - offset]++] = to - OFFSET;
  }
  for (lint i = 0; i < e_num; i++) graph->edge_info[i].edge_weight = (i % 2 == 0) ? 0.1 : 0.2;
}

void scan_csr_idx(t_csr *graph, char *file, char *dir, sint type) {
  scan_csr_idx_impl(graph, file, dir, type);
}
void read_graph_csr(t_csr *graph, char *file, char *dir, sint type) {
  read_graph_csr_impl(graph, file, dir, type);
}
void read_csr_bin(t_csr *graph, char *file, char *dir, sint type, int myrank) {
  read_csr_bin_impl(graph, file, dir, type, myrank);
}

// ---- SSSP -------------------------------------------------------------------

void run_sspath_arr(t_csr *gs, p_list *predecesor, lint source, double delta) {
  (void)predecesor;
  DevGraph dg;
  to_device(gs, dg, false, true);
  std::vector<double> dist;
  sssp_into(gs, source, delta, dg, dist);
  for (lint v = 0; v < gs->v_size; v++) gs->vet_info[v].weight = dist[v];
  g_v_num = gs->v_size;
}

void run_sspath(t_csr *gs, p_list *predecesor, lint source, double delta) {
  DevGraph dg;
  to_device(gs, dg, false, true);
  std::vector<double> dist;
  sssp_into(gs, source, delta, dg, dist);
  for (lint v = 0; v < gs->v_size; v++) gs->vet_info[v].weight = dist[v];
  if (!predecesor) return;
  const int32_t V = dg.V;
  const int64_t E = dg.E;
  DevBuf dd, tight;
  check(dd.ensure(sizeof(double) * V), "device allocation");
  check(tight.ensure(E + 1), "device allocation");
  hipcheck(hipMemcpy(dd.p, dist.data(), sizeof(double) * V, hipMemcpyHostToDevice), "upload");
  if (E > 0) {
    int blocks = (V + 255) / 256;
    hipLaunchKernelGGL(tight_edges_kernel, dim3(blocks), dim3(256), 0, 0, V, dg.rp.as<int32_t>(),
                       dg.col.as<int32_t>(), dg.w64.as<double>(), dd.as<double>(),
                       tight.as<uint8_t>());
    hipcheck(hipGetLastError(), "tight edges");
  }
  std::vector<uint8_t> h(E);
  if (E) hipcheck(hipMemcpy(h.data(), tight.p, E, hipMemcpyDeviceToHost), "download");
  lint e = 0;
  for (lint u = 0; u < gs->v_size; u++)
    for (; e < gs->vet_idx[u]; e++)
      if (h[e]) add_list_one(predecesor[gs->edge_idx[e]], u);
}

// ---- PageRank (pagerank_mpi.c:667-1044, one rank) ---------------------------

void run_pagerank_csr_mpi(char *file, char *dir) {
  if (num_nodes != 1)
    fatal("run_pagerank_csr_mpi: this build runs on one GPU (num_nodes must be 1)");
  const int myrank = 0;
  char f_buf[4096];
  // the in-edge file (pagerank_mpi.c:721-728): read_graph_csr RECV or
  // read_csr_bin RECV; the CSR is built on the device
  snprintf(f_buf, sizeof(f_buf), "%s%s_rc_%d_%d", data_extend, file, myrank, num_nodes);
  FileGraph fg;
  load_edges(fg, f_buf, dir, GB_LEGACY_RECV);
  const lint V = fg.v_num;
  // read_recip (csr_partition.c:166-185: V doubles, then distrib, text) or
  // read_recip_bin (csr_partition_bin_mpi.c:312-332: num_nodes lints of
  // distrib, then V doubles, binary)
  std::vector<double> recip(V);
  {
    if (!bin)
      snprintf(f_buf, sizeof(f_buf), "%s%s%s_recip_%d_%s", dir, par_extend, file, myrank,
               nodes_extension ? nodes_extension : "(null)");
    else
      snprintf(f_buf, sizeof(f_buf), "%s%s%s_recip_%d_%d", dir, par_extend, file, myrank,
               num_nodes);
    FILE *r = fopen(f_buf, bin ? "rb" : "r");
    if (!r) {
      printf("the file %s does not exists!\n", f_buf);
      fatal("missing recip file");
    }
    if (!bin) {
      for (lint i = 0; i < V; i++)
        if (fscanf(r, "%lf", &recip[i]) != 1) fatal("short recip file");
    } else {
      std::vector<lint> distrib(num_nodes);
      if ((lint)fread(distrib.data(), sizeof(lint), num_nodes, r) != (lint)num_nodes ||
          (lint)fread(recip.data(), sizeof(double), V, r) != V)
        fatal("short recip file");
    }
    fclose(r);
    if (!bin) printf("%1.10f\n", recip[0]);
  }
  // the RECV CSR is the in-CSR: wrap it as the reverse graph of the handle
  DevGraph dg;
  build_rows(fg, dg.rrp, dg.rcol);
  fg.row.release();
  fg.col.release();
  check(gb_graph_wrap(&dg.g, (int32_t)V, fg.e_num, nullptr, nullptr, nullptr,
                      dg.rrp.as<int32_t>(), dg.rcol.as<int32_t>(), nullptr),
        "graph handle");
  dg.V = (int32_t)V;
  dg.E = fg.e_num;
  DevBuf drecip, dx;
  upload(drecip, recip);
  check(dx.ensure(sizeof(double) * (V + 1)), "device allocation");
  int32_t iters = 0;
  hipEvent_t e0, e1;
  hipcheck(hipEventCreate(&e0), "event");
  hipcheck(hipEventCreate(&e1), "event");
  hipcheck(hipEventRecord(e0, 0), "event");
  gb_status s = gb_pagerank_ref_impl(dg.g, drecip.as<double>(), 0, dx.as<double>(), &iters, 0,
                                     pr_print_hook, &recip);
  hipcheck(hipEventRecord(e1, 0), "event");
  check(s, "pagerank");
  const int n = V < 10 ? (int)V : 10;
  double x[10];
  hipcheck(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
  float ms = 0;
  hipcheck(hipEventSynchronize(e1), "event");
  hipcheck(hipEventElapsedTime(&ms, e0, e1), "event");
  hipcheck(hipEventDestroy(e0), "event");
  hipcheck(hipEventDestroy(e1), "event");
  printf("Final pagerank values: \n");  // pagerank_mpi.c:1024-1030
  for (int ii = 0; ii < n; ii++) printf("%lld|%lf ", ii + fg.offset, x[ii]);
  printf("\n");
  printf("iterations taken: %d\n", iters);
  printf("gpu time: %f ms (%lld in-edges, %d iterations)\n", ms, (long long)fg.e_num, iters);
  fflush(stdout);
}

// ---- Brandes (csr_bc.c:16-68) --------------------------------------------------

void run_bc(char *file, char *dir) {
  // scan_csr_idx + read_graph_csr SEND, or read_csr_bin SEND (csr_bc.c:20-33)
  FileGraph fg;
  load_edges(fg, file, dir, GB_LEGACY_SEND);
  DevGraph dg;
  build_send_graph(fg, dg);
  fg.row.release();
  fg.col.release();
  const int32_t V = dg.V;
  std::vector<int32_t> sources(V);
  for (int32_t i = 0; i < V; i++) sources[i] = i;  // every vertex (csr_bc.c:40)
  DevBuf dbc;
  check(dbc.ensure(sizeof(double) * V), "device allocation");
  int64_t batch = ((int64_t)64 << 20) / (V > 0 ? V : 1);
  if (batch < 1) batch = 1;
  if (batch > 65535) batch = 65535;
  check(gb_bc_weighted(dg.g, sources.data(), V, (int32_t)batch, dbc.as<double>(), nullptr,
                       nullptr),
        "betweenness");
  std::vector<double> bc(V);
  hipcheck(hipMemcpy(bc.data(), dbc.p, sizeof(double) * V, hipMemcpyDeviceToHost), "download");
  printf("the final result for betweeness centrality is:\n");
  for (int32_t i = 0; i < V; i++) printf("%lld:%lf ", (lint)(i + 1), bc[i]);
  printf("\n");
  fflush(stdout);
}

// ---- light/heavy split and partitioning ------------------------------------------

void separate_csr(t_csr *graph, t_csr *gl, t_csr *gh, double delta) {
  DevGraph dg;
  to_device(graph, dg, false, true);
  check(gb_graph_split64(dg.g, delta, nullptr), "separate");
  const int32_t V = dg.V;
  const int64_t E = dg.E;
  std::vector<int32_t> col(E), le(V);
  std::vector<double> w(E);
  if (E) {
    hipcheck(hipMemcpy(col.data(), dg.g->split64_col.p, sizeof(int32_t) * E, hipMemcpyDeviceToHost),
             "download");
    hipcheck(hipMemcpy(w.data(), dg.g->split64_w.p, sizeof(double) * E, hipMemcpyDeviceToHost),
             "download");
  }
  hipcheck(hipMemcpy(le.data(), dg.g->split64_le.p, sizeof(int32_t) * V, hipMemcpyDeviceToHost),
           "download");
  lint nl = 0, nh = 0;
  for (int32_t v = 0; v < V; v++) {
    lint s = v ? graph->vet_idx[v - 1] : 0;
    nl += le[v] - s;
    nh += graph->vet_idx[v] - le[v];
  }
  for (t_csr *g : {gl, gh}) {
    g->v_size = V;
    g->vet_idx = (lint *)amalloc(sizeof(lint) * V);
    g->vet_info = (t_vi *)amalloc(sizeof(t_vi) * V);
    memset(g->vet_info, 0, sizeof(t_vi) * V);
  }
  gl->e_size = nl;
  gh->e_size = nh;
  gl->edge_idx = (lint *)amalloc(sizeof(lint) * (nl ? nl : 1));
  gh->edge_idx = (lint *)amalloc(sizeof(lint) * (nh ? nh : 1));
  gl->edge_info = (t_e *)amalloc(sizeof(t_e) * (nl ? nl : 1));
  gh->edge_info = (t_e *)amalloc(sizeof(t_e) * (nh ? nh : 1));
  lint pl = 0, ph = 0;
  for (int32_t v = 0; v < V; v++) {
    lint s = v ? graph->vet_idx[v - 1] : 0, end = graph->vet_idx[v];
    for (lint e = s; e < end; e++) {
      t_csr *g = e < le[v] ? gl : gh;
      lint &p = e < le[v] ? pl : ph;
      g->edge_idx[p] = col[e];
      g->edge_info[p].edge_weight = w[e];
      g->edge_info[p].node_id = 0;
      g->vet_info[v].vet_deg++;
      p++;
    }
    gl->vet_idx[v] = pl;
    gh->vet_idx[v] = ph;
  }
}

void comp_distrib_csr(t_csr *graph_send, t_csr *graph_recv, lint *distrib, lint *send_size,
                      lint *recv_size, sint type) {
  const int32_t V = (int32_t)graph_send->v_size;
  const int32_t p = num_nodes;
  if (p <= 0 || V <= 0) fatal("comp_distrib_csr: empty graph or num_nodes <= 0");
  auto prefix = [&](const t_csr *g, DevBuf &b) {
    // degrees from vet_deg (the reference's walk reads vet_info[i].vet_deg)
    std::vector<int32_t> rp(V + 1, 0);
    for (int32_t i = 0; i < V; i++) rp[i + 1] = rp[i] + (int32_t)g->vet_info[i].vet_deg;
    upload(b, rp);
  };
  DevBuf srp, rrp, dd, ds, dr;
  prefix(graph_send, srp);
  prefix(graph_recv, rrp);
  check(dd.ensure(sizeof(int64_t) * p), "device allocation");
  check(ds.ensure(sizeof(int64_t) * p), "device allocation");
  check(dr.ensure(sizeof(int64_t) * p), "device allocation");
  // the cut threshold is gr->e_size / num_nodes in both modes (csr_partition.c:9)
  check(gb_edge_balanced_parts_ex(V, type == GB_LEGACY_RECV ? rrp.as<int32_t>() : srp.as<int32_t>(),
                                  srp.as<int32_t>(), rrp.as<int32_t>(), p, dd.as<int64_t>(),
                                  ds.as<int64_t>(), dr.as<int64_t>(), nullptr),
        "comp_distrib_csr");
  hipcheck(hipMemcpy(distrib, dd.p, sizeof(int64_t) * p, hipMemcpyDeviceToHost), "download");
  hipcheck(hipMemcpy(send_size, ds.p, sizeof(int64_t) * p, hipMemcpyDeviceToHost), "download");
  hipcheck(hipMemcpy(recv_size, dr.p, sizeof(int64_t) * p, hipMemcpyDeviceToHost), "download");
}

void init_node_id(t_csr *graph, lint *distrib, sint nparts) {
  const int64_t E = graph->e_size;
  if (E == 0) return;
  if (nparts <= 0) fatal("init_node_id: no parts");
  std::vector<int32_t> col(E);
  for (int64_t e = 0; e < E; e++) col[e] = (int32_t)graph->edge_idx[e];
  DevBuf dc, dd, dout;
  upload(dc, col);
  std::vector<int64_t> dist(distrib, distrib + nparts);
  upload(dd, dist);
  check(dout.ensure(sizeof(int64_t) * E), "device allocation");
  check(gb_node_ids(E, dc.as<int32_t>(), dd.as<int64_t>(), nparts, dout.as<int64_t>(), nullptr),
        "init_node_id");
  std::vector<int64_t> out(E);
  hipcheck(hipMemcpy(out.data(), dout.p, sizeof(int64_t) * E, hipMemcpyDeviceToHost), "download");
  for (int64_t e = 0; e < E; e++) graph->edge_info[e].node_id = out[e];
}

void write_recip(char *file, char *dir, lint *distrib, t_csr *graph_send) {
  // partition/csr_partition.c:133-164: part j's file
  // <dir><par_extend><file>_recip_<j>_<nodes_extension> holds "%f " of the
  // recip of its vertices (the walk moves to the next part after vertex
  // distrib[j]), then "\n", every distrib entry as "%lld ", "\n", and a
  // closing "\n"
  const int p = num_nodes;
  if (p <= 0) fatal("write_recip: num_nodes <= 0");
  if (!nodes_extension) fatal("write_recip: nodes_extension is not set");
  std::vector<FILE *> w(p, nullptr);
  char buf[4096];
  for (int i = 0; i < p; i++) {
    snprintf(buf, sizeof(buf), "%s%s%s_recip_%d_%s", dir, par_extend, file, i, nodes_extension);
    if ((w[i] = fopen(buf, "w")) == nullptr) {
      printf("the file does not exists!\n");
      fatal("write_recip: cannot open an output file");
    }
  }
  lint j = 0;
  for (lint i = 0; i < graph_send->v_size; i++) {
    if (j >= p) fatal("write_recip: vertex past the last part (distrib[p-1] < v_size-1)");
    fprintf(w[j], "%f ", graph_send->vet_info[i].recip);
    if (i == distrib[j]) j++;
  }
  for (int i = 0; i < p; i++) {
    fprintf(w[i], "\n");
    for (int k = 0; k < p; k++) fprintf(w[i], "%lld ", (long long)distrib[k]);
    fprintf(w[i], "\n");
  }
  for (int i = 0; i < p; i++) {
    fprintf(w[i], "\n");
    fclose(w[i]);
  }
}

void chunk_files(char *file, char *dir, lint *distrib, lint *send_size, lint *recv_size) {
  // partition/csr_partition.c:32-125: every edge line "a u v w" of
  // <dir><file> goes, in file order, to the send file of its source's part
  // (<dir><data_extend><file>_sd_<i>_<nodes_extension>) and the recv file of
  // its target's part (_rc_), each headed by "p sp size edges off V" at every
  // 'p' line; the counts of edges whose two parts differ go to
  // <dir><par_extend><file>_send_<nodes_extension> as a p x p matrix.  The
  // owners and counts are computed on the device (gb_edge_parts), batch by
  // batch between 'p' lines; the text is written here.
  const int p = num_nodes;
  if (p <= 0 || p > GB_MAX_PARTS) fatal("chunk_files: num_nodes outside [1, 4096]");
  if (!nodes_extension) fatal("chunk_files: nodes_extension is not set");
  char buf[4096];
  snprintf(buf, sizeof(buf), "%s%s%s_send_%s", dir, par_extend, file, nodes_extension);
  FILE *send_f = fopen(buf, "w");
  if (!send_f) {
    printf("the file %s does not exists!\n", buf);
    fatal("chunk_files: cannot open the send file");
  }
  std::vector<FILE *> ws(p, nullptr), wr(p, nullptr);
  for (int i = 0; i < p; i++) {
    snprintf(buf, sizeof(buf), "%s%s%s_sd_%d_%s", dir, data_extend, file, i, nodes_extension);
    if ((ws[i] = fopen(buf, "w")) == nullptr) {
      printf("the file you input does not exist!");
      fatal("chunk_files: cannot open a part file");
    }
    snprintf(buf, sizeof(buf), "%s%s%s_rc_%d_%s", dir, data_extend, file, i, nodes_extension);
    if ((wr[i] = fopen(buf, "w")) == nullptr) {
      printf("the file you input does not exist!");
      fatal("chunk_files: cannot open a part file");
    }
  }
  snprintf(buf, sizeof(buf), "%s%s", dir, file);
  FILE *in = fopen(buf, "r");
  if (!in) {
    printf("the file you input does not exist!");
    fatal("chunk_files: cannot open the graph file");
  }
  std::vector<int64_t> dist(distrib, distrib + p);
  DevBuf d_dist, d_u, d_v, d_sp, d_rp, d_ub;
  upload(d_dist, dist);
  check(d_ub.ensure(sizeof(unsigned long long) * (size_t)p * p), "device allocation");
  std::vector<unsigned long long> send_ub((size_t)p * p, 0ull), ub((size_t)p * p);
  // edges read since the last flush: as written (u, v, w) and the endpoints
  // the parts are found for (u - 1, v - offset); flushed at every 'p' line,
  // every CHUNK_EDGES edges and at the end
  size_t CHUNK_EDGES = (size_t)1 << 24;
  if (const char *e = getenv("GB_CHUNK_EDGES"))  // smaller batches (tests)
    if (atoll(e) > 0) CHUNK_EDGES = (size_t)atoll(e);
  std::vector<long long> eu, ev, ew;
  std::vector<int32_t> su, sv;
  auto flush = [&]() {
    const int64_t n = (int64_t)eu.size();
    if (n == 0) return;
    upload(d_u, su);
    upload(d_v, sv);
    check(d_sp.ensure(sizeof(int32_t) * n), "device allocation");
    check(d_rp.ensure(sizeof(int32_t) * n), "device allocation");
    check(gb_edge_parts(n, d_u.as<int32_t>(), d_v.as<int32_t>(), d_dist.as<int64_t>(), p,
                        d_sp.as<int32_t>(), d_rp.as<int32_t>(), d_ub.as<unsigned long long>(),
                        nullptr),
          "chunk_files");
    std::vector<int32_t> sp(n), rp(n);
    hipcheck(hipMemcpy(sp.data(), d_sp.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "download");
    hipcheck(hipMemcpy(rp.data(), d_rp.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "download");
    hipcheck(hipMemcpy(ub.data(), d_ub.p, sizeof(unsigned long long) * ub.size(),
                       hipMemcpyDeviceToHost),
             "download");
    for (size_t k = 0; k < ub.size(); k++) send_ub[k] += ub[k];
    for (int64_t e = 0; e < n; e++) {
      fprintf(ws[sp[e]], "a %lld %lld %lld\n", eu[e], ev[e], ew[e]);
      fprintf(wr[rp[e]], "a %lld %lld %lld\n", eu[e], ev[e], ew[e]);
    }
    eu.clear();
    ev.clear();
    ew.clear();
    su.clear();
    sv.clear();
  };
  Tok t{in};
  char str[128];
  long long v_num = 0, e_num = 0, offset = 0, read = 0;
  bool header = false;
  while (t.next(str, sizeof(str))) {
    if (str[0] == 'p' && str[1] == '\0') {
      char kind[128];
      // "%s %lld %lld %lld": the fifth header field is read as a skipped token
      if (!t.next(kind, sizeof(kind)) || !t.num(&v_num) || !t.num(&e_num) || !t.num(&offset))
        fatal("chunk_files: bad 'p' line");
      flush();
      for (int i = 0; i < p; i++) {
        const long long size = i == 0 ? distrib[0] + 1 : distrib[i] - distrib[i - 1];
        const long long off = (i == 0 ? 0 : distrib[i - 1] + 1) + 1;
        fprintf(ws[i], "p sp %lld %lld %lld %lld\n", size, (long long)send_size[i], off, v_num);
        fprintf(wr[i], "p sp %lld %lld %lld %lld\n", size, (long long)recv_size[i], off, v_num);
      }
      header = true;
    } else if (str[0] == 'a' && str[1] == '\0') {
      long long u, v, w;
      if (!header) fatal("chunk_files: 'a' line before the 'p' line");
      if (!t.num(&u) || !t.num(&v) || !t.num(&w)) fatal("chunk_files: bad 'a' line");
      const long long x = u - 1, y = v - offset;
      if (x < INT32_MIN || x > INT32_MAX || y < INT32_MIN || y > INT32_MAX)
        fatal("chunk_files: vertex id outside the int32 range");
      eu.push_back(u);
      ev.push_back(v);
      ew.push_back(w);
      su.push_back((int32_t)x);
      sv.push_back((int32_t)y);
      if (++read == e_num) break;
      if (eu.size() >= CHUNK_EDGES) flush();  // bounded host memory; order is kept
    }
  }
  flush();
  for (int i = 0; i < p; i++) {
    for (int k = 0; k < p; k++) fprintf(send_f, "%lld ", (long long)send_ub[(size_t)i * p + k]);
    fprintf(send_f, "\n");
  }
  for (int i = 0; i < p; i++) {
    fclose(ws[i]);
    fclose(wr[i]);
  }
  fclose(in);
  fclose(send_f);
}

void remove_duplicate(char *file, char *dir, sint myrank) {
  // partition/csr_partition_bin_mpi.c:31-106: <dir><file>_<rank> holds the
  // edge count (lint) and (from, to) lint pairs; the pairs without repeats,
  // first occurrences in file order, go to
  // <dir><data_extend><file>_phy_<rank>_<num_nodes>, headed by the edge count
  // (and, for one node, g_v_num before it and OFFSET 0, g_v_num after it)
  char path[4096];
  snprintf(path, sizeof(path), "%s%s_%d", dir, file, myrank);
  FILE *reader = fopen(path, "rb");
  if (!reader) {
    printf("THE FILE %s does not exist!\n", path);
    fatal("remove_duplicate: cannot open the edge file");
  }
  snprintf(path, sizeof(path), "%s%s%s_phy_%d_%d", dir, data_extend, file, myrank, num_nodes);
  FILE *writer = fopen(path, "wb");
  if (!writer) {
    printf("THE FILE %s does not exist!\n", path);
    fatal("remove_duplicate: cannot open the output file");
  }
  lint edge_num = 0;
  if (fread(&edge_num, sizeof(lint), 1, reader) != 1 || edge_num < 0)
    fatal("remove_duplicate: short edge file");
  std::vector<lint> buf((size_t)edge_num * 2);
  if (edge_num > 0 && fread(buf.data(), sizeof(lint), buf.size(), reader) != buf.size())
    fatal("remove_duplicate: short edge file");
  fclose(reader);
  if (edge_num >= ((lint)1 << 31)) fatal("remove_duplicate: more than 2^31 - 1 edges");
  std::vector<int32_t> src(edge_num), dst(edge_num);
  for (lint i = 0; i < edge_num; i++) {
    const lint from = buf[2 * i], to = buf[2 * i + 1];
    if (from == -1 || to == -1) {
      printf("Node: %d ::: i: %lld, graph500 generated graph has -1 in it", myrank, i);
      fatal("remove_duplicate: vertex id -1");
    }
    if (from < 0 || to < 0 || from >= ((lint)1 << 31) || to >= ((lint)1 << 31))
      fatal("remove_duplicate: vertex id outside [0, 2^31)");
    src[i] = (int32_t)from;
    dst[i] = (int32_t)to;
  }
  DevBuf ds, dd, os, od;
  upload(ds, src);
  upload(dd, dst);
  check(os.ensure(sizeof(int32_t) * ((size_t)edge_num + 1)), "device allocation");
  check(od.ensure(sizeof(int32_t) * ((size_t)edge_num + 1)), "device allocation");
  int64_t kept = 0;
  check(gb_edges_unique(edge_num, ds.as<int32_t>(), dd.as<int32_t>(), os.as<int32_t>(),
                        od.as<int32_t>(), &kept, nullptr),
        "remove_duplicate");
  std::vector<int32_t> ks(kept), kd(kept);
  if (kept > 0) {
    hipcheck(hipMemcpy(ks.data(), os.p, sizeof(int32_t) * kept, hipMemcpyDeviceToHost), "download");
    hipcheck(hipMemcpy(kd.data(), od.p, sizeof(int32_t) * kept, hipMemcpyDeviceToHost), "download");
  }
  const lint nondup = kept, off = 0;
  if (num_nodes == 1) fwrite(&g_v_num, sizeof(lint), 1, writer);
  fwrite(&nondup, sizeof(lint), 1, writer);
  if (num_nodes == 1) {
    fwrite(&off, sizeof(lint), 1, writer);
    fwrite(&g_v_num, sizeof(lint), 1, writer);
  }
  std::vector<lint> out((size_t)kept * 2);
  for (int64_t i = 0; i < kept; i++) {
    out[2 * i] = ks[i];
    out[2 * i + 1] = kd[i];
  }
  if (kept > 0) fwrite(out.data(), sizeof(lint), out.size(), writer);
  fclose(writer);
}

lint bin_search(lint *distrib, lint val, lint start, lint end) {
  // utils/utils.c:222-239
  for (;;) {
    lint mid = start + (end - start) / 2;
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

}  // extern "C"
