/*
 * gb_legacy.h -- drop-in C entry points with the reference's own signatures
 * and graph struct, implemented on the GPU by libgraphbench.so.
 *
 * The structs below are layout-identical to the reference's
 * (include/graph.h:95-127, include/list.h:3-9), so a caller built against the
 * reference's headers links against this library unchanged.  Do not include
 * this header together with the reference's graph.h in one translation unit.
 *
 * The library reads (never defines strongly) the reference's globals
 * num_threads, num_nodes, bin, buck_size, buff_size, prefetch_dis,
 * use_partition, g_v_num, par_extend, data_extend, idx_extend, base_dir,
 * nodes_extension (graph.h:79-91, defined in main/main.c:8-22).  Weak
 * definitions with main.c's defaults are provided for standalone use.
 *
 * Errors follow the reference: a message, then exit(123) (ERROR_PRINT,
 * graph.h:74).  GPU results replace the reference's CPU loops; see the
 * per-function notes for semantics that differ from the reference on purpose.
 */
#ifndef GB_LEGACY_H
#define GB_LEGACY_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int sint;
typedef long long int lint;

/* graph.h:95-102 */
struct vet_info {
  double weight; /* distance (SSSP) or rank (PageRank) */
  double wb;
  double recip;
  lint vet_deg;
};
typedef struct vet_info t_vi;

/* graph.h:105-110 */
struct edge_info {
  lint node_id;
  double edge_weight;
};
typedef struct edge_info t_e;

/* graph.h:113-125: vet_idx[i] is the END offset of row i (implicit 0 first) */
struct CSR {
  lint *vet_idx;
  lint *edge_idx;
  t_vi *vet_info;
  t_e *edge_info;
  lint v_size;
  lint e_size;
  lint offset;
  sint csr_type;
  lint global_v_size;
};
typedef struct CSR t_csr;

/* list.h:3-9 */
struct list_elem {
  lint int_val;
  lint int_val1;
  double double_val;
  struct list_elem *next;
};
typedef struct list_elem t_list;
typedef struct list_elem *p_list;

#define GB_LEGACY_RECV 0 /* graph.h:22 RECV */
#define GB_LEGACY_SEND 1 /* graph.h:23 SEND */

/* Replaces run_sspath_arr, sssp/csr_sspath_array.c:149-151 (called by
 * main.c:223).  Writes gs->vet_info[v].weight = shortest distance from source
 * (INFINITY if unreached) for every v; predecesor is ignored, as in the
 * reference.  Integer weights in [0, 2^32) run on uint32 distances, other
 * weights on f64; delta is the bucket width (integer path: max(1, floor)). */
void run_sspath_arr(t_csr *gs, p_list *predecesor, lint source, double delta);

/* Replaces run_sspath, sssp/csr_sspath.c:38-83 (list.h:31-33).  Distances as
 * above (f64 weights); predecesor[v] (list heads made by creat_list) receives
 * one add_list_one-style node per tight in-edge (u,v): d[u] + w == d[v].
 * The reference appends duplicates (re-relaxed tight edges); this records
 * each tight edge once. */
void run_sspath(t_csr *gs, p_list *predecesor, lint source, double delta);

/* Replaces run_pagerank_csr_mpi, pagerank/pagerank_mpi.c:667-1044
 * (pagerank_mpi.h:24) for num_nodes == 1: reads <dir><data_extend><file>_rc_0_1
 * (in-CSR, text or binary per `bin`) and the recip file
 * <dir><par_extend><file>_recip_0_<nodes_extension>, runs the reference loop
 * (x=1, wb=0.3, wb += 0.7*diff, stop when no |diff|>1e-6) on the GPU in FP64
 * with the reference's per-row order, and prints what the reference prints. */
void run_pagerank_csr_mpi(char *file, char *dir);

/* Replaces run_bc, btwcentral/csr_bc.c:16-68 (csr_bc.h:4-5): Brandes over all
 * V sources on the graph file's weights (text: v_to/V, binary: 0.1/0.2),
 * printed as "%lld:%lf " per vertex, sums not halved.  Each tight edge counts
 * once (the reference's duplicated predecessors are not reproduced). */
void run_bc(char *file, char *dir);

/* Readers, graph/csr.c:4-66, 133-202, 211-299 (csr.h:4-16): same file formats,
 * same fields filled, same error handling. */
void scan_csr_idx(t_csr *graph, char *file, char *dir, sint type);
void read_graph_csr(t_csr *graph, char *file, char *dir, sint type);
void read_csr_bin(t_csr *graph, char *file, char *dir, sint type, int myrank);

/* separate_csr, graph/csr.c:75-130: light (w <= delta) / heavy split of every
 * row, computed on the GPU; gl/gh receive newly allocated arrays. */
void separate_csr(t_csr *graph, t_csr *gl, t_csr *gh, double delta);

/* comp_distrib_csr, partition/csr_partition.c:5-30, on the GPU (num_nodes
 * parts; type SEND cuts on gs's degrees, RECV on gr's). */
void comp_distrib_csr(t_csr *graph_send, t_csr *graph_recv, lint *distrib,
                      lint *send_size, lint *recv_size, sint type);
/* init_node_id, partition/csr_partition.c:127-131: owner part of every edge
 * target (bin_search of utils.c:222-239), on the GPU. */
void init_node_id(t_csr *graph, lint *distrib, sint num_nodes);
/* Replace write_recip and chunk_files, partition/csr_partition.c:133-164 and
 * 32-125 (csr_partition.h), the file writers of the serial chunking step
 * (main.c:116-136: scan_csr_idx SEND and RECV, comp_distrib_csr, write_recip,
 * chunk_files).  write_recip writes
 * <dir><par_extend><file>_recip_<j>_<nodes_extension> per part j: "%f " of
 * each of its vertices' recip, then the distrib line.  chunk_files splits the
 * text graph <dir><file> into <dir><data_extend><file>_sd_<j>_<nodes_extension>
 * (edges by source part) and _rc_ (by target part) in file order, and writes
 * the cross-part edge counts to <dir><par_extend><file>_send_<nodes_extension>.
 * The edge owners and counts are computed on the GPU (gb_edge_parts).  Both
 * need num_nodes and nodes_extension set as main.c does for -n; the files are
 * byte-identical to the reference's.  Endpoint ids must fit int32.  The
 * edges go to the GPU in batches of 2^24 (GB_CHUNK_EDGES in the environment
 * sets another size), so host memory stays bounded. */
void write_recip(char *file, char *dir, lint *distrib, t_csr *graph_send);
void chunk_files(char *file, char *dir, lint *distrib, lint *send_size, lint *recv_size);
/* Replaces remove_duplicate, partition/csr_partition_bin_mpi.c:31-106
 * (csr_partition_bin_mpi.h): reads <dir><file>_<myrank> (lint edge count,
 * then (from, to) lint pairs), drops repeated pairs on the GPU (first
 * occurrence kept, file order kept) and writes
 * <dir><data_extend><file>_phy_<myrank>_<num_nodes> in the reference's
 * format (for num_nodes == 1: g_v_num, count, 0, g_v_num, pairs; otherwise
 * count, pairs).  Ids must lie in [0, 2^31); -1 stops as the reference does. */
void remove_duplicate(char *file, char *dir, sint myrank);
/* bin_search, utils/utils.c:222-239 (single lookup, host). */
lint bin_search(lint *distrib, lint val, lint start, lint end);

#ifdef __cplusplus
}
#endif
#endif
