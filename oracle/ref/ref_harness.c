/*
 * ref_harness.c -- driver around the reference build (TEST INFRASTRUCTURE).
 *
 * Links against the reference's own objects (oracle/ref/Makefile) and runs
 * them on small inputs to produce golden vectors:
 *   cli <args...>                 the reference's main() (PR -k/-r/-i/-e ...)
 *   sssp_list V E src delta       run_sspath (sssp/csr_sspath.c:38-83) on an
 *                                 in-memory t_csr; edges "u v w" (0-based,
 *                                 grouped by u) on stdin; prints distances
 *                                 and predecessor lists
 *   distrib V p type              comp_distrib_csr (partition/csr_partition.c:5-30)
 *                                 on "send_deg recv_deg" lines, then
 *                                 bin_search (utils/utils.c:222-239) of 0..V-1
 *   separate V E delta            separate_csr (graph/csr.c:75-130)
 * Output is plain text, one value per token.  Needs ./data/log/ (created).
 */
#include <malloc.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "graph.h"
#include "list.h"

int gbref_reference_main(int argc, char *argv[]);
void comp_distrib_csr(t_csr *gs, t_csr *gr, lint *distrib, lint *send_size,
                      lint *recv_size, sint type);
lint bin_search(lint *distrib, lint val, lint start, lint end);
void separate_csr(t_csr *graph, t_csr *gl, t_csr *gh, double delta);
void creat_list(p_list list);

static t_csr *read_edges(lint V, lint E) {
  t_csr *g = (t_csr *)calloc(1, sizeof(t_csr));
  g->v_size = V;
  g->e_size = E;
  g->vet_idx = (lint *)calloc(V ? V : 1, sizeof(lint));
  g->edge_idx = (lint *)calloc(E ? E : 1, sizeof(lint));
  g->vet_info = (t_vi *)calloc(V ? V : 1, sizeof(t_vi));
  g->edge_info = (t_e *)calloc(E ? E : 1, sizeof(t_e));
  lint *deg = (lint *)calloc(V ? V : 1, sizeof(lint));
  for (lint e = 0; e < E; e++) {
    lint u, v;
    double w;
    if (scanf("%lld %lld %lf", &u, &v, &w) != 3) {
      fprintf(stderr, "bad edge line %lld\n", e);
      exit(2);
    }
    deg[u]++;
    g->edge_idx[e] = v;
    g->edge_info[e].edge_weight = w;
  }
  lint s = 0;
  for (lint i = 0; i < V; i++) {
    s += deg[i];
    g->vet_idx[i] = s;
    g->vet_info[i].vet_deg = deg[i];
  }
  free(deg);
  return g;
}

int main(int argc, char **argv) {
  /* every allocation from fresh (zeroed) pages: separate_csr reads
   * never-initialised degree counters (graph/csr.c:76-90) */
  mallopt(M_MMAP_THRESHOLD, 0);
  mkdir("data", 0755);
  mkdir("data/log", 0755);
  if (argc < 2) {
    fprintf(stderr, "usage: ref_harness cli|sssp_list|distrib|separate ...\n");
    return 2;
  }
  if (!strcmp(argv[1], "cli")) return gbref_reference_main(argc - 1, argv + 1);
  if (!strcmp(argv[1], "sssp_list") && argc == 6) {
    lint V = atoll(argv[2]), E = atoll(argv[3]), src = atoll(argv[4]);
    double delta = atof(argv[5]);
    t_csr *g = read_edges(V, E);
    p_list *pred = (p_list *)malloc(sizeof(p_list) * V);
    for (lint i = 0; i < V; i++) {
      pred[i] = (p_list)malloc(sizeof(t_list));
      creat_list(pred[i]);
    }
    run_sspath(g, pred, src, delta);
    printf("dist");
    for (lint i = 0; i < V; i++) printf(" %.17g", g->vet_info[i].weight);
    printf("\n");
    for (lint i = 0; i < V; i++) {
      printf("pred %lld", i);
      for (p_list p = pred[i]->next; p; p = p->next) printf(" %lld", p->int_val);
      printf("\n");
    }
    return 0;
  }
  if (!strcmp(argv[1], "distrib") && argc == 5) {
    lint V = atoll(argv[2]);
    num_nodes = atoi(argv[3]);
    sint type = atoi(argv[4]);
    t_csr gs = {0}, gr = {0};
    gs.v_size = gr.v_size = V;
    gs.vet_info = (t_vi *)calloc(V, sizeof(t_vi));
    gr.vet_info = (t_vi *)calloc(V, sizeof(t_vi));
    lint E = 0;
    for (lint i = 0; i < V; i++) {
      lint a, b;
      if (scanf("%lld %lld", &a, &b) != 2) return 2;
      gs.vet_info[i].vet_deg = a;
      gr.vet_info[i].vet_deg = b;
      E += a;
    }
    gs.e_size = gr.e_size = E;
    lint *distrib = (lint *)calloc(num_nodes, sizeof(lint));
    lint *ss = (lint *)calloc(num_nodes, sizeof(lint));
    lint *rs = (lint *)calloc(num_nodes, sizeof(lint));
    comp_distrib_csr(&gs, &gr, distrib, ss, rs, type);
    printf("distrib");
    for (int k = 0; k < num_nodes; k++) printf(" %lld", distrib[k]);
    printf("\nsend");
    for (int k = 0; k < num_nodes; k++) printf(" %lld", ss[k]);
    printf("\nrecv");
    for (int k = 0; k < num_nodes; k++) printf(" %lld", rs[k]);
    printf("\npart");
    for (lint v = 0; v < V; v++) printf(" %lld", bin_search(distrib, v, 0, num_nodes - 1));
    printf("\n");
    return 0;
  }
  if (!strcmp(argv[1], "separate") && argc == 5) {
    lint V = atoll(argv[2]), E = atoll(argv[3]);
    double delta = atof(argv[4]);
    t_csr *g = read_edges(V, E);
    t_csr gl = {0}, gh = {0};
    separate_csr(g, &gl, &gh, delta);
    t_csr *gg[2] = {&gl, &gh};
    for (int k = 0; k < 2; k++) {
      printf("%s_idx", k ? "heavy" : "light");
      for (lint i = 0; i < V; i++) printf(" %lld", gg[k]->vet_idx[i]);
      printf("\n%s_col", k ? "heavy" : "light");
      for (lint e = 0; e < gg[k]->e_size; e++) printf(" %lld", gg[k]->edge_idx[e]);
      printf("\n%s_w", k ? "heavy" : "light");
      for (lint e = 0; e < gg[k]->e_size; e++) printf(" %.17g", gg[k]->edge_info[e].edge_weight);
      printf("\n");
    }
    return 0;
  }
  fprintf(stderr, "bad command\n");
  return 2;
}
