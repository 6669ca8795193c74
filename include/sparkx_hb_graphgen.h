/*
 * sparkx_hb_graphgen.h -- seeded synthetic graph generators (host, C++ threads).
 *
 * No reference counterpart: the reference reads graphs from MTX / GGCAT files
 * (src/graph/cache/mod.rs:588-642).  These build the synthetic workloads the
 * HyperBall engine is measured on.  Every generator returns a symmetric CSR
 * (u32 offsets[n+1], u32 neighbours[arcs]) with each adjacency list sorted
 * ascending, duplicates and self-loops removed.  Output depends only on the
 * arguments (not on the thread count).
 */
#ifndef SPARKX_HB_GRAPHGEN_H
#define SPARKX_HB_GRAPHGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hb_graph hb_graph;

/* Erdos-Renyi G(n, m): `samples` undirected edges with endpoints uniform in [0, n). */
int hb_graph_er(hb_graph **out, uint64_t n, uint64_t samples, uint64_t seed, int threads);

/* R-MAT: 2^scale nodes, `samples` edges, quadrant probabilities a, b, c (d = 1-a-b-c). */
int hb_graph_rmat(hb_graph **out, int scale, uint64_t samples, double a, double b, double c,
                  uint64_t seed, int threads);

/* k-mer de Bruijn graph: a uniform ACGT genome of `genome_len` bases, `copies`
 * copies with independent substitutions at rate `sub_rate`; nodes are the
 * distinct canonical k-mers (k <= 31) numbered in ascending 2-bit-code order;
 * an edge joins k-mers adjacent in any copy. */
int hb_graph_kmer(hb_graph **out, uint64_t genome_len, int copies, double sub_rate, int k,
                  uint64_t seed, int threads);

uint64_t hb_graph_nodes(const hb_graph *g);
uint64_t hb_graph_arcs(const hb_graph *g);
const uint32_t *hb_graph_offsets(const hb_graph *g);
const uint32_t *hb_graph_neighbours(const hb_graph *g);
void hb_graph_free(hb_graph *g);

#ifdef __cplusplus
}
#endif
#endif
