/*
 * fsig_host.h - C ABI of libfsig_host.so: the host side of the FSIG drop-in
 * (graph file IO, CSR / PCSR / signature builders, inverted index, seeded
 * synthetic generators).  No GPU is touched.  Used by the FSIG driver's
 * tests and benchmarks through ctypes; C++ callers can use the classes of
 * fsig_amd/host directly.
 */
#ifndef FSIG_HOST_H
#define FSIG_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A host graph in CSR form: n vertices, off[n+1], adj[off[n]] ascending per
 * vertex without duplicates or self-loops, labels[n], num_labels = max+1. */
typedef struct fsig_hgraph fsig_hgraph;

/* Reads a `t / v / e` file (README.md:28-35) through the plain-graph reader
 * and the CSR transformer (IO/graph_reader.cc:12-55,
 * graph_format_transformer.cc:12-37).  *status gets 0 or a negative IoStatus. */
fsig_hgraph* fsig_hg_read(const char* path, int is_labeled, int* status);
fsig_hgraph* fsig_hg_from_csr(uint32_t n, const uint32_t* off, const uint32_t* adj, const int32_t* labels);
/* Edge list (pairs, any order, duplicates and self-loops dropped). */
fsig_hgraph* fsig_hg_from_edges(uint32_t n, uint64_t m, const uint32_t* pairs, const int32_t* labels);
fsig_hgraph* fsig_hg_chung_lu(uint32_t n, uint64_t m, double gamma, uint64_t seed);
fsig_hgraph* fsig_hg_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed);
/* kind 0 uniform, 1 Zipf(s) */
int fsig_hg_set_labels(fsig_hgraph* g, int kind, int32_t num_labels, double s, uint64_t seed);
/* 0, or -1 (cannot open), -6 (short write or failed close: the file is
 * incomplete). */
int fsig_hg_write(const fsig_hgraph* g, const char* path);
void fsig_hg_info(const fsig_hgraph* g, uint32_t* n, uint64_t* nnz, int32_t* num_labels, uint32_t* max_degree);
void fsig_hg_export(const fsig_hgraph* g, uint32_t* off, uint32_t* adj, int32_t* labels);
/* Zero-copy views, valid until fsig_hg_free. */
const uint32_t* fsig_hg_off(const fsig_hgraph* g);
const uint32_t* fsig_hg_adj(const fsig_hgraph* g);
const int32_t* fsig_hg_labels(const fsig_hgraph* g);
/* index[num_labels+1], list[n] (utility/inverted_index.h:16-44) */
void fsig_hg_inverted_index(const fsig_hgraph* g, uint32_t* index, uint32_t* list);
/* [16][n] when column_oriented, else [n][16] (graph_format_transformer.cc:217-281). */
void fsig_hg_signature(const fsig_hgraph* g, int column_oriented, int x86_shift, uint32_t* out);
void fsig_hg_free(fsig_hgraph* g);

/* Random-walk induced query: k vertices, edge count in [min_e, max_e].
 * Writes labels[k], edges[2*max_e] (u<v pairs), walk[k] (the data vertex of
 * each query vertex: one embedding).  Returns the edge count, or -1. */
int fsig_hg_walk_query(const fsig_hgraph* g, uint32_t k, uint32_t min_e, uint32_t max_e, uint64_t seed,
                       int bias, uint32_t max_attempts, int32_t* labels, uint32_t* edges, uint32_t* walk);

/* PCSR of a host graph (graph_format_transformer.cc:39-215).  NULL for a
 * graph with more than 4096 labels (one unit per label pair). */
typedef struct fsig_hpcsr fsig_hpcsr;
fsig_hpcsr* fsig_pcsr_build(const fsig_hgraph* g, int reference_unit_ids);
uint32_t fsig_pcsr_num_units(const fsig_hpcsr* p);
void fsig_pcsr_unit_info(const fsig_hpcsr* p, uint32_t u, uint32_t* keys, uint32_t* nedges);
const uint32_t* fsig_pcsr_unit_offset(const fsig_hpcsr* p, uint32_t u);
const uint32_t* fsig_pcsr_unit_edges(const fsig_hpcsr* p, uint32_t u);
void fsig_pcsr_free(fsig_hpcsr* p);

/* Binary caches, all fields 4 bytes in native order.  PCSR: the layout of
 * the reference's commented STORE / READ blocks
 * (graph_format_transformer/graph_format_transformer.cc:47-74, 197-213):
 * num_vertices, num_edges, num_labels, labels[num_vertices], num_units, then
 * per unit num_keys, num_edges, offset[32*num_keys], edges[num_edges].  A
 * loaded PCSR carries the header and labels (fsig_pcsr_graph_info /
 * fsig_pcsr_labels).  CSR: the same header and labels, then offset[n+1] and
 * the 2*num_edges adjacency entries.  Loads return NULL with *status one of
 * the reader's error codes on a missing, short, inconsistent or overlong
 * file.  Replace re-parsing the text format (the reference's cache blocks
 * are commented out). */
int fsig_hg_save_bin(const fsig_hgraph* g, const char* path);
fsig_hgraph* fsig_hg_load_bin(const char* path, int* status);
int fsig_pcsr_save_bin(const fsig_hpcsr* p, const fsig_hgraph* g, const char* path);
fsig_hpcsr* fsig_pcsr_load_bin(const char* path, int* status);
void fsig_pcsr_graph_info(const fsig_hpcsr* p, uint32_t* n, uint32_t* num_edges, int32_t* num_labels);
const int32_t* fsig_pcsr_labels(const fsig_hpcsr* p);

uint32_t fsig_murmur2(const void* key, int len, uint32_t seed);

/* Canonical embedding file: `#embedding: <count>` then the rows sorted
 * lexicographically, one per line.  0, -1 (cannot open) or -6 (write failed). */
int fsig_write_embeddings(const char* path, uint32_t k, uint64_t count, const uint32_t* rows, uint64_t nrows);

#ifdef __cplusplus
}
#endif
#endif /* FSIG_HOST_H */
