/*
 * fsig.h - C ABI of libfsig_hip.so, the MI355X (gfx950) subgraph-matching
 * engine that replaces the GPU side of FSIG's matchers.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative FSIG_E* code; the message of the last failure is kept per context
 * (fsig_last_error).  Nothing in this library calls exit().
 *
 * Semantics of a match (the output contract of the reference's MMatch / GSI):
 * an embedding is a map phi from query vertices to data vertices that is
 * injective, keeps labels (all labels are 0 for an unlabeled graph) and maps
 * every query edge to a data edge.  Non-edges of the query are unconstrained
 * and automorphic images are distinct embeddings.  Rows handed back are k
 * data-vertex ids in query-vertex-id order.
 */
#ifndef FSIG_H
#define FSIG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsig_ctx fsig_ctx;
typedef struct fsig_graph fsig_graph;
typedef struct fsig_server fsig_server;

enum {
  FSIG_OK = 0,
  FSIG_EINVAL = -1,       /* bad argument, malformed graph or query        */
  FSIG_ENOMEM = -2,       /* request exceeds the HBM budget of the context */
  FSIG_EHIP = -3,         /* HIP runtime error                             */
  FSIG_EUNSUPPORTED = -4, /* e.g. more than 32 query vertices              */
  FSIG_ENOTFOUND = -5,    /* unknown ticket / missing PCSR                 */
  FSIG_EAGAIN = -6,       /* server: not finished yet / queue full         */
  FSIG_ELIMIT = -7        /* the context's work limit was exceeded         */
};

enum { FSIG_ALG_MMATCH = 1, FSIG_ALG_GSI = 2 };
enum { FSIG_MODE_COUNT = 0, FSIG_MODE_ENUMERATE = 1 };

#define FSIG_MAX_QUERY_VERTICES 32

/* Opens device `device`, creates the context's HIP stream and reserves an
 * arena of `hbm_budget_bytes` of device memory for per-query state
 * (candidate sets, candidate-edge lists, intermediate embedding tables).
 * 0 selects 3/4 of the device memory free at the call.
 * Replaces GPUSubgraphMatcher::init_GPU (match/subgraph_match.h:51-80) and
 * the MemoryManager budget of totalGlobalMem (match/MMatch.h:158-160,
 * memory_manager/memory_manager.h:16-28). */
int fsig_ctx_create(int device, uint64_t hbm_budget_bytes, fsig_ctx** out);
void fsig_ctx_destroy(fsig_ctx* ctx);
const char* fsig_last_error(const fsig_ctx* ctx);
/* Waits for all work on the context's device (timing brackets). */
int fsig_ctx_synchronize(fsig_ctx* ctx);

/* Uploads a data graph in CSR form; it stays resident until fsig_graph_free.
 * off[n+1], adj[off[n]] strictly ascending per vertex (no duplicates),
 * labels[n] in [0, num_labels), inv_index[num_labels+1] / inv_list[n] the
 * label -> vertex inverted index (vertices ascending within a label).
 * Replaces MMatch::copy_graph_to_GPU (match/MMatch.cu:14-32); the inverted
 * index is the one main() builds (main/run.cc:122,130). */
int fsig_graph_upload_csr(fsig_ctx* ctx, uint32_t n, const uint32_t* off, const uint32_t* adj,
                          const int32_t* labels, uint32_t num_labels, const uint32_t* inv_index,
                          const uint32_t* inv_list, fsig_graph** out);

/* Adds GSI's storage to an uploaded graph: num_units PCSR units (one per
 * unordered label pair; unit u has unit_keys[u] groups of 32 words in
 * unit_offset[u] and unit_nedges[u] neighbour ids in unit_edges[u], the
 * layout of graph/graph.h:90-95) and the column-major [16][n] signature
 * table of graph_format_transformer.cc:217-281.  reference_unit_ids != 0
 * says the units are numbered as graph/graph.h:107-122 numbers them.
 * Replaces the per-join unit uploads of GSI::join_embeddings
 * (match/GSI.cu:234-243) and the signature upload of GSI::signature_filter
 * (match/GSI.cu:38-45): everything is uploaded once and stays resident. */
int fsig_graph_upload_pcsr(fsig_ctx* ctx, fsig_graph* g, uint32_t num_units,
                           const uint32_t* unit_keys, const uint32_t* const* unit_offset,
                           const uint32_t* unit_nedges, const uint32_t* const* unit_edges,
                           const uint32_t* sig_colmajor, int reference_unit_ids);
void fsig_graph_free(fsig_ctx* ctx, fsig_graph* g);

/* Matches one query (k <= 32 vertices; qoff[k+1] / qadj CSR, each
 * undirected edge stored in both directions; qsig[k*16] row-major query
 * signatures, used by alg 2 only, may be NULL for alg 1).
 *   alg  FSIG_ALG_MMATCH: label/degree filter over the inverted index,
 *        count-based refinement, candidate-edge CSR, memory-blocked join.
 *        Replaces MMatch::match (match/MMatch.cu:2077-2122).
 *        FSIG_ALG_GSI: signature filter, join over the resident PCSR.
 *        Replaces GSI::match (match/GSI.cu:329-358).
 *   mode FSIG_MODE_COUNT: *count_out = number of embeddings (u64).
 *        FSIG_MODE_ENUMERATE: embeddings are materialised in HBM; the first
 *        min(count, emb_cap) rows are copied to emb_out (host, k words per
 *        row, query-vertex order) when emb_out != NULL.
 *   max_embeddings (0 = no bound) bounds both modes: the work stops once
 *        that many embeddings are found and *count_out = min(count,
 *        max_embeddings); pass L+1 to ask "more than L?".  Which embeddings
 *        form a bounded prefix is unspecified.
 */
int fsig_match(fsig_ctx* ctx, const fsig_graph* g, uint32_t k, const int32_t* qlabels,
               const uint32_t* qoff, const uint32_t* qadj, const uint32_t* qsig, int alg, int mode,
               uint64_t max_embeddings, uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap);

/* Enumerates every embedding of one query (as fsig_match with
 * FSIG_MODE_ENUMERATE, windowed through the arena, nothing copied to the
 * host) and folds each into an order-independent digest of the set: a row
 * r = (r_0 .. r_{k-1}) of data vertex ids in query-vertex order hashes to
 *   x = fold_{i=0..k-1} x <- mix64(x + r_i + (i << 32)),  x_0 = 0x9e3779b97f4a7c15 * (k + 1)
 * with mix64 the splitmix64 finaliser, and the set's digest is
 *   digest_out[0] = sum of x,  digest_out[1] = sum of mix64(x ^ 0xd6e8feb86659fd93)
 * modulo 2^64.  Two sets with equal digests and counts are equal up to a
 * 2^-64-scale collision; the digest does not depend on the order rows are
 * found in, so it compares engines and join orders.  *count_out is the
 * number of embeddings.  No counterpart in the reference, whose output is
 * the count and a few rows (match/MMatch.cu:2046-2074, match/GSI.cu:310-326). */
int fsig_match_digest(fsig_ctx* ctx, const fsig_graph* g, uint32_t k, const int32_t* qlabels,
                      const uint32_t* qoff, const uint32_t* qadj, const uint32_t* qsig, int alg,
                      uint64_t* count_out, uint64_t* digest_out);

/* fsig_match for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The match synchronises its stream between levels, so the call
 * cannot itself be captured into a HIP graph.  Same arguments, results
 * and errors as fsig_match, but every buffer of the match beyond its inputs
 * and outputs (candidate masks and arrays, candidate-edge lists, join
 * tables, scan partials, counters, the staging of emb_out) comes from one
 * static device array of FSIG_NOALLOC_SCRATCH_BYTES, which each GPU that
 * loads the library holds until the process ends; the call allocates no
 * device memory and records no profiling events.  The graph and the
 * context's stream and pinned host staging are the caller's, made before.
 * A graph too large for the scratch is rejected with FSIG_EINVAL (at least
 * 12 bytes a vertex plus 1 MiB must fit); a match that runs out of scratch
 * returns FSIG_ENOMEM.  All calls on one GPU share its scratch, which keeps
 * the last call's contents: calls that overlap on one GPU (any contexts or
 * threads) race and must be serialised by the caller. */
#define FSIG_NOALLOC_SCRATCH_BYTES (256ull << 20)
int fsig_match_noalloc(fsig_ctx* ctx, const fsig_graph* g, uint32_t k, const int32_t* qlabels,
                       const uint32_t* qoff, const uint32_t* qadj, const uint32_t* qsig, int alg, int mode,
                       uint64_t max_embeddings, uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap);

/* Candidate sets of the last fsig_match / fsig_match_digest of a context,
 * kept on the host while fsig_set_keep_candidates(ctx, 1) is in force (a
 * copy per match; off by default).  stage 0: after the filter (alg 1: label
 * and degree, plus the neighbour-label filter unless fsig_tune.nlf_filter
 * is 1, over the inverted index, as match/MMatch_kernels.cuh:6-16; alg 2:
 * the signature filter of match/GSI_kernels.cuh:8-24), stage 1: after
 * refinement (alg 1; match/MMatch_kernels.cuh:28-117, 190-214) or the same
 * sets (alg 2).  The set of query vertex u is ascending data vertex ids;
 * *size is its length and min(size, cap) ids are copied to out (may be
 * NULL).  FSIG_ENOTFOUND when no set of u is kept. */
int fsig_set_keep_candidates(fsig_ctx* ctx, int on);
int fsig_last_candidate_set(const fsig_ctx* ctx, uint32_t u, int stage, uint32_t* out, uint32_t cap,
                            uint32_t* size);

/* Counts the embeddings of nq queries (alg 1 semantics) against one
 * resident graph.  Filter (with the neighbour-label filter), refinement and
 * candidate edges run for the whole batch at once.  The join is either the
 * lockstep batched join (every query's level p in one count / scan / write
 * sequence, the last two levels of count queries fused) or one persistent
 * kernel, launched as a captured hipGraph, whose lane groups take
 * (query, root range) tasks from a device work queue and walk them
 * depth-first.  An autotuner times the engines and their lane counts and
 * LDS staging (lockstep shapes on the whole batch, the walk on a sample)
 * unless fsig_set_tune fixes them; the choice is cached per batch shape.
 * counts_out[i] is query i's count; *join_ms (may be NULL) the join's device
 * time.  No counterpart in the reference (it runs one query per process). */
int fsig_match_batch(fsig_ctx* ctx, const fsig_graph* g, uint32_t nq, const uint32_t* k,
                     const int32_t* const* qlabels, const uint32_t* const* qoff,
                     const uint32_t* const* qadj, uint64_t* counts_out, double* join_ms);
/* fsig_match_batch with a bound per query: max_embeddings[i] = 0 counts
 * query i exactly; otherwise counts_out[i] = min(count, max_embeddings[i])
 * (fsig_match's max_embeddings: pass L+1 to ask "more than L?") and its
 * search stops once that many embeddings are found.  A bounded query joins
 * the lockstep batch; when its next level would exceed the beam (2^16
 * children, fsig_tune.beam_log2) only its first parent rows, whose children
 * start within the beam, go on, and the count found from then on is a lower
 * bound that settles the query once it reaches the bound.  The cut rows (a
 * suffix of the query's rows at that level) stay in the batch's tables; a
 * beamed query that ends below its bound has their subtrees counted by
 * fsig_match's memory-blocked depth-first driver, starting from those rows
 * and stopping at what the bound leaves.  That driver honours the context's
 * work limit (fsig_set_work_limit): a
 * query over it is refused, status_out[i] = FSIG_ELIMIT and counts_out[i]
 * = 0, while the other queries complete (status_out NULL: the call then
 * returns FSIG_ELIMIT).  No autotuning (lockstep with lanes per query).  No
 * counterpart in the reference. */
int fsig_match_batch_bounded(fsig_ctx* ctx, const fsig_graph* g, uint32_t nq, const uint32_t* k,
                             const int32_t* const* qlabels, const uint32_t* const* qoff,
                             const uint32_t* const* qadj, const uint64_t* max_embeddings, uint64_t* counts_out,
                             int32_t* status_out, double* join_ms);
/* Configuration the autotuner chose for the last batch: engine 1 = lockstep
 * batched join (every query's level p in one count/scan/write sequence),
 * 2 = persistent work-queue walk; lanes per row or group; LDS tile (0:
 * none).  With fsig_set_tune(lanes_per_row != 0) the batch uses strategy 0:
 * lockstep, 1: persistent, 2: persistent with LDS staging. */
int fsig_last_batch_tune(const fsig_ctx* ctx, uint32_t* lanes_per_row, uint32_t* lds_tile, uint32_t* engine);

/* Per-kernel-family statistics of the context, accumulated since the last
 * fsig_stats_reset.  With profiling on, every launch of the family is
 * bracketed by HIP events on the context stream and its time is summed. */
enum {
  FSIG_K_FILTER = 0,   /* label/degree or signature filter + compaction  */
  FSIG_K_REFINE = 1,   /* count-based refinement passes                  */
  FSIG_K_CEDGE = 2,    /* candidate-edge CSR count + fill                */
  FSIG_K_JOIN_COUNT = 3,
  FSIG_K_JOIN_WRITE = 4,
  FSIG_K_SCAN = 5,
  FSIG_K_BATCH = 6,    /* batched join (lockstep or persistent)          */
  FSIG_K_JOIN_TAIL = 7,/* fused count of the last two levels            */
  FSIG_K_NUM = 8
};
int fsig_set_profiling(fsig_ctx* ctx, int on);
int fsig_stats_reset(fsig_ctx* ctx);
/* launches, summed device ms (profiling on), algorithmic bytes (every
 * logical element a kernel reads or writes counted once per use, at its
 * natural width) and units processed (rows, candidates or children; for
 * FSIG_K_CEDGE the adjacency entries of the candidate rows swept). */
int fsig_kernel_stats(const fsig_ctx* ctx, int family, uint64_t* launches, double* ms,
                      uint64_t* alg_bytes, uint64_t* units);
/* Candidate-set sizes of the last fsig_match, per query vertex (k entries). */
int fsig_last_candidates(const fsig_ctx* ctx, uint32_t k, uint32_t* sizes);
/* Peak arena bytes of the last fsig_match. */
uint64_t fsig_last_peak_bytes(const fsig_ctx* ctx);

/* Join tuning, all fields set at once; 0 selects the automatic choice.
 * lanes_per_row in {4,8,16,32,64}: lanes of a wavefront that share one
 * partial embedding (auto: the power of two at or above the mean list
 * length).  strategy (fsig_match_batch only; with lanes_per_row != 0): 0
 * lockstep join, 1 persistent walk, 2 persistent walk with the second list
 * staged in LDS (lds_tile words per group); 3 (any lanes_per_row): lockstep
 * join with each query's lanes from its mean candidate-list length, no
 * autotuning.  For fsig_match, lds_tile caps the LDS tile of the fused tail
 * (words per lane group).  refine_passes: cap on refinement passes (0:
 * 2). */
typedef struct fsig_tune {
  uint32_t lanes_per_row;
  uint32_t strategy;
  uint32_t lds_tile;
  uint32_t refine_passes; /* cap on refinement passes (0: default) */
  uint32_t fused_tail;    /* 0: count the last two levels of a count-mode
                             alg-1 match in one pass (the next-to-last level
                             is never written) when the last query vertex has
                             a backward neighbour and no work limit is set;
                             1: level by level */
  uint32_t nlf_filter;    /* 0: alg-1 filtering also requires, per label, at
                             least as many neighbours of that label as the
                             query vertex has (graphs with 2..255 labels);
                             1: label and degree only, as the reference */
  uint32_t beam_log2;     /* fsig_match_batch_bounded: beam of a bounded query
                             in the lockstep join, 2^beam_log2 children
                             (0: 16; 1..40) */
  uint32_t symmetry;      /* 0: a count-mode fsig_match counts one embedding
                             per orbit of the query's automorphism group
                             (symmetry-breaking conditions phi(a) < phi(b)
                             on the join, Grochow and Kellis 2007) and
                             multiplies by the group's order; 1: every
                             embedding.  Same result either way. */
} fsig_tune;
int fsig_set_tune(fsig_ctx* ctx, const fsig_tune* t);
int fsig_get_tune(const fsig_ctx* ctx, fsig_tune* t);

/* Work limit of fsig_match: when more than max_rows partial embeddings
 * (rows of all intermediate tables together) would be produced, the match
 * stops and returns FSIG_ELIMIT.  0 removes the limit.  Lets a caller reject
 * queries whose search space is too large before paying for it. */
int fsig_set_work_limit(fsig_ctx* ctx, uint64_t max_rows);

/* Achievable-bandwidth probe: a plain copy kernel (16-byte loads and stores)
 * of `bytes` from one arena buffer to another, run `iters` times on the
 * context stream; *gbps = 2*bytes*iters / elapsed.  Also the known-size
 * kernel the HBM byte counters are scaled by.  No counterpart in the
 * reference. */
int fsig_stream_copy(fsig_ctx* ctx, uint64_t bytes, uint32_t iters, double* gbps);

/* ---- Served request loop (continuous batching of count queries).
 *
 * `workers` worker contexts on `device` (each its own HIP stream and an arena
 * of hbm_budget_bytes / workers; 0 = 3/4 of the free device memory) share
 * the resident graph g.  A worker that is free takes the requests waiting at
 * that moment, oldest first, up to max_batch of them and while their
 * projected intermediate-table bytes (fsig_project_bytes) fit 3/4 of its
 * arena (the first request is always taken), and counts them as one batch
 * with fsig_match_batch's kernels.  Requests that arrive while batches run
 * wait only for the next free worker, and the batches of different workers
 * overlap on the device.  A batch that still runs out of arena falls back
 * to fsig_match per query, whose join is memory-blocked.  tune (NULL:
 * lockstep join with lanes per query, no autotuning) is set on every worker.
 * The graph may have at most 4096 labels (admission keeps an L x L table of
 * label-pair adjacency counts; fsig_graph_label_pairs has the same limit).
 * Replaces the reference's one query per process with a static device
 * budget (main/run.cc:101-140, memory_manager/memory_manager.h:220-242). */
typedef struct fsig_server_stats {
  uint64_t submitted, completed, batches, fallbacks, max_batch_seen;
  uint64_t refused;        /* requests over the workers' work limit (FSIG_ELIMIT) */
  uint32_t max_concurrent; /* most batches in flight at once */
  uint32_t pad;
  double busy_ms;          /* summed batch wall time over the workers */
  double wall_ms;          /* first submit to last completion (0 before one) */
} fsig_server_stats;
int fsig_server_create(int device, const fsig_graph* g, uint32_t workers, uint64_t hbm_budget_bytes,
                       uint32_t max_batch, const fsig_tune* tune, fsig_server** out);
/* Queues one count query (alg 1 semantics; the arrays are copied) and
 * returns its ticket.  The submit time starts the query's latency. */
int fsig_server_submit(fsig_server* s, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                       const uint32_t* qadj, uint64_t* ticket);
/* fsig_server_submit with a bound: the request's count is min(count,
 * max_embeddings) (0: exact), as fsig_match_batch_bounded counts it; its
 * admission projection keeps each stored level under the 2^20 rows after
 * which a bounded query leaves the lockstep join. */
int fsig_server_submit_bounded(fsig_server* s, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                               const uint32_t* qadj, uint64_t max_embeddings, uint64_t* ticket);
/* Work limit (partial embeddings, as fsig_set_work_limit) of every worker;
 * a request over it completes with FSIG_ELIMIT (refused) and counts in
 * fsig_server_stats.refused.  Set before submitting. */
int fsig_server_set_work_limit(fsig_server* s, uint64_t max_rows);
/* Waits up to timeout_ms (< 0: without limit) for a ticket; returns the
 * query's status (FSIG_EAGAIN: not finished, the ticket stays valid) and on
 * completion its count and submit-to-completion latency.  A completed
 * ticket is released by the call that returns it; any other call waiting on
 * it then returns FSIG_ENOTFOUND. */
int fsig_server_wait(fsig_server* s, uint64_t ticket, int64_t timeout_ms, uint64_t* count, double* latency_ms);
int fsig_server_get_stats(fsig_server* s, fsig_server_stats* out);
/* Projected bytes the server admits the query with. */
int fsig_server_projected(fsig_server* s, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                          const uint32_t* qadj, uint64_t* bytes);
const char* fsig_server_last_error(fsig_server* s);
/* Finishes every queued request, stops the workers and frees them. */
void fsig_server_destroy(fsig_server* s);

/* Admission projection (host only): arena bytes of the batched engine for
 * one count query on a graph of n vertices whose label l has label_count[l]
 * vertices and whose adjacency holds pair_edges[a * num_labels + b] entries
 * from a label-a to a label-b vertex.  Setup tables plus every stored join
 * level; level sizes follow an independence estimate (first backward
 * neighbour: mean label-matching degree; each further one: edge
 * probability).  Generalises the reference's per-level block sizing
 * (memory_manager/memory_manager.h:220-242) to a whole query. */
int fsig_project_bytes(uint32_t n, uint32_t num_labels, const uint64_t* label_count, const uint64_t* pair_edges,
                       uint32_t k, const int32_t* qlabels, const uint32_t* qoff, const uint32_t* qadj,
                       uint64_t* bytes);
/* Label-pair adjacency counts of a resident graph (num_labels^2 u64). */
int fsig_graph_label_pairs(fsig_ctx* ctx, const fsig_graph* g, uint64_t* pair_edges);

#ifdef __cplusplus
}
#endif
#endif /* FSIG_H */
