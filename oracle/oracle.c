/*
 * oracle.c - CPU restatement of the reference's matching semantics, used ONLY
 * as the checker by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  Nothing in the product links or calls it.
 *
 * Functions and the reference lines they restate:
 *   oracle_murmur2      MurmurHash2 as restated at
 *                       graph_format_transformer/graph_format_transformer.h:59-93
 *                       (pinned by oracle/_ref known answers)
 *   oracle_signature    build_signature, graph_format_transformer.cc:217-281,
 *                       including the shift 2*(2h mod 32) reduced mod 32 that
 *                       x86 applies (pinned by oracle/_ref tables)
 *   oracle_ldf          initialise_kernel, match/MMatch_kernels.cuh:6-16:
 *                       label equal and degree >= query degree
 *                       (pinned by its closed form, checked in numpy)
 *   oracle_sig_filter   filter_kernel, match/GSI_kernels.cuh:8-24
 *                       (pinned by soundness against the matcher)
 *   oracle_refine       count rule of count_neighbors_kernel + check_re_kernel,
 *                       match/MMatch_kernels.cuh:28-117, 190-214, over
 *                       colour-refinement classes, Jacobi passes
 *                       (pinned by invariants: it contains every embedding's
 *                       image and is contained in the LDF sets)
 *   oracle_match        the output contract of MMatch / GSI: every injective,
 *                       label-preserving, edge-preserving map; edges are checked
 *                       through backward neighbours as compute_kernel does
 *                       (match/MMatch_kernels.cuh:456-708), injectivity as at
 *                       :641-707.  Pinned by closed forms (tr(A^3), 4-cycles,
 *                       K_n counts), brute force over all injective maps on
 *                       small graphs, and networkx VF2 golden fixtures.
 *   oracle_match_digest the same enumeration folded into the order-independent
 *                       set digest of fsig_match_digest (row hash: splitmix64
 *                       finaliser chained over the row in query-vertex
 *                       order); pinned by a numpy restatement of the digest
 *                       over the rows oracle_match lists.
 *   oracle_check_rows   validity of given rows (labels, injectivity, every
 *                       query edge present) and their row hashes; pinned by
 *                       planted mistakes in known embeddings.
 *   oracle_match_budget oracle_match_limit under a step budget (candidates
 *                       tried), for checks that skip the queries whose search
 *                       is too large; a finished search is oracle_match's.
 *   oracle_triangles    number of triangles by degree-ordered orientation
 *                       (each triangle once; the triangle query's embeddings
 *                       are 6 times it); pinned by oracle_match and tr(A^3).
 * Parity unpinned: none of the functions above.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OMAXQ 32
#define OINVALID 0xffffffffu

uint32_t oracle_murmur2(const void* key, int len, uint32_t seed) {
  const uint32_t m = 0x5bd1e995u;
  const unsigned char* p = (const unsigned char*)key;
  uint32_t h = seed ^ (uint32_t)len;
  while (len >= 4) {
    uint32_t k = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    k *= m;
    k ^= k >> 24;
    k *= m;
    h *= m;
    h ^= k;
    p += 4;
    len -= 4;
  }
  switch (len) {
    case 3: h ^= (uint32_t)p[2] << 16; /* fall through */
    case 2: h ^= (uint32_t)p[1] << 8;  /* fall through */
    case 1: h ^= p[0]; h *= m;
  }
  h ^= h >> 13;
  h *= m;
  h ^= h >> 15;
  return h;
}

/* out: [16][n] if colmajor else [n][16] */
void oracle_signature(uint32_t n, const uint32_t* off, const uint32_t* adj, const int32_t* labels, int colmajor,
                      uint32_t* out) {
  uint32_t sig[16];
  for (uint32_t v = 0; v < n; v++) {
    memset(sig, 0, sizeof sig);
    sig[0] = (uint32_t)labels[v];
    for (uint32_t i = off[v]; i < off[v + 1]; i++) {
      int32_t l = labels[adj[i]];
      uint32_t hp = oracle_murmur2(&l, 4, 17) % 240u;
      uint32_t a = 2 * hp / 32, b = 2 * hp % 32;
      uint32_t sh = (2 * b) % 32;
      uint32_t t = sig[1 + a];
      uint32_t c = (t >> sh) & 3u;
      c = c == 0 ? 1u : 3u;
      sig[1 + a] = t | (c << sh);
    }
    for (int w = 0; w < 16; w++) {
      if (colmajor) out[(uint64_t)w * n + v] = sig[w];
      else out[(uint64_t)v * 16 + w] = sig[w];
    }
  }
}

/* Writes the ascending candidates of (label, degree) into out; returns count. */
uint32_t oracle_ldf(uint32_t n, const uint32_t* off, const int32_t* labels, int32_t label, uint32_t mindeg,
                    uint32_t* out) {
  uint32_t c = 0;
  for (uint32_t v = 0; v < n; v++)
    if (labels[v] == label && off[v + 1] - off[v] >= mindeg) out[c++] = v;
  return c;
}

/* qsig: 16 words of one query vertex; data sig column-major [16][n]. */
uint32_t oracle_sig_filter(uint32_t n, const uint32_t* dsig, const uint32_t* qsig, uint32_t* out) {
  uint32_t c = 0;
  for (uint32_t v = 0; v < n; v++) {
    int ok = dsig[v] == qsig[0];
    for (int w = 1; w < 16 && ok; w++) ok = (qsig[w] & dsig[(uint64_t)w * n + v]) == qsig[w];
    if (ok) out[c++] = v;
  }
  return c;
}

/* ---------------------------------------------------------------------------
 * Colour-refinement classes of a query (k <= 32): start from (label, degree),
 * refine by the multiset of neighbour colours until the number of colours is
 * stable; classes numbered by first occurrence. */
static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

static uint32_t recolour(uint32_t k, uint32_t keys[][OMAXQ + 2], uint32_t* len, uint32_t* color) {
  /* colour = rank of key among distinct keys in lexicographic order */
  uint32_t ncol = 0;
  for (uint32_t u = 0; u < k; u++) {
    uint32_t rank = 0;
    int seen_smaller[OMAXQ] = {0};
    for (uint32_t w = 0; w < k; w++) {
      /* compare keys[w] < keys[u] */
      uint32_t m = len[w] < len[u] ? len[w] : len[u];
      int cmp = 0;
      for (uint32_t i = 0; i < m && !cmp; i++) cmp = keys[w][i] < keys[u][i] ? -1 : (keys[w][i] > keys[u][i]);
      if (!cmp) cmp = len[w] < len[u] ? -1 : (len[w] > len[u]);
      if (cmp < 0) {
        /* count distinct smaller keys: only the first w with this key */
        int dup = 0;
        for (uint32_t x = 0; x < w && !dup; x++) {
          if (!seen_smaller[x]) continue;
          if (len[x] != len[w]) continue;
          dup = memcmp(keys[x], keys[w], len[w] * 4) == 0;
        }
        seen_smaller[w] = 1;
        if (!dup) rank++;
      }
    }
    color[u] = rank;
    if (rank + 1 > ncol) ncol = rank + 1;
  }
  return ncol;
}

uint32_t oracle_classes(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t* cls) {
  uint32_t keys[OMAXQ][OMAXQ + 2], len[OMAXQ], color[OMAXQ];
  for (uint32_t u = 0; u < k; u++) {
    keys[u][0] = (uint32_t)qlabels[u];
    keys[u][1] = (uint32_t)__builtin_popcount(qadjm[u]);
    len[u] = 2;
  }
  uint32_t ncol = recolour(k, keys, len, color);
  for (;;) {
    for (uint32_t u = 0; u < k; u++) {
      keys[u][0] = color[u];
      len[u] = 1;
      for (uint32_t w = 0; w < k; w++)
        if ((qadjm[u] >> w) & 1u) keys[u][len[u]++] = color[w];
      qsort(&keys[u][1], len[u] - 1, 4, cmp_u32);
    }
    uint32_t nc = recolour(k, keys, len, color);
    if (nc == ncol) break;
    ncol = nc;
  }
  uint32_t remap[OMAXQ], nclass = 0;
  for (uint32_t i = 0; i < OMAXQ; i++) remap[i] = OINVALID;
  for (uint32_t u = 0; u < k; u++) {
    if (remap[color[u]] == OINVALID) remap[color[u]] = nclass++;
    cls[u] = remap[color[u]];
  }
  return nclass;
}

/* Candidate sets after LDF and `passes` Jacobi passes of the count rule:
 * v in C(a) survives iff |N(v) ∩ C(b)| >= (neighbours a member of a has in
 * b) for every class b.  sizes[u] receives |C(class of u)|; if cand != NULL it
 * receives, per query vertex u, n flags (cand[u*n + v]). */
void oracle_refine(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n, const uint32_t* off,
                   const uint32_t* adj, const int32_t* labels, uint32_t passes, uint32_t* sizes, uint8_t* cand) {
  uint32_t cls[OMAXQ];
  uint32_t nc = oracle_classes(k, qlabels, qadjm, cls);
  uint32_t rep[OMAXQ];
  for (uint32_t u = k; u-- > 0;) rep[cls[u]] = u;
  uint32_t thr[OMAXQ][OMAXQ];
  memset(thr, 0, sizeof thr);
  for (uint32_t a = 0; a < nc; a++)
    for (uint32_t w = 0; w < k; w++)
      if ((qadjm[rep[a]] >> w) & 1u) thr[a][cls[w]]++;
  uint32_t* mask = (uint32_t*)calloc(n ? n : 1, 4);
  uint32_t* next = (uint32_t*)calloc(n ? n : 1, 4);
  for (uint32_t v = 0; v < n; v++)
    for (uint32_t a = 0; a < nc; a++)
      if (labels[v] == qlabels[rep[a]] && off[v + 1] - off[v] >= (uint32_t)__builtin_popcount(qadjm[rep[a]]))
        mask[v] |= 1u << a;
  int any = 0;
  for (uint32_t a = 0; a < nc; a++)
    for (uint32_t b = 0; b < nc; b++) any |= thr[a][b] != 0;
  for (uint32_t p = 0; any && p < passes; p++) {
    int changed = 0, empty = 0;
    uint32_t cnt[OMAXQ];
    for (uint32_t a = 0; a < nc; a++) cnt[a] = 0;
    for (uint32_t v = 0; v < n; v++)
      for (uint32_t a = 0; a < nc; a++) cnt[a] += (mask[v] >> a) & 1u;
    for (uint32_t a = 0; a < nc; a++) empty |= cnt[a] == 0;
    if (empty) break;
    for (uint32_t v = 0; v < n; v++) {
      uint32_t m = mask[v], keep = m;
      for (uint32_t a = 0; a < nc; a++) {
        if (!((m >> a) & 1u)) continue;
        for (uint32_t b = 0; b < nc && ((keep >> a) & 1u); b++) {
          if (!thr[a][b]) continue;
          uint32_t got = 0;
          for (uint32_t i = off[v]; i < off[v + 1]; i++) got += (mask[adj[i]] >> b) & 1u;
          if (got < thr[a][b]) keep &= ~(1u << a);
        }
      }
      next[v] = keep;
      changed |= keep != m;
    }
    memcpy(mask, next, (size_t)n * 4);
    if (!changed) break;
  }
  for (uint32_t u = 0; u < k; u++) {
    uint32_t s = 0;
    for (uint32_t v = 0; v < n; v++) {
      uint32_t bit = (mask[v] >> cls[u]) & 1u;
      s += bit;
      if (cand) cand[(uint64_t)u * n + v] = (uint8_t)bit;
    }
    sizes[u] = s;
  }
  free(mask);
  free(next);
}

/* ---------------------------------------------------------------------------
 * Backtracking matcher. */
typedef struct {
  uint32_t k, n;
  const int32_t* ql;
  const uint32_t* qadjm;
  const uint32_t* off;
  const uint32_t* adj;
  const int32_t* labels;
  uint32_t order[OMAXQ];
  uint32_t back[OMAXQ][OMAXQ]; /* backward neighbours (positions) of position p */
  uint32_t nback[OMAXQ];
  uint32_t qdeg[OMAXQ];
  /* output */
  uint64_t max_rows;   /* enumeration bound, 0 = none */
  uint32_t* rows;      /* k words per row, query-vertex order */
  uint64_t cap;        /* rows capacity */
  uint64_t count;
  uint64_t written;
  pthread_mutex_t lock;
  /* dynamic root distribution */
  uint32_t next_root;
  uint32_t nroots;
  uint32_t* roots;
  /* abort limit: stop once this many embeddings were seen (0 = none) */
  uint64_t limit;
  uint64_t seen;
  int stop;
  int digest;          /* fold every embedding into dig[] */
  /* step budget: stop once this many candidates were tried (0 = none) */
  uint64_t max_steps;
  uint64_t steps;
  int over_budget;
} Match;

static int has_edge(const Match* M, uint32_t a, uint32_t b) {
  uint32_t lo = M->off[a], hi = M->off[a + 1];
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (M->adj[mid] < b) lo = mid + 1; else hi = mid;
  }
  return lo < M->off[a + 1] && M->adj[lo] == b;
}

typedef struct {
  Match* M;
  uint32_t phi[OMAXQ];
  uint64_t count;
  uint32_t buf[64 * OMAXQ];
  uint32_t nbuf;
  uint64_t dig[2];
  uint64_t steps;
} Worker;

/* Row hash of the set digest: splitmix64's finaliser (public domain) chained
 * over the row's words in query-vertex order. */
static uint64_t omix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

uint64_t oracle_row_hash(uint32_t k, const uint32_t* row) {
  uint64_t x = 0x9e3779b97f4a7c15ull * (uint64_t)(k + 1);
  for (uint32_t i = 0; i < k; i++) x = omix64(x + (uint64_t)row[i] + ((uint64_t)i << 32));
  return x;
}

static void flush(Worker* W) {
  Match* M = W->M;
  if (!W->nbuf) return;
  pthread_mutex_lock(&M->lock);
  for (uint32_t i = 0; i < W->nbuf && M->written < M->cap; i++, M->written++)
    memcpy(M->rows + M->written * M->k, W->buf + i * M->k, M->k * 4);
  pthread_mutex_unlock(&M->lock);
  W->nbuf = 0;
}

static void emit(Worker* W) {
  Match* M = W->M;
  W->count++;
  if (M->limit && (W->count & 1023u) == 0) {
    uint64_t seen = __atomic_add_fetch(&M->seen, 1024, __ATOMIC_RELAXED);
    if (seen > M->limit) __atomic_store_n(&M->stop, 1, __ATOMIC_RELAXED);
  }
  if (M->digest) {
    uint32_t r[OMAXQ];
    for (uint32_t p = 0; p < M->k; p++) r[M->order[p]] = W->phi[p];
    const uint64_t x = oracle_row_hash(M->k, r);
    W->dig[0] += x;
    W->dig[1] += omix64(x ^ 0xd6e8feb86659fd93ull);
  }
  if (!M->rows) return;
  uint32_t* r = W->buf + W->nbuf * M->k;
  for (uint32_t p = 0; p < M->k; p++) r[M->order[p]] = W->phi[p];
  if (++W->nbuf == 64) flush(W);
}

static void extend(Worker* W, uint32_t p) {
  Match* M = W->M;
  if (p == M->k) { emit(W); return; }
  if ((M->limit || M->max_steps) && __atomic_load_n(&M->stop, __ATOMIC_RELAXED)) return;
  const uint32_t u = M->order[p];
  /* iterate the neighbours of the backward neighbour with the smallest image degree */
  uint32_t pivot = OINVALID, best = OINVALID;
  for (uint32_t i = 0; i < M->nback[p]; i++) {
    uint32_t v = W->phi[M->back[p][i]];
    uint32_t d = M->off[v + 1] - M->off[v];
    if (d < best) { best = d; pivot = v; }
  }
  uint32_t lo, hi;
  if (pivot == OINVALID) { lo = 0; hi = M->n; } else { lo = M->off[pivot]; hi = M->off[pivot + 1]; }
  for (uint32_t i = lo; i < hi; i++) {
    if (M->max_steps && (++W->steps & 65535u) == 0) {
      const uint64_t s = __atomic_add_fetch(&M->steps, 65536, __ATOMIC_RELAXED);
      if (s > M->max_steps) {
        __atomic_store_n(&M->over_budget, 1, __ATOMIC_RELAXED);
        __atomic_store_n(&M->stop, 1, __ATOMIC_RELAXED);
      }
      if (__atomic_load_n(&M->stop, __ATOMIC_RELAXED)) return;
    }
    uint32_t x = pivot == OINVALID ? i : M->adj[i];
    if (M->labels[x] != M->ql[u]) continue;
    if (M->off[x + 1] - M->off[x] < M->qdeg[u]) continue;
    int ok = 1;
    for (uint32_t j = 0; j < p && ok; j++) ok = W->phi[j] != x;
    for (uint32_t j = 0; j < M->nback[p] && ok; j++) {
      uint32_t v = W->phi[M->back[p][j]];
      if (v != pivot) ok = has_edge(M, v, x);
    }
    if (!ok) continue;
    W->phi[p] = x;
    extend(W, p + 1);
  }
}

static void* worker_main(void* arg) {
  Worker* W = (Worker*)arg;
  Match* M = W->M;
  for (;;) {
    uint32_t r = __atomic_fetch_add(&M->next_root, 1, __ATOMIC_RELAXED);
    if (r >= M->nroots || ((M->limit || M->max_steps) && __atomic_load_n(&M->stop, __ATOMIC_RELAXED))) break;
    W->phi[0] = M->roots[r];
    extend(W, 1);
  }
  flush(W);
  return NULL;
}

/* Counts (and, when rows != NULL, lists up to cap rows of) the embeddings of
 * the query (k vertices, adjacency bitmasks qadjm) in the data graph.  With
 * threads > 1 the roots are handed out one at a time to a pthread pool.
 * With limit > 0 the search stops once more than `limit` embeddings were
 * seen; *stopped tells whether it did (the count is then partial). */
static uint64_t match_run(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n,
                          const uint32_t* off, const uint32_t* adj, const int32_t* labels, uint32_t threads,
                          uint32_t* rows, uint64_t cap, uint64_t* written, uint64_t limit, int* stopped,
                          uint64_t* digest, uint64_t max_steps, int* over_budget) {
  if (digest) digest[0] = digest[1] = 0;
  if (k == 0 || k > OMAXQ) return 0;
  Match M;
  memset(&M, 0, sizeof M);
  M.digest = digest != NULL;
  M.k = k;
  M.n = n;
  M.ql = qlabels;
  M.qadjm = qadjm;
  M.off = off;
  M.adj = adj;
  M.labels = labels;
  M.rows = rows;
  M.cap = rows ? cap : 0;
  M.limit = limit;
  M.max_steps = max_steps;
  for (uint32_t u = 0; u < k; u++) M.qdeg[u] = (uint32_t)__builtin_popcount(qadjm[u]);
  /* order: max degree first, then the vertex with most ordered neighbours */
  uint32_t used = 0;
  for (uint32_t p = 0; p < k; p++) {
    uint32_t best = OINVALID;
    int bk = -1, bd = -1;
    for (uint32_t u = 0; u < k; u++) {
      if ((used >> u) & 1u) continue;
      int back = __builtin_popcount(qadjm[u] & used);
      int d = (int)M.qdeg[u];
      if (back > bk || (back == bk && d > bd)) { bk = back; bd = d; best = u; }
    }
    M.order[p] = best;
    used |= 1u << best;
  }
  for (uint32_t p = 0; p < k; p++)
    for (uint32_t q = 0; q < p; q++)
      if ((qadjm[M.order[p]] >> M.order[q]) & 1u) M.back[p][M.nback[p]++] = q;
  M.roots = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  const uint32_t u0 = M.order[0];
  for (uint32_t v = 0; v < n; v++)
    if (labels[v] == qlabels[u0] && off[v + 1] - off[v] >= M.qdeg[u0]) M.roots[M.nroots++] = v;
  pthread_mutex_init(&M.lock, NULL);
  if (threads < 1) threads = 1;
  Worker* W = (Worker*)calloc(threads, sizeof(Worker));
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  for (uint32_t t = 0; t < threads; t++) W[t].M = &M;
  if (threads == 1) {
    worker_main(&W[0]);
  } else {
    for (uint32_t t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker_main, &W[t]);
    for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
  }
  uint64_t total = 0;
  for (uint32_t t = 0; t < threads; t++) {
    total += W[t].count;
    if (digest) {
      digest[0] += W[t].dig[0];
      digest[1] += W[t].dig[1];
    }
  }
  if (stopped) *stopped = M.stop && !M.over_budget;
  if (over_budget) *over_budget = M.over_budget;
  if (written) *written = M.written;
  pthread_mutex_destroy(&M.lock);
  free(W);
  free(th);
  free(M.roots);
  return total;
}

uint64_t oracle_match_limit(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n,
                            const uint32_t* off, const uint32_t* adj, const int32_t* labels, uint32_t threads,
                            uint32_t* rows, uint64_t cap, uint64_t* written, uint64_t limit, int* stopped) {
  return match_run(k, qlabels, qadjm, n, off, adj, labels, threads, rows, cap, written, limit, stopped, NULL, 0,
                   NULL);
}

/* oracle_match_limit with a step budget: the search also stops once about
 * max_steps candidates were tried, and *over_budget tells whether it did
 * (the count is then partial and says nothing). */
uint64_t oracle_match_budget(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n,
                             const uint32_t* off, const uint32_t* adj, const int32_t* labels, uint32_t threads,
                             uint64_t limit, uint64_t max_steps, int* stopped, int* over_budget) {
  return match_run(k, qlabels, qadjm, n, off, adj, labels, threads, NULL, 0, NULL, limit, stopped, NULL, max_steps,
                   over_budget);
}

uint64_t oracle_match(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n, const uint32_t* off,
                      const uint32_t* adj, const int32_t* labels, uint32_t threads, uint32_t* rows, uint64_t cap,
                      uint64_t* written) {
  return match_run(k, qlabels, qadjm, n, off, adj, labels, threads, rows, cap, written, 0, NULL, NULL, 0, NULL);
}

/* Count and set digest (digest[0] = sum of row hashes x, digest[1] = sum of
 * omix64(x ^ salt), both mod 2^64) of every embedding. */
uint64_t oracle_match_digest(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n,
                             const uint32_t* off, const uint32_t* adj, const int32_t* labels, uint32_t threads,
                             uint64_t* digest) {
  return match_run(k, qlabels, qadjm, n, off, adj, labels, threads, NULL, 0, NULL, 0, NULL, digest, 0, NULL);
}

/* ---------------------------------------------------------------------------
 * Checks rows (k words each, query-vertex order) against the contract:
 * every word a vertex of the graph with the query vertex's label, the words
 * distinct, and every query edge a data edge.  Returns the number of rows
 * that fail; hashes (may be NULL) receives each row's oracle_row_hash. */
typedef struct {
  uint32_t k, n;
  const int32_t* ql;
  const uint32_t* qadjm;
  const uint32_t* off;
  const uint32_t* adj;
  const int32_t* labels;
  const uint32_t* rows;
  uint64_t nrows;
  uint64_t* hashes;
  uint64_t next;
  uint64_t bad;
} Check;

static int row_ok(const Check* K, const uint32_t* r) {
  Match M;
  M.off = K->off;
  M.adj = K->adj;
  for (uint32_t i = 0; i < K->k; i++) {
    if (r[i] >= K->n || K->labels[r[i]] != K->ql[i]) return 0;
    for (uint32_t j = 0; j < i; j++)
      if (r[j] == r[i]) return 0;
  }
  for (uint32_t i = 0; i < K->k; i++)
    for (uint32_t j = i + 1; j < K->k; j++)
      if (((K->qadjm[i] >> j) & 1u) && !has_edge(&M, r[i], r[j])) return 0;
  return 1;
}

static void* check_main(void* arg) {
  Check* K = (Check*)arg;
  uint64_t bad = 0;
  for (;;) {
    uint64_t at = __atomic_fetch_add(&K->next, 65536, __ATOMIC_RELAXED);
    if (at >= K->nrows) break;
    uint64_t end = at + 65536 < K->nrows ? at + 65536 : K->nrows;
    for (uint64_t r = at; r < end; r++) {
      const uint32_t* row = K->rows + r * K->k;
      bad += !row_ok(K, row);
      if (K->hashes) K->hashes[r] = oracle_row_hash(K->k, row);
    }
  }
  __atomic_fetch_add(&K->bad, bad, __ATOMIC_RELAXED);
  return NULL;
}

uint64_t oracle_check_rows(uint32_t k, const int32_t* qlabels, const uint32_t* qadjm, uint32_t n,
                           const uint32_t* off, const uint32_t* adj, const int32_t* labels, const uint32_t* rows,
                           uint64_t nrows, uint32_t threads, uint64_t* hashes) {
  Check K;
  memset(&K, 0, sizeof K);
  K.k = k; K.n = n; K.ql = qlabels; K.qadjm = qadjm; K.off = off; K.adj = adj; K.labels = labels;
  K.rows = rows; K.nrows = nrows; K.hashes = hashes;
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  for (uint32_t t = 0; t < threads; t++) pthread_create(&th[t], NULL, check_main, &K);
  for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(th);
  return K.bad;
}

/* ---------------------------------------------------------------------------
 * Triangles by degree-ordered orientation: each edge points from the vertex
 * of lower (degree, id) rank to the higher; a triangle is counted once, at
 * its lowest vertex v and middle vertex u, as a common out-neighbour of both. */
typedef struct {
  uint32_t n;
  const uint64_t* ooff;
  const uint32_t* oadj;
  uint32_t next;
  uint64_t total;
} Tri;

static int rank_less(const uint32_t* off, uint32_t a, uint32_t b) {
  uint32_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
  return da < db || (da == db && a < b);
}

static void* tri_main(void* arg) {
  Tri* T = (Tri*)arg;
  uint64_t cnt = 0;
  for (;;) {
    uint32_t v0 = __atomic_fetch_add(&T->next, 256, __ATOMIC_RELAXED);
    if (v0 >= T->n) break;
    uint32_t v1 = v0 + 256 < T->n ? v0 + 256 : T->n;
    for (uint32_t v = v0; v < v1; v++)
      for (uint64_t i = T->ooff[v]; i < T->ooff[v + 1]; i++) {
        uint32_t u = T->oadj[i];
        uint64_t a = T->ooff[v], ae = T->ooff[v + 1], b = T->ooff[u], be = T->ooff[u + 1];
        while (a < ae && b < be) {
          if (T->oadj[a] < T->oadj[b]) a++;
          else if (T->oadj[a] > T->oadj[b]) b++;
          else { cnt++; a++; b++; }
        }
      }
  }
  __atomic_fetch_add(&T->total, cnt, __ATOMIC_RELAXED);
  return NULL;
}

uint64_t oracle_triangles(uint32_t n, const uint32_t* off, const uint32_t* adj, uint32_t threads) {
  uint64_t* ooff = (uint64_t*)calloc((size_t)n + 1, sizeof(uint64_t));
  for (uint32_t v = 0; v < n; v++) {
    uint64_t d = 0;
    for (uint32_t i = off[v]; i < off[v + 1]; i++) d += rank_less(off, v, adj[i]);
    ooff[v + 1] = ooff[v] + d;
  }
  uint32_t* oadj = (uint32_t*)malloc(sizeof(uint32_t) * (ooff[n] ? ooff[n] : 1));
  for (uint32_t v = 0; v < n; v++) {
    uint64_t at = ooff[v];
    for (uint32_t i = off[v]; i < off[v + 1]; i++)
      if (rank_less(off, v, adj[i])) oadj[at++] = adj[i]; /* ascending ids, as adj */
  }
  Tri T;
  memset(&T, 0, sizeof T);
  T.n = n;
  T.ooff = ooff;
  T.oadj = oadj;
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  for (uint32_t t = 0; t < threads; t++) pthread_create(&th[t], NULL, tri_main, &T);
  for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(th);
  free(ooff);
  free(oadj);
  return T.total;
}
