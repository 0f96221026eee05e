// Served request loop: continuous batching of count queries over one resident
// graph, with admission by projected intermediate-table bytes.
//
// The reference answers one query per process (main/run.cc:101-140) and its
// MemoryManager sizes blocks of one query's intermediate table against the
// device memory (memory_manager/memory_manager.h:220-242).  Here a pool of
// worker contexts (one HIP stream and one arena slice each) shares the
// resident graph.  A worker that becomes free takes the requests waiting at
// that moment, front first, while their projected bytes fit its arena, and
// runs them as one batch with the batched engine's kernels.  Requests that
// arrive while batches are on the device are taken by the next free worker,
// so the work of different batches overlaps on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "common.h"
#include "device.h"
#include "internal.h"

namespace {

// Largest label count the served loop and fsig_graph_label_pairs take: the
// label-pair table holds L * L u64 counts (128 MiB at 4096).
constexpr uint32_t kServeMaxLabels = 4096;

// Directed label-pair edge counts: pairs[a * L + b] = number of adjacency
// entries (v, w) with label(v) = a and label(w) = b.  One wavefront per
// vertex; per-block LDS histogram when L <= 64.
constexpr uint32_t kPairLds = 64;

__global__ void __launch_bounds__(FSIG_BLOCK) k_label_pairs(const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj,
                                                            const int32_t* __restrict__ lab, uint32_t n,
                                                            uint32_t L, unsigned long long* __restrict__ pairs) {
  __shared__ uint32_t hist[kPairLds * kPairLds];
  const bool use_lds = L <= kPairLds;
  if (use_lds)
    for (uint32_t i = threadIdx.x; i < L * L; i += FSIG_BLOCK) hist[i] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t waves = FSIG_BLOCK / 64;
  for (uint64_t v = (uint64_t)blockIdx.x * waves + wave; v < n; v += (uint64_t)gridDim.x * waves) {
    const uint32_t a = (uint32_t)lab[v];
    for (uint32_t j = off[v] + lane; j < off[v + 1]; j += 64) {
      const uint32_t b = (uint32_t)lab[adj[j]];
      if (use_lds) atomicAdd(&hist[a * L + b], 1u);
      else atomicAdd(&pairs[(uint64_t)a * L + b], 1ull);
    }
  }
  __syncthreads();
  if (use_lds)
    for (uint32_t i = threadIdx.x; i < L * L; i += FSIG_BLOCK)
      if (hist[i]) atomicAdd(&pairs[i], (unsigned long long)hist[i]);
}

}  // namespace

// Projected arena bytes of one count query in the batched engine: the
// per-query setup tables plus every intermediate level the lockstep join
// keeps, with level sizes from an independence estimate over label-pair edge
// counts (the first backward neighbour contributes the mean number of
// label-matching neighbours, each further one the probability of an edge).
static uint64_t project_bytes(uint32_t n, uint32_t L, const uint64_t* lcount, const uint64_t* pairs,
                              const QueryPlan& P, bool bounded = false) {
  const uint32_t k = P.k;
  for (uint32_t u = 0; u < k; u++)
    if (P.label[u] < 0 || (uint32_t)P.label[u] >= L) return 4096;  // label absent: the query is dead
  uint64_t bytes = 4ull * n;  // candidate mask words
  uint32_t csize[MAXQ];
  for (uint32_t a = 0; a < P.nclass; a++) {
    const uint64_t c = lcount[P.class_label[a]];
    csize[a] = (uint32_t)std::min<uint64_t>(c, 0xffffffffu);
    bytes += 9ull * c + 512;  // two candidate arrays and the keep flags
  }
  for (uint32_t a = 0; a < P.nclass; a++)
    for (uint32_t b = 0; b < P.nclass; b++) {
      bool adjacent = false;
      for (uint32_t u = 0; u < k && !adjacent; u++)
        if (P.cls[u] == a)
          for (uint32_t w = 0; w < k; w++)
            if (P.cls[w] == b && ((P.adjm[u] >> w) & 1u)) { adjacent = true; break; }
      if (adjacent)
        bytes += 4ull * (lcount[P.class_label[a]] + 1) + 4ull * pairs[(uint64_t)P.class_label[a] * L + P.class_label[b]];
    }
  QueryPlan Q = P;
  plan_order_mmatch(Q, csize);
  double rows = (double)lcount[P.label[Q.order[0]]];
  double level_bytes = 0;
  for (uint32_t p = 1; p < k; p++) {
    level_bytes += rows * (4.0 * p + 12.0);
    const uint32_t u = Q.order[p];
    const uint32_t lu = (uint32_t)P.label[u];
    bool first = true;
    double f = 1.0;
    for (uint32_t cp = 0; cp < p; cp++) {
      const uint32_t b = Q.order[cp];
      if (!((P.adjm[u] >> b) & 1u)) continue;
      const uint32_t lb = (uint32_t)P.label[b];
      const double e = (double)pairs[(uint64_t)lb * L + lu];
      const double nb = std::max<double>(1.0, (double)lcount[lb]);
      const double nu = std::max<double>(1.0, (double)lcount[lu]);
      f *= first ? e / nb : e / (nb * nu);
      first = false;
    }
    rows = std::min(rows * f, 1e15);
    // a bounded query leaves the lockstep join before a level passes 2^20
    // rows (the depth-first driver then keeps windows of the free arena)
    if (bounded) rows = std::min(rows, (double)(1u << 20));
    if (p + 1 == k) break;  // the last level is counted, not stored
  }
  const double total = (double)bytes + level_bytes;
  return total > 1.8e19 ? ~0ull : (uint64_t)total;
}

extern "C" int fsig_project_bytes(uint32_t n, uint32_t num_labels, const uint64_t* label_count,
                                  const uint64_t* pair_edges, uint32_t k, const int32_t* qlabels,
                                  const uint32_t* qoff, const uint32_t* qadj, uint64_t* bytes) {
  if (!label_count || !pair_edges || !qlabels || !qoff || !qadj || !bytes || num_labels == 0) return FSIG_EINVAL;
  QueryPlan P;
  std::string err;
  const int rc = plan_classes(P, k, qlabels, qoff, qadj, nullptr, FSIG_ALG_MMATCH, err);
  if (rc) return rc;
  *bytes = project_bytes(n, num_labels, label_count, pair_edges, P);
  return FSIG_OK;
}

extern "C" int fsig_graph_label_pairs(fsig_ctx* c, const fsig_graph* g, uint64_t* pair_edges) {
  if (!c) return FSIG_EINVAL;
  if (!g || !pair_edges) return set_err(c, FSIG_EINVAL, "label pairs: null argument");
  if (g->num_labels > kServeMaxLabels) return set_err(c, FSIG_EINVAL, "label pairs: more than 4096 labels");
  FSIG_HIP(c, hipSetDevice(c->device));
  const uint32_t L = g->num_labels;
  const size_t mk = c->arena.mark();
  unsigned long long* d = (unsigned long long*)c->arena.alloc(8ull * L * L);
  if (!d) return set_err(c, FSIG_ENOMEM, "label pairs: arena exhausted");
  k_fill_u32<<<std::max<uint32_t>(1, (2 * L * L + FSIG_BLOCK - 1) / FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(
      (uint32_t*)d, 2ull * L * L, 0);
  if (g->n)
    k_label_pairs<<<(uint32_t)c->num_cus * 8, FSIG_BLOCK, 0, c->stream>>>(g->d_off, g->d_adj, g->d_labels, g->n, L, d);
  FSIG_HIP(c, hipGetLastError());
  std::vector<uint64_t> h((size_t)L * L);
  int rc = FSIG_OK;
  for (size_t i = 0; i < h.size() && !rc; i += kPinnedWords)
    rc = read_u64(c, d + i, h.data() + i, std::min(kPinnedWords, h.size() - i));
  c->arena.release(mk);
  if (rc) return rc;
  std::memcpy(pair_edges, h.data(), 8 * h.size());
  return FSIG_OK;
}

// ---------------------------------------------------------------------------
// The server.
struct Request {
  uint64_t ticket;
  uint32_t k;
  uint64_t bound;  // 0: exact count, else min(count, bound)
  std::vector<int32_t> labels;
  std::vector<uint32_t> off, adj;
  uint64_t projected;
  std::chrono::steady_clock::time_point t_submit;
};

struct Result {
  int rc = FSIG_EAGAIN;
  uint64_t count = 0;
  double latency_ms = 0;
};

struct fsig_server {
  int device = 0;
  const fsig_graph* g = nullptr;
  std::vector<fsig_ctx*> workers;
  std::vector<std::thread> threads;
  uint64_t share = 0;  // admission budget of one worker's batch
  uint32_t max_batch = 0;
  std::vector<uint64_t> lcount, pairs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> pending;
  std::unordered_map<uint64_t, Result> results;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::string err;
  // statistics
  fsig_server_stats st{};
  uint32_t running = 0;
  bool started = false;
  std::chrono::steady_clock::time_point t_first, t_last;
};

static void worker_loop(fsig_server* s, uint32_t wi) {
  fsig_ctx* c = s->workers[wi];
  for (;;) {
    std::vector<Request> batch;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
      if (s->pending.empty()) return;  // stop requested and nothing queued
      uint64_t proj = 0;
      while (!s->pending.empty() && batch.size() < s->max_batch) {
        const Request& r = s->pending.front();
        if (!batch.empty() && proj + r.projected > s->share) break;
        proj += r.projected;
        batch.push_back(std::move(s->pending.front()));
        s->pending.pop_front();
      }
      s->running++;
      s->st.max_concurrent = std::max(s->st.max_concurrent, s->running);
      s->st.batches++;
      s->st.max_batch_seen = std::max<uint64_t>(s->st.max_batch_seen, batch.size());
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t nb = (uint32_t)batch.size();
    std::vector<uint32_t> ks(nb);
    std::vector<const int32_t*> ql(nb);
    std::vector<const uint32_t*> qo(nb), qa(nb);
    for (uint32_t i = 0; i < nb; i++) {
      ks[i] = batch[i].k;
      ql[i] = batch[i].labels.data();
      qo[i] = batch[i].off.data();
      qa[i] = batch[i].adj.data();
    }
    std::vector<uint64_t> counts(nb, 0), bounds(nb, 0);
    for (uint32_t i = 0; i < nb; i++) bounds[i] = batch[i].bound;
    std::vector<int32_t> st(nb, FSIG_OK);
    std::vector<int> rcs(nb, FSIG_OK);
    int rc = fsig_match_batch_bounded(c, s->g, nb, ks.data(), ql.data(), qo.data(), qa.data(), bounds.data(),
                                      counts.data(), st.data(), nullptr);
    uint64_t fallbacks = 0, refused = 0;
    if (rc == FSIG_ENOMEM) {
      // the projection was short: each query alone through the plain path,
      // whose join is memory-blocked
      for (uint32_t i = 0; i < nb; i++) {
        rcs[i] = fsig_match(c, s->g, ks[i], ql[i], qo[i], qa[i], nullptr, FSIG_ALG_MMATCH, FSIG_MODE_COUNT,
                            bounds[i], &counts[i], nullptr, 0);
        fallbacks++;
      }
    } else if (rc) {
      for (uint32_t i = 0; i < nb; i++) rcs[i] = rc;
    } else {
      for (uint32_t i = 0; i < nb; i++) rcs[i] = st[i];
    }
    for (uint32_t i = 0; i < nb; i++) refused += rcs[i] == FSIG_ELIMIT;
    const auto t1 = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (uint32_t i = 0; i < nb; i++) {
        Result& r = s->results[batch[i].ticket];
        r.rc = rcs[i];
        r.count = counts[i];
        r.latency_ms = std::chrono::duration<double, std::milli>(t1 - batch[i].t_submit).count();
      }
      if (rc && rc != FSIG_ENOMEM) s->err = fsig_last_error(c);
      s->running--;
      s->st.completed += nb;
      s->st.fallbacks += fallbacks;
      s->st.refused += refused;
      s->st.busy_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
      s->t_last = t1;
    }
    s->cv_done.notify_all();
  }
}

extern "C" int fsig_server_create(int device, const fsig_graph* g, uint32_t workers, uint64_t hbm_budget,
                                  uint32_t max_batch, const fsig_tune* tune, fsig_server** out) {
  if (!out || !g || workers == 0 || workers > 16 || max_batch == 0) return FSIG_EINVAL;
  *out = nullptr;
  // admission keeps an L x L table of label-pair adjacency counts on the host
  if (g->num_labels > kServeMaxLabels) return FSIG_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return FSIG_EHIP;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return FSIG_EHIP;
  if (hbm_budget == 0) hbm_budget = (uint64_t)freeb / 4 * 3;
  fsig_server* s = new fsig_server();
  s->device = device;
  s->g = g;
  s->max_batch = max_batch;
  const uint64_t per = hbm_budget / workers;
  // a batch is admitted against 3/4 of its worker's arena: the rest covers
  // the tables and scan temporaries the projection leaves out
  s->share = per / 4 * 3;
  fsig_tune t{0, 3, 0, 0, 0, 0, 0};  // lockstep join, lanes per query, no autotuning
  if (tune) t = *tune;
  for (uint32_t w = 0; w < workers; w++) {
    fsig_ctx* c = nullptr;
    int rc = fsig_ctx_create(device, per, &c);
    if (rc == FSIG_OK) rc = fsig_set_tune(c, &t);
    if (rc) {
      if (c) fsig_ctx_destroy(c);
      for (fsig_ctx* x : s->workers) fsig_ctx_destroy(x);
      delete s;
      return rc;
    }
    s->workers.push_back(c);
  }
  const uint32_t L = g->num_labels;
  s->lcount.resize(L);
  for (uint32_t a = 0; a < L; a++) s->lcount[a] = g->h_inv_index[a + 1] - g->h_inv_index[a];
  s->pairs.resize((size_t)L * L);
  int rc = fsig_graph_label_pairs(s->workers[0], g, s->pairs.data());
  if (rc) {
    for (fsig_ctx* x : s->workers) fsig_ctx_destroy(x);
    delete s;
    return rc;
  }
  for (uint32_t w = 0; w < workers; w++) s->threads.emplace_back(worker_loop, s, w);
  *out = s;
  return FSIG_OK;
}

extern "C" int fsig_server_submit(fsig_server* s, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                                  const uint32_t* qadj, uint64_t* ticket) {
  return fsig_server_submit_bounded(s, k, qlabels, qoff, qadj, 0, ticket);
}

extern "C" int fsig_server_submit_bounded(fsig_server* s, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                                          const uint32_t* qadj, uint64_t max_embeddings, uint64_t* ticket) {
  if (!s || !qlabels || !qoff || !qadj || !ticket) return FSIG_EINVAL;
  QueryPlan P;
  std::string err;
  const int rc = plan_classes(P, k, qlabels, qoff, qadj, nullptr, FSIG_ALG_MMATCH, err);
  if (rc) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = err;
    return rc;
  }
  Request r;
  r.k = k;
  r.bound = max_embeddings;
  r.labels.assign(qlabels, qlabels + k);
  r.off.assign(qoff, qoff + k + 1);
  r.adj.assign(qadj, qadj + qoff[k]);
  r.projected = project_bytes(s->g->n, s->g->num_labels, s->lcount.data(), s->pairs.data(), P, max_embeddings != 0);
  r.t_submit = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    if (!s->started) {
      s->started = true;
      s->t_first = r.t_submit;
    }
    s->st.submitted++;
    s->results[r.ticket] = Result{};
    s->pending.push_back(std::move(r));
  }
  s->cv_work.notify_one();
  return FSIG_OK;
}

extern "C" int fsig_server_wait(fsig_server* s, uint64_t ticket, int64_t timeout_ms, uint64_t* count,
                                double* latency_ms) {
  if (!s) return FSIG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->results.find(ticket) == s->results.end()) return FSIG_ENOTFOUND;
  // ready when finished, or gone (another waiter on the same ticket took it)
  auto ready = [&] {
    auto it = s->results.find(ticket);
    return it == s->results.end() || it->second.rc != FSIG_EAGAIN;
  };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return FSIG_EAGAIN;
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return FSIG_ENOTFOUND;
  const Result r = it->second;
  s->results.erase(it);
  s->cv_done.notify_all();
  if (count) *count = r.count;
  if (latency_ms) *latency_ms = r.latency_ms;
  return r.rc;
}

extern "C" int fsig_server_set_work_limit(fsig_server* s, uint64_t max_rows) {
  if (!s) return FSIG_EINVAL;
  // workers read the limit when a batch starts; set it before submitting
  for (fsig_ctx* c : s->workers) fsig_set_work_limit(c, max_rows);
  return FSIG_OK;
}

extern "C" int fsig_server_get_stats(fsig_server* s, fsig_server_stats* out) {
  if (!s || !out) return FSIG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  // first submit to last completion; 0 until a batch has completed
  out->wall_ms = s->started && s->st.completed
                     ? std::chrono::duration<double, std::milli>(s->t_last - s->t_first).count()
                     : 0;
  return FSIG_OK;
}

extern "C" int fsig_server_projected(fsig_server* s, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                                     const uint32_t* qadj, uint64_t* bytes) {
  if (!s) return FSIG_EINVAL;
  return fsig_project_bytes(s->g->n, s->g->num_labels, s->lcount.data(), s->pairs.data(), k, qlabels, qoff, qadj,
                            bytes);
}

extern "C" const char* fsig_server_last_error(fsig_server* s) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> lk(s->mu);
  thread_local std::string copy;
  copy = s->err;
  return copy.c_str();
}

extern "C" void fsig_server_destroy(fsig_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads) t.join();
  for (fsig_ctx* c : s->workers) fsig_ctx_destroy(c);
  delete s;
}
