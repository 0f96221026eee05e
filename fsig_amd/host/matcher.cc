// GPUSubgraphMatcher / MMatch / GSI on top of the C ABI.
#include "matcher.h"

#include <algorithm>
#include <cstdio>

namespace fsig {

void TimeRecorder::event_start(const std::string& name) {
  events_.push_back(name);
  times_.push_back(0);
  start_.push_back(std::chrono::steady_clock::now());
}

void TimeRecorder::event_end(const std::string& name) {
  auto now = std::chrono::steady_clock::now();
  for (size_t i = 0; i < events_.size(); i++)
    if (events_[i] == name) {
      times_[i] = std::chrono::duration<double, std::milli>(now - start_[i]).count();
      return;
    }
  printf("warning: no event %s\n", name.c_str());
}

void TimeRecorder::add(const std::string& name, double ms) {
  events_.push_back(name);
  times_.push_back(ms);
  start_.push_back(std::chrono::steady_clock::now());
}

void TimeRecorder::print_all_events() const {
  for (size_t i = 0; i < events_.size(); i++) printf("time cost(%s): %gms\n", events_[i].c_str(), times_[i]);
}

void query_to_csr(const UndirectedPlainGraph& q, std::vector<uint32_t>& off, std::vector<uint32_t>& adj) {
  const Vertex k = q.num_vertices_;
  off.assign((size_t)k + 1, 0);
  adj.clear();
  for (Vertex u = 0; u < k; u++) {
    std::vector<Vertex> nb = u < q.edges_.size() ? q.edges_[u] : std::vector<Vertex>();
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    nb.erase(std::remove(nb.begin(), nb.end(), u), nb.end());
    adj.insert(adj.end(), nb.begin(), nb.end());
    off[u + 1] = (uint32_t)adj.size();
  }
}

GPUSubgraphMatcher::~GPUSubgraphMatcher() {
  if (dgraph_) fsig_graph_free(ctx_, dgraph_);
  if (ctx_) fsig_ctx_destroy(ctx_);
}

int GPUSubgraphMatcher::init_GPU(int dev_id, uint64_t budget) {
  dev_id_ = dev_id;
  int rc = fsig_ctx_create(dev_id, budget, &ctx_);
  if (rc) {
    error_ = "Error: cannot open device " + std::to_string(dev_id);
    return rc;
  }
  fsig_set_profiling(ctx_, 1);
  return 0;
}

int GPUSubgraphMatcher::run(const fsig_graph* g, int alg, const uint32_t* qsig) {
  std::vector<uint32_t> qoff, qadj;
  query_to_csr(*query_graph_, qoff, qadj);
  const uint32_t k = query_graph_->num_vertices_;
  embeddings_.clear();
  const bool copy = mode_ == FSIG_MODE_ENUMERATE && host_rows_cap_ > 0;
  if (copy) {
    // size the host buffer from an exact count first
    uint64_t total = 0;
    int rc = fsig_match(ctx_, g, k, query_graph_->labels_.data(), qoff.data(), qadj.data(), qsig, alg,
                        FSIG_MODE_COUNT, 0, &total, nullptr, 0);
    if (rc) { error_ = fsig_last_error(ctx_); return rc; }
    if (max_embeddings_) total = std::min(total, max_embeddings_);
    embeddings_.resize((size_t)std::min(total, host_rows_cap_) * k);
  }
  fsig_stats_reset(ctx_);
  int rc = fsig_match(ctx_, g, k, query_graph_->labels_.data(), qoff.data(), qadj.data(), qsig, alg, mode_,
                      max_embeddings_, &num_embeddings_, copy ? embeddings_.data() : nullptr,
                      copy ? embeddings_.size() / std::max<uint32_t>(k, 1) : 0);
  if (rc) { error_ = fsig_last_error(ctx_); return rc; }
  if (copy) embeddings_.resize((size_t)std::min<uint64_t>(num_embeddings_, embeddings_.size() / std::max<uint32_t>(k, 1)) * k);
  if (time_recorder_) {
    static const char* names[FSIG_K_NUM] = {"filter kernels", "refine kernels", "candidate edge kernels",
                                            "join count kernels", "join write kernels", "scan kernels",
                                            "batch kernels"};
    for (int f = 0; f < FSIG_K_NUM; f++) {
      uint64_t launches;
      double ms;
      fsig_kernel_stats(ctx_, f, &launches, &ms, nullptr, nullptr);
      if (launches) time_recorder_->add(names[f], ms);
    }
  }
  return 0;
}

MMatch::MMatch(const UndirectedPlainGraph* query, const UndirectedCSRGraph* data,
               const InvertedIndex<Label, Vertex>* inv, int dev_id)
    : GPUSubgraphMatcher(query, data), csr_(data), inv_(inv) {
  dev_id_ = dev_id;
}

int MMatch::match() {
  TimeRecorder local;
  TimeRecorder* tr = time_recorder_ ? time_recorder_ : &local;
  time_recorder_ = tr;
  int rc;
  if (!ctx_ && (rc = init_GPU(dev_id_))) return rc;
  tr->event_start("copy graph");
  rc = fsig_graph_upload_csr(ctx_, csr_->num_vertices_, csr_->offset_.data(), csr_->edges_.data(),
                             csr_->labels_.data(), (uint32_t)std::max<Label>(csr_->num_labels_, 1),
                             inv_->index_.data(), inv_->list_.data(), &dgraph_);
  tr->event_end("copy graph");
  if (rc) { error_ = fsig_last_error(ctx_); return rc; }
  tr->event_start("total time");
  rc = run(dgraph_, FSIG_ALG_MMATCH, nullptr);
  tr->event_end("total time");
  if (rc) return rc;
  if (verbose_) printf("#embedding: %llu\n", (unsigned long long)num_embeddings_);
  return 0;
}

GSI::GSI(const UndirectedPlainGraph* query, const UndirectedPCSRGraph* data, const InvertedIndex<Label, Vertex>* inv,
         const uint32_t* qsig, const uint32_t* dsig, int dev_id)
    : GPUSubgraphMatcher(query, data), pcsr_(data), inv_(inv), qsig_(qsig), dsig_(dsig) {
  dev_id_ = dev_id;
}

int GSI::match() {
  TimeRecorder local;
  TimeRecorder* tr = time_recorder_ ? time_recorder_ : &local;
  time_recorder_ = tr;
  int rc;
  if (!ctx_ && (rc = init_GPU(dev_id_))) return rc;
  tr->event_start("copy graph");
  // Alg 2 reads no CSR adjacency: the graph is registered with empty lists
  // and the PCSR units plus signatures carry the structure.
  const Vertex n = pcsr_->num_vertices_;
  std::vector<uint32_t> zero_off((size_t)n + 1, 0);
  rc = fsig_graph_upload_csr(ctx_, n, zero_off.data(), nullptr, pcsr_->labels_.data(),
                             (uint32_t)std::max<Label>(pcsr_->num_labels_, 1), inv_->index_.data(),
                             inv_->list_.data(), &dgraph_);
  if (rc == 0) {
    const uint32_t U = (uint32_t)pcsr_->units_.size();
    std::vector<uint32_t> keys(U), nedges(U);
    std::vector<const uint32_t*> offs(U), edges(U);
    for (uint32_t u = 0; u < U; u++) {
      keys[u] = pcsr_->units_[u].num_keys_;
      nedges[u] = pcsr_->units_[u].num_edges_;
      offs[u] = pcsr_->units_[u].offset_.data();
      edges[u] = pcsr_->units_[u].edges_.data();
    }
    rc = fsig_graph_upload_pcsr(ctx_, dgraph_, U, keys.data(), offs.data(), nedges.data(), edges.data(), dsig_,
                                pcsr_->reference_unit_ids_ ? 1 : 0);
  }
  tr->event_end("copy graph");
  if (rc) { error_ = fsig_last_error(ctx_); return rc; }
  tr->event_start("total time");
  rc = run(dgraph_, FSIG_ALG_GSI, qsig_);
  tr->event_end("total time");
  if (rc) return rc;
  if (verbose_) printf("#embeddings: %llu\n", (unsigned long long)num_embeddings_);
  return 0;
}

}  // namespace fsig
