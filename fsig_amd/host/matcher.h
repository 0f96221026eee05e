// The matcher interface of the FSIG drop-in: GPUSubgraphMatcher with its MMatch
// (alg 1) and GSI (alg 2) subclasses, as constructed by main()
// (match/subgraph_match.h:39-81, match/MMatch.h:155-176, match/GSI.h:67-91).
// The GPU work happens behind the C ABI of include/fsig.h; these classes own
// the device context and the resident graph and report the reference's
// stdout lines.  Errors come back as negative status codes (the reference
// prints and calls exit(0)).
#pragma once

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fsig.h"
#include "graph.h"

namespace fsig {

// utility/time_recorder.h:26-57: named wall-clock intervals, printed as
// `time cost(<event>): <ms>ms`.
class TimeRecorder {
 public:
  void event_start(const std::string& name);
  void event_end(const std::string& name);
  void add(const std::string& name, double ms);
  void print_all_events() const;
 private:
  std::vector<std::string> events_;
  std::vector<double> times_;
  std::vector<std::chrono::steady_clock::time_point> start_;
};

class GPUSubgraphMatcher {
 public:
  GPUSubgraphMatcher(const UndirectedPlainGraph* query, const Graph* data) : query_graph_(query), data_graph_(data) {}
  virtual ~GPUSubgraphMatcher();
  // Runs the match; 0 or a negative FSIG_E* code.
  virtual int match() = 0;
  // Opens the device and reserves the arena (match/subgraph_match.h:51-80).
  int init_GPU(int dev_id, uint64_t hbm_budget = 0);

  // Options set by the driver before match().
  int mode_ = FSIG_MODE_ENUMERATE;
  uint64_t max_embeddings_ = 0;      // ENUMERATE bound, 0 = none
  uint64_t host_rows_cap_ = 0;       // rows copied to embeddings_, 0 = none
  bool verbose_ = true;

  // Results.
  uint64_t num_embeddings_ = 0;
  std::vector<uint32_t> embeddings_;  // host_rows_cap_ rows at most, k words each
  std::string error_;
  TimeRecorder* time_recorder_ = nullptr;

 protected:
  int run(const fsig_graph* g, int alg, const uint32_t* qsig);
  const UndirectedPlainGraph* query_graph_;
  const Graph* data_graph_;
  fsig_ctx* ctx_ = nullptr;
  fsig_graph* dgraph_ = nullptr;
  int dev_id_ = 0;
};

class MMatch : public GPUSubgraphMatcher {
 public:
  MMatch(const UndirectedPlainGraph* query, const UndirectedCSRGraph* data,
         const InvertedIndex<Label, Vertex>* inverted_data_label, int dev_id = 0);
  int match() override;
 private:
  const UndirectedCSRGraph* csr_;
  const InvertedIndex<Label, Vertex>* inv_;
};

class GSI : public GPUSubgraphMatcher {
 public:
  GSI(const UndirectedPlainGraph* query, const UndirectedPCSRGraph* data,
      const InvertedIndex<Label, Vertex>* inverted_data_label, const uint32_t* query_signature_table,
      const uint32_t* data_signature_table, int dev_id = 0);
  int match() override;
 private:
  const UndirectedPCSRGraph* pcsr_;
  const InvertedIndex<Label, Vertex>* inv_;
  const uint32_t* qsig_;
  const uint32_t* dsig_;
};

// Query graph in the CSR form the C ABI takes (duplicates and self-loops
// dropped, lists ascending).
void query_to_csr(const UndirectedPlainGraph& q, std::vector<uint32_t>& off, std::vector<uint32_t>& adj);

}  // namespace fsig
