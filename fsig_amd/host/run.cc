// FSIG driver: `FSIG <data> <query> [alg=1] [labeled=1] [out] [dev=0] [max]`.
// Same positional arguments as main/run.cc:27-64; argv[5], which the
// reference accepts but never writes, receives the canonical embedding file
// (`#embedding: <count>` then the sorted rows, query-vertex order) when given;
// argv[7] bounds the enumeration (0 = all).  A data graph named *.csr.bin is
// read from the binary CSR cache instead of the text format.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "../../include/fsig_host.h"
#include "graph.h"
#include "io.h"
#include "matcher.h"
#include "transform.h"

using namespace fsig;

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    printf("invalid arguments!\n");
    printf("usage: FSIG <data graph> <query graph> [alg 1|2] [labeled 0|1] [out file] [device] [max embeddings]\n");
    return -1;
  }
  const std::string data_file = argv[1], query_file = argv[2];
  const int alg = argc >= 4 ? atoi(argv[3]) : 1;
  const bool labeled = argc >= 5 ? atoi(argv[4]) != 0 : true;
  const std::string out_file = argc >= 6 ? argv[5] : "";
  const int dev = argc >= 7 ? atoi(argv[6]) : 0;
  const uint64_t max_emb = argc >= 8 ? strtoull(argv[7], nullptr, 10) : 0;
  if (alg != 1 && alg != 2) {
    printf("invalid arguments! algorithm type must be 1 (FSIG) or 2 (GSI)\n");
    return -1;
  }
  TimeRecorder tr;
  UndirectedPlainGraph data, query;
  // a data graph ending in .csr.bin is the binary CSR cache (alg 1 only: the
  // GSI path builds its signatures from the plain graph)
  const std::string bin_suffix = ".csr.bin";
  const bool from_bin = data_file.size() > bin_suffix.size() &&
                        data_file.compare(data_file.size() - bin_suffix.size(), bin_suffix.size(), bin_suffix) == 0;
  if (from_bin && alg != 1) {
    printf("invalid arguments! a binary CSR cache serves algorithm 1 only\n");
    return -1;
  }
  std::unique_ptr<UndirectedCSRGraph> cached;
  int rc = 0;
  tr.event_start("read data");
  if (from_bin) {
    fsig_hgraph* hg = fsig_hg_load_bin(data_file.c_str(), &rc);
    if (hg) {
      uint32_t n = 0, md = 0;
      uint64_t nnz = 0;
      int32_t nl = 0;
      fsig_hg_info(hg, &n, &nnz, &nl, &md);
      cached.reset(new UndirectedCSRGraph());
      cached->num_vertices_ = n;
      cached->num_edges_ = (Edge)(nnz / 2);
      cached->num_labels_ = nl;
      cached->offset_.assign(fsig_hg_off(hg), fsig_hg_off(hg) + n + 1);
      cached->edges_.assign(fsig_hg_adj(hg), fsig_hg_adj(hg) + nnz);
      cached->labels_.assign(fsig_hg_labels(hg), fsig_hg_labels(hg) + n);
      fsig_hg_free(hg);
    }
  } else {
    rc = read_plain_graph(data_file, labeled, data);
  }
  tr.event_end("read data");
  if (rc) { printf("%s\n", io_status_string(rc)); return 1; }
  tr.event_start("read query");
  rc = read_plain_graph(query_file, labeled, query);
  tr.event_end("read query");
  if (rc) { printf("%s\n", io_status_string(rc)); return 1; }
  if (cached)
    printf("|V|: %u\n|E|: %u\n|L|: %d\n", cached->num_vertices_, cached->num_edges_, cached->num_labels_);
  else
    printf("|V|: %u\n|E|: %u\n|L|: %d\n", data.num_vertices_, data.num_edges_, data.num_labels_);
  printf("|V|: %u\n|E|: %u\n|L|: %d\n", query.num_vertices_, query.num_edges_, query.num_labels_);

  Graph* organised = nullptr;
  tr.event_start("organise graph");
  std::unique_ptr<GraphFormatTransformer> t;
  if (alg == 1) t.reset(new UndirectedPlainToCSRTransformer(true));
  else t.reset(new UndirectedPlainToPCSRTransformer(false));
  if (cached) organised = cached.release();
  else rc = t->transform(&data, organised);
  tr.event_end("organise graph");
  if (rc || !organised) { printf("Error: transform plain graph\n"); return 1; }
  std::unique_ptr<Graph> owned(organised);
  tr.print_all_events();

  TimeRecorder mtr;
  std::unique_ptr<GPUSubgraphMatcher> matcher;
  std::vector<uint32_t> qsig, dsig;
  std::unique_ptr<InvertedIndex<Label, Vertex>> inv;
  if (alg == 1) {
    auto* csr = static_cast<UndirectedCSRGraph*>(organised);
    inv.reset(new InvertedIndex<Label, Vertex>(csr->num_vertices_, std::max<Label>(csr->num_labels_, 1), csr->labels_.data()));
    matcher.reset(new MMatch(&query, csr, inv.get(), dev));
  } else {
    auto* pcsr = static_cast<UndirectedPCSRGraph*>(organised);
    inv.reset(new InvertedIndex<Label, Vertex>(pcsr->num_vertices_, std::max<Label>(pcsr->num_labels_, 1), pcsr->labels_.data()));
    qsig = UndirectedPlainToPCSRTransformer::build_signature(false, query);
    dsig = UndirectedPlainToPCSRTransformer::build_signature(true, data);
    matcher.reset(new GSI(&query, pcsr, inv.get(), qsig.data(), dsig.data(), dev));
  }
  matcher->time_recorder_ = &mtr;
  matcher->max_embeddings_ = max_emb;
  if (!out_file.empty()) matcher->host_rows_cap_ = max_emb ? max_emb : 100000000ull;
  rc = matcher->match();
  if (rc) {
    printf("%s (status %d)\n", matcher->error_.c_str(), rc);
    return 1;
  }
  mtr.print_all_events();
  if (!out_file.empty()) {
    rc = write_embeddings(out_file, query.num_vertices_, matcher->num_embeddings_, matcher->embeddings_);
    if (rc) { printf("%s\n", io_status_string(rc)); return 1; }
  }
  return 0;
}
