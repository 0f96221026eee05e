// Seeded synthetic inputs: Chung-Lu and R-MAT data graphs, uniform and Zipf
// labels, and queries taken as induced subgraphs of random walks.  All
// randomness comes from a project-local xoshiro256** stream so the same seed
// gives the same graph on every machine.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace fsig {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed);
  uint64_t next();
  double uniform();                       // [0, 1)
  uint32_t below(uint32_t bound);         // [0, bound)
};

typedef std::vector<std::pair<uint32_t, uint32_t>> EdgeList;

// Chung-Lu: expected degrees d_i proportional to (i+1)^(-1/(gamma-1)),
// scaled so that sum d = 2m and capped at sqrt(2m) (the admissibility bound
// d_max^2 <= sum d); endpoints drawn independently proportional to d, the
// first m distinct non-loop pairs kept; vertex ids shuffled.
EdgeList gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint64_t seed);

// R-MAT with quadrant probabilities (a, b, c, 1-a-b-c) on 2^scale vertices,
// the first m distinct non-loop pairs kept, vertex ids shuffled.
EdgeList gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed,
                  int threads = 0);

// Renames vertices so that vertex n-1 is not isolated (a file read by the
// reference sizes its adjacency by the largest endpoint).
void ensure_last_vertex_has_edge(uint32_t n, EdgeList& edges);

std::vector<int32_t> labels_uniform(uint32_t n, int32_t num_labels, uint64_t seed);
// P(label = i) proportional to 1/(i+1)^s, i < num_labels.
std::vector<int32_t> labels_zipf(uint32_t n, int32_t num_labels, double s, uint64_t seed);

struct WalkQuery {
  uint32_t k = 0;
  std::vector<int32_t> labels;                          // [k]
  std::vector<std::pair<uint32_t, uint32_t>> edges;     // u < v, query ids
  std::vector<uint32_t> data_vertices;                  // walk vertex of query id i
};

// Random walk from a random non-isolated vertex until k distinct vertices are
// visited; the query is their induced subgraph with their labels, ids given
// in visit order.  Walks are retried until the edge count lies in
// [min_edges, max_edges].  bias > 0 steps to the highest-degree of
// (1 + bias) uniformly drawn neighbours (reaches the denser core); bias < 0
// to the lowest-degree of (1 - bias) (stays in the periphery).
// Returns 0, or -1 when max_attempts walks all failed.
int random_walk_query(uint32_t n, const uint32_t* off, const uint32_t* adj, const int32_t* labels,
                      uint32_t k, uint32_t min_edges, uint32_t max_edges, uint64_t seed, int bias,
                      uint32_t max_attempts, WalkQuery& out,
                      const std::vector<uint32_t>* nonisolated = nullptr);

}  // namespace fsig
