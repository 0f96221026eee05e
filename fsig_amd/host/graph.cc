// Host graph model: adjacency-list graph and PCSR unit numbering.
// Behaviour follows graph/graph.cc:12-67 and graph/graph.h:107-122.
#include "graph.h"

#include <algorithm>

namespace fsig {

// graph/graph.cc:12-21: grow the label array to the id, raise num_labels.
void UndirectedPlainGraph::add_vertex(Vertex id, Label label) {
  if (labels_.size() <= id) {
    labels_.resize((size_t)id + 1, kInvalidLabel);
    num_vertices_ = id + 1;
  }
  if (num_labels_ <= label) num_labels_ = label + 1;
  labels_[id] = label;
}

// graph/graph.cc:23-33: store both directions, count one undirected edge.
void UndirectedPlainGraph::add_edge(Vertex a, Vertex b) {
  Vertex hi = std::max(a, b);
  if (edges_.size() <= hi) edges_.resize((size_t)hi + 1);
  edges_[a].push_back(b);
  edges_[b].push_back(a);
  num_edges_++;
}

bool UndirectedPlainGraph::check_edge_existence(Vertex a, Vertex b) const {
  if (a >= edges_.size()) return false;
  const auto& n = edges_[a];
  return std::find(n.begin(), n.end(), b) != n.end();
}

void UndirectedPlainGraph::finalize() {
  if (edges_.size() < num_vertices_) edges_.resize(num_vertices_);
}

// graph/graph.cc:60-67
void UndirectedPlainGraph::set_min_degree() {
  min_degree = num_vertices_;
  for (Vertex v = 0; v < num_vertices_; v++) min_degree = std::min(min_degree, degree(v));
}

// graph/graph.h:107-122 verbatim in meaning: larger label L, smaller s,
// unit = s if L == 0 else L(L-1)/2 + s.  Not injective: (l+1, 0) and (l, l)
// share a unit.
Label UndirectedPCSRGraph::reference_edge_label(Label a, Label b) {
  Label hi = std::max(a, b), lo = std::min(a, b);
  return hi == 0 ? lo : (hi - 1) * hi / 2 + lo;
}

// Injective numbering of unordered label pairs: L(L+1)/2 + s, s <= L.
// It spans the same L(L+1)/2 units.
Label UndirectedPCSRGraph::injective_edge_label(Label a, Label b) {
  Label hi = std::max(a, b), lo = std::min(a, b);
  return hi * (hi + 1) / 2 + lo;
}

}  // namespace fsig
