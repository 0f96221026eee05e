// Host-side graph model of the FSIG drop-in.
//
// Mirrors the reference's data classes so that a caller of FSIG finds the
// same names and meanings:
//   Graph / UndirectedPlainGraph / UndirectedCSRGraph   graph/graph.h:27-85
//   PCSRUnit / UndirectedPCSRGraph                       graph/graph.h:90-123
//   InvertedIndex<Label, Vertex>                         utility/inverted_index.h:16-44
//   scalar types and sentinels                           utility/defines.h:16-31
// The storage is std::vector owned by the object (the reference uses raw new[]
// and never frees); the field meanings and array layouts are the same.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fsig {

typedef uint32_t Vertex;  // utility/defines.h:16
typedef uint32_t Edge;    // utility/defines.h:17
typedef int32_t Label;    // utility/defines.h:18

constexpr uint32_t kInvalid = 0xffffffffu;  // INVALID, utility/defines.h:28
constexpr uint32_t kHashSeed = 17;          // HASHSEED, utility/defines.h:29
constexpr uint32_t kSigBits = 512;          // SIGLEN, utility/defines.h:30
constexpr uint32_t kSigWords = 16;          // SIGNUM, utility/defines.h:31
constexpr Label kInvalidLabel = -1;         // INVALID_LABEL read as int, utility/defines.h:22

class Graph {
 public:
  virtual ~Graph() = default;
  Vertex num_vertices() const { return num_vertices_; }
  Edge num_edges() const { return num_edges_; }
  Label num_labels() const { return num_labels_; }
  bool is_directed() const { return false; }

  Vertex num_vertices_ = 0;
  Edge num_edges_ = 0;   // undirected edges, each stored twice in adjacency
  Label num_labels_ = 0; // max label + 1
  bool is_labeled_ = true;
};

// Adjacency-list graph as produced by the text reader (graph/graph.h:50-69).
class UndirectedPlainGraph : public Graph {
 public:
  std::vector<Label> labels_;
  std::vector<std::vector<Vertex>> edges_;
  Vertex min_degree = 0;

  void add_vertex(Vertex id, Label label);
  void add_edge(Vertex a, Vertex b);
  Vertex degree(Vertex v) const { return v < edges_.size() ? (Vertex)edges_[v].size() : 0; }
  const std::vector<Vertex>& neighbors(Vertex v) const { return edges_[v]; }
  bool check_edge_existence(Vertex a, Vertex b) const;
  void set_min_degree();
  // Sizes the adjacency to num_vertices_ (the reference sizes it by the largest
  // edge endpoint, graph/graph.cc:23-33, which leaves an isolated max id out).
  void finalize();
};

// CSR: offset_[n+1], edges_[2m] ascending per vertex, labels_[n]
// (graph/graph.h:74-85, built by graph_format_transformer.cc:12-37).
class UndirectedCSRGraph : public Graph {
 public:
  std::vector<Label> labels_;
  std::vector<Edge> offset_;
  std::vector<Vertex> edges_;
  Vertex min_degree = 0;
  Vertex degree(Vertex v) const { return offset_[v + 1] - offset_[v]; }
  const Vertex* neighbors(Vertex v) const { return edges_.data() + offset_[v]; }
};

// One label-pair partition of the PCSR (graph/graph.h:90-95).
// offset_ holds num_keys_ groups of 32 words: words 2s / 2s+1 are (key, start)
// for slot s < 15, the start of slot s+1 doubles as the end of slot s, word 30
// is the overflow group id (kInvalid if none) and word 31 is the end of slot 14.
struct PCSRUnit {
  std::vector<Edge> offset_;
  std::vector<Vertex> edges_;
  Vertex num_keys_ = 0;
  Vertex num_edges_ = 0;
};

class UndirectedPCSRGraph : public Graph {
 public:
  std::vector<Label> labels_;
  std::vector<PCSRUnit> units_;
  Vertex num_units_ = 0;
  // true: the reference's unit numbering (graph/graph.h:107-122), which maps
  // (l+1, 0) and (l, l) to one unit; false: the injective l(l+1)/2 + s.
  bool reference_unit_ids_ = false;

  static Label reference_edge_label(Label a, Label b);
  static Label injective_edge_label(Label a, Label b);
  Label get_edge_label(Label a, Label b) const {
    return reference_unit_ids_ ? reference_edge_label(a, b) : injective_edge_label(a, b);
  }
};

// Label -> vertices grouped by label, ascending within each label
// (utility/inverted_index.h:16-44).
template <class T, class N>
class InvertedIndex {
 public:
  N length_ = 0;
  T num_keys_ = 0;
  std::vector<N> index_;  // [num_keys + 1]
  std::vector<N> list_;   // [length]

  InvertedIndex(N length, T num_keys, const T* forward) : length_(length), num_keys_(num_keys) {
    index_.assign((size_t)num_keys + 1, 0);
    for (N i = 0; i < length; i++) index_[(size_t)forward[i] + 1]++;
    for (T k = 0; k < num_keys; k++) index_[(size_t)k + 1] += index_[(size_t)k];
    list_.resize(length);
    std::vector<N> cursor(index_.begin(), index_.end() - 1);
    for (N i = 0; i < length; i++) list_[cursor[(size_t)forward[i]]++] = i;
  }
  N num_values(T key) const { return index_[(size_t)key + 1] - index_[(size_t)key]; }
};

}  // namespace fsig
