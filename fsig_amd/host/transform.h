// Storage transforms: plain -> CSR, plain -> PCSR, neighbour-label signatures
// and MurmurHash2. Interface follows graph_format_transformer/
// graph_format_transformer.h:37-95.
#pragma once

#include <cstdint>
#include <vector>

#include "graph.h"

namespace fsig {

// MurmurHash2, 32-bit, of `len` bytes (Austin Appleby's published, public
// domain algorithm: multiplier 0x5bd1e995, shift 24); the reference restates it
// at graph_format_transformer/graph_format_transformer.h:59-93.
uint32_t murmur_hash2(const void* key, int len, uint32_t seed);

class GraphFormatTransformer {
 public:
  virtual ~GraphFormatTransformer() = default;
  // Returns 0 on success, negative on a wrong input type.
  virtual int transform(const Graph* input, Graph*& output) = 0;
};

// graph_format_transformer.cc:12-37: sort each adjacency list ascending and
// pack offsets/edges.  When `dedupe` is set, repeated neighbours and
// self-loops are dropped (the matcher needs simple adjacency; the reference
// keeps them and its join would then repeat embeddings).
class UndirectedPlainToCSRTransformer : public GraphFormatTransformer {
 public:
  explicit UndirectedPlainToCSRTransformer(bool dedupe = true) : dedupe_(dedupe) {}
  int transform(const Graph* input, Graph*& output) override;
  bool dedupe_;
};

// graph_format_transformer.cc:39-215: one unit per unordered label pair,
// keys = vertices with a neighbour of the partner label, MurmurHash2 buckets
// of 15 (key, start) slots per 32-word group, overflow chained into empty
// groups in FIFO order, neighbour lists ascending.  Built with index maps,
// O(sum of degrees), instead of the reference's per-key linear find.
class UndirectedPlainToPCSRTransformer : public GraphFormatTransformer {
 public:
  explicit UndirectedPlainToPCSRTransformer(bool reference_unit_ids = false)
      : reference_unit_ids_(reference_unit_ids) {}
  int transform(const Graph* input, Graph*& output) override;
  bool reference_unit_ids_;

  // graph_format_transformer.cc:217-281: 16 words per vertex, word 0 = label,
  // words 1..15 hold 240 two-bit saturating counters (0 -> 1 -> 3) at slot
  // MurmurHash2(label(nbr)) % 240.  The reference shifts by 2*(2h mod 32),
  // which x86 reduces mod 32; `x86_shift` reproduces that layout exactly, and
  // false uses the intended shift 2h mod 32.  column_oriented gives [16][n]
  // (data graph), otherwise [n][16] (query graph).
  static std::vector<uint32_t> build_signature(bool column_oriented, const UndirectedPlainGraph& g,
                                               bool x86_shift = true);
};

// Packs an edge list (both orders or one order; each pair listed once or
// more) into sorted CSR with duplicates and self-loops removed.
void csr_from_edge_list(uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>>& edges,
                        std::vector<uint32_t>& offset, std::vector<uint32_t>& adj, int threads = 0);

// Largest label count a PCSR is built for: one unit per label pair, so 4096
// labels already make 8.4M units.
constexpr int32_t kPcsrMaxLabels = 4096;

// PCSR built directly from CSR arrays (same layout as the transformer).
void pcsr_from_csr(uint32_t n, const uint32_t* offset, const uint32_t* adj, const int32_t* labels,
                   int32_t num_labels, bool reference_unit_ids, std::vector<PCSRUnit>& units);

// Signature built directly from CSR arrays.
std::vector<uint32_t> signature_from_csr(uint32_t n, const uint32_t* offset, const uint32_t* adj,
                                         const int32_t* labels, bool column_oriented,
                                         bool x86_shift = true);

}  // namespace fsig
