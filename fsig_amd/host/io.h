// Text IO of the FSIG `t / v / e` graph format (README.md:28-35) and the
// embedding output file.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "graph.h"

namespace fsig {

enum IoStatus {
  kIoOk = 0,
  kIoCannotOpen = -1,      // "cannot open file" (IO/graph_reader.cc:16-19)
  kIoBadHeader = -2,       // first token is not 't' (IO/graph_reader.cc:21-29)
  kIoBadRecord = -3,       // truncated or non-numeric record
  kIoMissingVertex = -4,   // an id below the vertex count has no `v` line
  kIoEdgeOutOfRange = -5,  // an edge endpoint has no `v` line
  kIoWriteFailed = -6,     // an output file could not be written in full
};

const char* io_status_string(int status);

// Reads the survey format into an adjacency-list graph, the role of
// UndirectedSurveyPlainReader::read (IO/graph_reader.cc:12-55):
//  * `t n m` header, its counts are not trusted;
//  * `v id label deg`: deg ignored, label forced to 0 when !is_labeled;
//  * `e u v`: one undirected edge, both directions stored.
// Unlike the reference, a trailing newline adds no phantom edge, and every id
// must have a `v` line (errors come back as IoStatus, never exit()).
int read_plain_graph(const std::string& path, bool is_labeled, UndirectedPlainGraph& out);

// Writes the survey format: every vertex gets a `v` line (degree = number of
// stored neighbours), then every undirected edge u<v once, with no newline
// after the last record.
int write_graph(const std::string& path, Vertex n, const Label* labels,
                const Edge* offset, const Vertex* adj);

// Canonical embedding file: first line `#embedding: <count>`, then one line
// per embedding, k vertex ids in query-vertex order, lines sorted
// lexicographically. rows is n_rows x k, row-major, sorted in place.
int write_embeddings(const std::string& path, uint32_t k, uint64_t count,
                     std::vector<uint32_t>& rows);

}  // namespace fsig
