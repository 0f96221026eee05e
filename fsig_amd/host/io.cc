// Survey-format reader/writer and the embedding file writer.
#include "io.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <thread>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace fsig {

const char* io_status_string(int s) {
  switch (s) {
    case kIoOk: return "ok";
    case kIoCannotOpen: return "Error: cannot open file; undirected survey file";
    case kIoBadHeader: return "Error: graph file format is not correct; undirected survey file";
    case kIoBadRecord: return "Error: truncated or malformed record; undirected survey file";
    case kIoMissingVertex: return "Error: a vertex id has no v line; undirected survey file";
    case kIoEdgeOutOfRange: return "Error: an edge endpoint has no v line; undirected survey file";
    case kIoWriteFailed: return "Error: the output file could not be written";
  }
  return "Error: unknown";
}

namespace {

struct Cursor {
  const char* p;
  const char* end;
  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
  }
  void skip_line() {
    while (p < end && *p != '\n') p++;
  }
  bool next_char(char& c) {
    skip_ws();
    if (p >= end) return false;
    c = *p++;
    return true;
  }
  bool next_u64(uint64_t& v) {
    skip_ws();
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
      const uint64_t d = (uint64_t)(*p++ - '0');
      if (x > (UINT64_MAX - d) / 10) return false;  // a number past 2^64 - 1 is a malformed record
      x = x * 10 + d;
    }
    v = x;
    return true;
  }
  bool next_i64(int64_t& v) {
    skip_ws();
    bool neg = false;
    if (p < end && *p == '-') { neg = true; p++; }
    uint64_t u;
    if (!next_u64(u)) return false;
    v = neg ? -(int64_t)u : (int64_t)u;
    return true;
  }
};

// Read-only mapping of a whole file: the parser threads page in their own
// slices in parallel instead of one thread copying the file first.
struct MappedFile {
  const char* data = nullptr;
  size_t size = 0;
  bool ok = false;
  explicit MappedFile(const std::string& path) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return;
    struct stat st;
    if (::fstat(fd, &st) == 0) {
      size = (size_t)st.st_size;
      if (size == 0) {
        ok = true;
      } else {
        void* p = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p != MAP_FAILED) {
          data = (const char*)p;
          ok = true;
        }
      }
    }
    ::close(fd);
  }
  ~MappedFile() {
    if (data) ::munmap((void*)data, size);
  }
  MappedFile(const MappedFile&) = delete;
  MappedFile& operator=(const MappedFile&) = delete;
};

}  // namespace

namespace {

// Records of one slice of the file body, in file order.
struct ParsedSlice {
  std::vector<std::pair<Vertex, Label>> vertices;
  std::vector<std::pair<Vertex, Vertex>> edges;
  int rc = kIoOk;
};

void parse_slice(const char* b, const char* e, bool is_labeled, ParsedSlice& out) {
  Cursor c{b, e};
  char type;
  while (c.next_char(type)) {
    if (type == 'v') {
      uint64_t id, deg;
      int64_t label;
      if (!c.next_u64(id) || !c.next_i64(label) || !c.next_u64(deg)) { out.rc = kIoBadRecord; return; }
      if (id >= kInvalid || label < 0 || label > 0x7fffffff) { out.rc = kIoBadRecord; return; }
      out.vertices.emplace_back((Vertex)id, is_labeled ? (Label)label : 0);
    } else if (type == 'e') {
      uint64_t a, bb;
      if (!c.next_u64(a) || !c.next_u64(bb)) { out.rc = kIoBadRecord; return; }
      if (a >= kInvalid || bb >= kInvalid) { out.rc = kIoBadRecord; return; }
      out.edges.emplace_back((Vertex)a, (Vertex)bb);
    } else {
      c.skip_line();  // unknown record kind: ignored, as the reference does
    }
  }
}

}  // namespace

// Files above 16 MiB are parsed by up to 16 threads: the body is cut at line
// starts of `v` / `e` records, every slice is parsed on its own and the
// slices are applied in file order, so the result and the first error
// reported are those of a sequential parse.
int read_plain_graph(const std::string& path, bool is_labeled, UndirectedPlainGraph& g) {
  MappedFile file(path);
  if (!file.ok) return kIoCannotOpen;
  Cursor c{file.data, file.data + file.size};
  char type;
  uint64_t hn, hm;
  if (!c.next_char(type) || type != 't') return kIoBadHeader;
  if (!c.next_u64(hn) || !c.next_u64(hm)) return kIoBadHeader;
  g = UndirectedPlainGraph();
  g.is_labeled_ = is_labeled;
  const char* body = c.p;
  const char* end = file.data + file.size;
  const size_t len = (size_t)(end - body);
  unsigned T = 1;
  if (len > (16u << 20)) T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<const char*> cut{body};
  for (unsigned t = 1; t < T; t++) {
    const char* q = std::max(cut.back(), body + len / T * t);
    while (q < end && !(q[0] == '\n' && q + 1 < end && (q[1] == 'v' || q[1] == 'e'))) q++;
    cut.push_back(q);
  }
  cut.push_back(end);
  std::vector<ParsedSlice> slices(T);
  if (T == 1) {
    parse_slice(body, end, is_labeled, slices[0]);
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back(parse_slice, cut[t], cut[t + 1], is_labeled, std::ref(slices[t]));
    for (auto& x : th) x.join();
  }
  for (auto& sl : slices)
    if (sl.rc) return sl.rc;
  // every id up to the largest needs a v line, so an id at or past the number
  // of v lines has a gap below it: reject it before sizing the vertex arrays
  uint64_t vlines = 0;
  for (auto& sl : slices) vlines += sl.vertices.size();
  for (auto& sl : slices)
    for (auto& v : sl.vertices)
      if (v.first >= vlines) return kIoMissingVertex;
  for (auto& sl : slices)
    for (auto& v : sl.vertices) g.add_vertex(v.first, v.second);
  for (Vertex v = 0; v < g.num_vertices_; v++)
    if (g.labels_[v] == kInvalidLabel) return kIoMissingVertex;
  for (auto& sl : slices)
    for (auto& e : sl.edges)
      if (e.first >= g.num_vertices_ || e.second >= g.num_vertices_) return kIoEdgeOutOfRange;
  // Pre-size adjacency from degrees so huge files do not re-grow vectors.
  const Vertex n = g.num_vertices_;
  std::vector<uint32_t> deg(n, 0);
  uint64_t nedges = 0;
  for (auto& sl : slices) nedges += sl.edges.size();
  g.edges_.assign(n, {});
  if (T == 1) {
    for (auto& sl : slices)
      for (auto& e : sl.edges) { deg[e.first]++; deg[e.second]++; }
    for (Vertex v = 0; v < n; v++) g.edges_[v].reserve(deg[v]);
  } else {
    // degrees: each slice's thread counts its own edges with relaxed atomic
    // increments; then the lists are reserved by vertex ranges in parallel
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back([&, t] {
        for (auto& e : slices[t].edges) {
          __atomic_fetch_add(&deg[e.first], 1u, __ATOMIC_RELAXED);
          __atomic_fetch_add(&deg[e.second], 1u, __ATOMIC_RELAXED);
        }
      });
    for (auto& x : th) x.join();
    th.clear();
    for (unsigned t = 0; t < T; t++)
      th.emplace_back([&, t] {
        const Vertex a = (Vertex)((uint64_t)n * t / T), b = (Vertex)((uint64_t)n * (t + 1) / T);
        for (Vertex v = a; v < b; v++) g.edges_[v].reserve(deg[v]);
      });
    for (auto& x : th) x.join();
  }
  if (T == 1) {
    for (auto& sl : slices)
      for (auto& e : sl.edges) g.add_edge(e.first, e.second);
  } else {
    // Each thread owns a vertex range of about equal adjacency and appends,
    // in file order, the entries of every edge with an endpoint in its range:
    // every list comes out exactly as a sequential add_edge loop builds it.
    std::vector<Vertex> lo(T + 1, n);
    lo[0] = 0;
    uint64_t acc = 0;
    unsigned t = 1;
    for (Vertex v = 0; v < n && t < T; v++) {
      acc += deg[v];
      while (t < T && acc >= 2 * nedges * t / T) lo[t++] = v + 1;
    }
    std::vector<std::thread> th;
    for (unsigned w = 0; w < T; w++)
      th.emplace_back([&, w] {
        const Vertex a0 = lo[w], a1 = lo[w + 1];
        for (auto& sl : slices)
          for (auto& e : sl.edges) {
            if (e.first >= a0 && e.first < a1) g.edges_[e.first].push_back(e.second);
            if (e.second >= a0 && e.second < a1) g.edges_[e.second].push_back(e.first);
          }
      });
    for (auto& x : th) x.join();
    g.num_edges_ += (Edge)nedges;
  }
  for (auto& sl : slices) std::vector<std::pair<Vertex, Vertex>>().swap(sl.edges);
  g.finalize();
  g.set_min_degree();
  return kIoOk;
}

int write_graph(const std::string& path, Vertex n, const Label* labels, const Edge* off,
                const Vertex* adj) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return kIoCannotOpen;
  uint64_t m = 0;
  for (Vertex v = 0; v < n; v++)
    for (Edge i = off[v]; i < off[v + 1]; i++)
      if (adj[i] > v) m++;
  std::vector<char> out;
  out.reserve(1 << 20);
  char line[96];
  bool short_write = false;
  auto flush = [&](bool force) {
    if (force || out.size() > (1 << 20)) {
      if (!out.empty() && fwrite(out.data(), 1, out.size(), f) != out.size()) short_write = true;
      out.clear();
    }
  };
  bool first = true;
  auto emit = [&](int len) {
    if (!first) out.push_back('\n');
    first = false;
    out.insert(out.end(), line, line + len);
    flush(false);
  };
  emit(snprintf(line, sizeof line, "t %u %llu", n, (unsigned long long)m));
  for (Vertex v = 0; v < n; v++)
    emit(snprintf(line, sizeof line, "v %u %d %u", v, labels[v], off[v + 1] - off[v]));
  for (Vertex v = 0; v < n; v++)
    for (Edge i = off[v]; i < off[v + 1]; i++)
      if (adj[i] > v) emit(snprintf(line, sizeof line, "e %u %u", v, adj[i]));
  flush(true);
  // a short write or a failed close (buffered data that never reached the
  // file) leaves a truncated graph: report it
  if (fclose(f) != 0) short_write = true;
  return short_write ? kIoWriteFailed : kIoOk;
}

// Rows are sorted lexicographically (an index sort: chunks on threads, then
// pairwise merges), gathered into sorted order, and printed block by block,
// each block formatted on threads and written in order.
int write_embeddings(const std::string& path, uint32_t k, uint64_t count, std::vector<uint32_t>& rows) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return kIoCannotOpen;
  const size_t n = k ? rows.size() / k : 0;
  const unsigned T = n > (1u << 18) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
  auto run = [&](unsigned parts, auto&& fn) {
    if (parts == 1) return fn(0u);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < parts; t++) th.emplace_back(fn, t);
    for (auto& x : th) x.join();
  };
  const uint32_t* R = rows.data();
  auto less = [&](size_t a, size_t b) {
    const uint32_t* x = R + a * k;
    const uint32_t* y = R + b * k;
    for (uint32_t j = 0; j < k; j++)
      if (x[j] != y[j]) return x[j] < y[j];
    return false;
  };
  std::vector<size_t> idx(n), tmp;
  for (size_t i = 0; i < n; i++) idx[i] = i;
  std::vector<size_t> cut(T + 1);
  for (unsigned t = 0; t <= T; t++) cut[t] = n * t / T;
  run(T, [&](unsigned t) { std::sort(idx.begin() + cut[t], idx.begin() + cut[t + 1], less); });
  if (T > 1) {
    tmp.resize(n);
    for (unsigned width = 1; width < T; width *= 2) {
      const unsigned pairs = (T + 2 * width - 1) / (2 * width);
      run(pairs, [&](unsigned pr) {
        const unsigned a = pr * 2 * width, m = std::min(T, a + width), b = std::min(T, a + 2 * width);
        std::merge(idx.begin() + cut[a], idx.begin() + cut[m], idx.begin() + cut[m], idx.begin() + cut[b],
                   tmp.begin() + cut[a], less);
      });
      idx.swap(tmp);
    }
  }
  std::vector<uint32_t> sorted(rows.size());
  run(T, [&](unsigned t) {
    for (size_t i = cut[t]; i < cut[t + 1]; i++) std::memcpy(&sorted[i * k], R + idx[i] * k, k * sizeof(uint32_t));
  });
  std::vector<size_t>().swap(idx);
  std::vector<size_t>().swap(tmp);
  fprintf(f, "#embedding: %llu\n", (unsigned long long)count);
  // decimal text, at most 11 bytes per number (10 digits and a separator)
  const size_t block = (size_t)1 << 20;
  std::vector<std::vector<char>> buf(T);
  bool ok = true;
  for (size_t b0 = 0; b0 < n && ok; b0 += block) {
    const size_t b1 = std::min(n, b0 + block);
    run(T, [&](unsigned t) {
      const size_t r0 = b0 + (b1 - b0) * t / T, r1 = b0 + (b1 - b0) * (t + 1) / T;
      std::vector<char>& out = buf[t];
      out.resize((r1 - r0) * k * 11);
      char* o = out.data();
      for (size_t i = r0; i < r1; i++) {
        for (uint32_t j = 0; j < k; j++) {
          uint32_t v = sorted[i * k + j];
          char d[10];
          int nd = 0;
          do {
            d[nd++] = (char)('0' + v % 10);
            v /= 10;
          } while (v);
          while (nd) *o++ = d[--nd];
          *o++ = j + 1 < k ? ' ' : '\n';
        }
      }
      out.resize((size_t)(o - out.data()));
    });
    for (unsigned t = 0; t < T && ok; t++)
      ok = buf[t].empty() || fwrite(buf[t].data(), 1, buf[t].size(), f) == buf[t].size();
  }
  rows.swap(sorted);
  if (fclose(f) != 0) ok = false;
  return ok ? kIoOk : kIoWriteFailed;
}

}  // namespace fsig
