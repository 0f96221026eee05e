// Internal structures of libfsig_hip.so shared by its translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/fsig.h"

#define MAXQ 32
constexpr size_t kPinnedWords = 65536;
#define FSIG_INVALID 0xffffffffu

// ---------------------------------------------------------------------------
// Device-memory arena: one hipMalloc per context, bump allocation with
// mark/release.  The matcher is a depth-first walk over levels and chunks, so
// every per-query allocation is released in LIFO order.  This is the role of
// memory_manager/memory_manager.h (per-request cudaMalloc + name log) rebuilt
// as a single reservation sized for the device's HBM.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t top = 0;
  size_t peak = 0;
  static size_t align(size_t x) { return (x + 255) & ~(size_t)255; }
  void* alloc(size_t bytes) {
    size_t at = align(top);
    if (bytes > cap || at > cap - bytes) return nullptr;
    top = at + bytes;
    if (top > peak) peak = top;
    return base + at;
  }
  size_t mark() const { return top; }
  void release(size_t m) { top = m; }
  size_t avail() const {
    size_t at = align(top);
    return at >= cap ? 0 : cap - at;
  }
};

struct KStat {
  uint64_t launches = 0;
  uint64_t alg_bytes = 0;
  uint64_t units = 0;
  double ms = 0;
};

struct TimedLaunch {
  int fam;
  hipEvent_t a, b;
};

struct fsig_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Arena arena;
  std::string err;
  KStat kstat[FSIG_K_NUM];
  bool profiling = false;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  // Profiling: event pairs recorded around launches on the context stream and
  // resolved after the stream drains (no synchronisation per launch).
  std::vector<TimedLaunch> timed;
  size_t timed_used = 0;
  fsig_tune tune{0, 0, 0, 0, 0, 0, 0};
  std::vector<uint32_t> last_cand;
  // candidate sets of the last fsig_match per query vertex, kept when
  // keep_cand is set: [0] after the filter, [1] after refinement
  bool keep_cand = false;
  std::vector<std::vector<uint32_t>> cand_sets[2];
  uint64_t last_peak = 0;
  // device counters (64 words): [0..7] per-launch results, [16 + 4f ..] per
  // kernel family f accumulated device-side statistics ([0] algorithmic
  // bytes, [1] units).  Host staging is pinned.
  unsigned long long* d_counters = nullptr;
  unsigned long long* h_pinned = nullptr;  // kPinnedWords words of pinned staging
  int num_cus = 256;
  // choice of the batched engine's autotuner on its last batch
  uint64_t work_limit = 0;  // max intermediate rows per fsig_match (0 = none)
  uint32_t last_tune_G = 0;
  uint32_t last_tune_lds = 0;
  uint32_t last_tune_engine = 0;  // 1 lockstep batched BFS, 2 persistent walk
  // the batched engine's autotuner decisions, per graph and batch shape
  // ({lanes, lds, tile, lockstep}), and its persistent walk's instantiated
  // graphs, per batch shape (updated from each new capture, destroyed with
  // the context)
  std::map<std::pair<const fsig_graph*, uint64_t>, std::array<uint32_t, 4>> batch_tuned;
  std::map<uint64_t, hipGraphExec_t> batch_exec;
};

struct UnitMeta {
  uint64_t group_base;  // word offset of the unit's first group
  uint64_t edge_base;   // word offset of the unit's first neighbour id
  uint32_t keys;        // number of 32-word groups (= keys)
  uint32_t pad;
};

struct fsig_graph {
  uint32_t n = 0;
  uint32_t num_labels = 0;
  uint64_t nnz = 0;
  uint32_t max_degree = 0;
  uint32_t* d_off = nullptr;
  uint32_t* d_adj = nullptr;
  int32_t* d_labels = nullptr;
  uint32_t* d_inv_index = nullptr;
  uint32_t* d_inv_list = nullptr;
  std::vector<uint32_t> h_inv_index;
  // neighbour-label profile: nlf[v * num_labels + l] = number of neighbours
  // of v with label l, saturated at 255 (null for one label or > 255 labels)
  uint8_t* d_nlf = nullptr;
  // PCSR and signatures (alg 2)
  bool has_pcsr = false;
  int ref_unit_ids = 0;
  uint32_t num_units = 0;
  uint32_t* d_groups = nullptr;
  uint32_t* d_pedges = nullptr;
  UnitMeta* d_units = nullptr;
  uint32_t* d_sig = nullptr;  // [16][n]
  std::vector<UnitMeta> h_units;
};

// ---------------------------------------------------------------------------
// Host-side query plan.
struct QueryPlan {
  uint32_t k = 0;
  int32_t label[MAXQ];
  uint32_t deg[MAXQ];
  uint32_t adjm[MAXQ];  // neighbour bitmask
  uint32_t nclass = 0;
  uint32_t cls[MAXQ];          // class of query vertex
  uint32_t rep[MAXQ];          // representative vertex of a class
  int32_t class_label[MAXQ];
  uint32_t class_deg[MAXQ];
  uint32_t thr[MAXQ][MAXQ];    // refinement threshold thr[a][b]
  uint32_t class_sig[MAXQ][16];  // alg 2: query signature of a class
  uint32_t order[MAXQ];        // join position -> query vertex
  uint32_t pos[MAXQ];          // query vertex -> join position
  // neighbour-label requirements of a class: (label << 8) | count, one entry
  // per label among the representative's neighbours
  uint32_t nlf_n[MAXQ];
  uint32_t nlf_req[MAXQ][MAXQ];
  // symmetry breaking (count mode): bit b of sym_lt[a] asks phi(a) < phi(b);
  // the count is aut times the number of embeddings that satisfy all of them
  uint64_t aut = 1;
  uint32_t sym_lt[MAXQ] = {0};
};

int plan_classes(QueryPlan& P, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                 const uint32_t* qadj, const uint32_t* qsig, int alg, std::string& err);
void plan_order_mmatch(QueryPlan& P, const uint32_t* class_size);
// Fills aut and sym_lt (after plan_classes); false when the automorphism
// search ran past its step budget (then aut = 1 and no conditions).
bool plan_symmetry(QueryPlan& P);
void plan_order_gsi(QueryPlan& P, const uint32_t* class_size);

// ---------------------------------------------------------------------------
// Error helpers.
int set_err(fsig_ctx* c, int code, const std::string& msg);
#define FSIG_HIP(ctx, call)                                                        \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return set_err((ctx), FSIG_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Timing of one kernel family around a launch sequence: with profiling on,
// an event pair brackets the launches on the context stream; the pairs are
// resolved by resolve_timings() once the stream has drained.
struct KTimer {
  fsig_ctx* c;
  int fam;
  size_t slot = ~(size_t)0;  // index, not pointer: nested timers may grow the vector
  explicit KTimer(fsig_ctx* ctx, int family) : c(ctx), fam(family) {
    if (c->profiling) {
      if (c->timed_used == c->timed.size()) {
        TimedLaunch t{family, nullptr, nullptr};
        (void)hipEventCreate(&t.a);
        (void)hipEventCreate(&t.b);
        c->timed.push_back(t);
      }
      slot = c->timed_used++;
      c->timed[slot].fam = family;
      (void)hipEventRecord(c->timed[slot].a, c->stream);
    }
  }
  // alg_bytes / units: host-known figures (kernels that account on the
  // device pass 0 and add into the family's device counters instead).
  void done(uint64_t alg_bytes, uint64_t units) {
    KStat& s = c->kstat[fam];
    s.launches++;
    s.alg_bytes += alg_bytes;
    s.units += units;
    if (slot != ~(size_t)0) (void)hipEventRecord(c->timed[slot].b, c->stream);
  }
};
// Adds the recorded event intervals and the device-side family counters into
// the context statistics; the stream must be idle.
int resolve_timings(fsig_ctx* c);
inline unsigned long long* family_counters(fsig_ctx* c, int fam) { return c->d_counters + 16 + 4 * fam; }
// Refinement passes: the tuned cap, else 2 (one query or a batch).  After
// the neighbour-label filter the second pass still pays for itself on large
// graphs (smaller candidate edges and joins); later ones remove a few
// candidates at the price of a full pass.
inline uint32_t refine_passes(const fsig_ctx* c) { return c->tune.refine_passes ? c->tune.refine_passes : 2u; }

// Device-wide exclusive scan of n u32 counts into out[0..n] (u64, out[n] is
// the total).  Temporary storage comes from the arena.
int scan_u32_to_u64(fsig_ctx* c, const uint32_t* in, uint64_t* out, uint64_t n);

// Reads `count` u64 words from device to host through the pinned staging area.
int read_u64(fsig_ctx* c, const void* dev, uint64_t* host, size_t count);

// The per-query matcher pipelines.
// digest_acc (two zeroed device words, may be null): enumerate every
// embedding and fold it into the order-independent set digest there instead
// of copying rows to emb_out.
int run_mmatch(fsig_ctx* c, const fsig_graph* g, QueryPlan& P, int mode, uint64_t max_emb,
               uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap, unsigned long long* digest_acc);
int run_gsi(fsig_ctx* c, const fsig_graph* g, QueryPlan& P, int mode, uint64_t max_emb,
            uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap, unsigned long long* digest_acc);
