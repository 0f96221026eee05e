// libfsig_hip.so: context, arena, resident graph, device scan and the C ABI
// entry points of include/fsig.h.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "device.h"
#include "internal.h"

int set_err(fsig_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (u32 in, u64 out): per-tile sums, one-workgroup
// scan of the tile sums, then per-tile scans.  Tile = 256 threads x 4 items.
namespace {

constexpr uint32_t kTile = FSIG_BLOCK * 4;

__global__ void __launch_bounds__(FSIG_BLOCK) k_scan_tile_sums(const uint32_t* __restrict__ in, uint64_t n,
                                                               uint64_t* __restrict__ sums) {
  __shared__ uint64_t lds[4];
  const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x * 4ull;
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (base + j < n) s += in[base + j];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

__global__ void __launch_bounds__(FSIG_BLOCK) k_scan_tile_offsets(uint64_t* __restrict__ sums, uint64_t ntiles,
                                                                  uint64_t* __restrict__ total_out) {
  __shared__ uint64_t lds[4];
  uint64_t carry = 0;
  for (uint64_t b = 0; b < ntiles; b += FSIG_BLOCK) {
    uint64_t i = b + threadIdx.x;
    uint64_t v = i < ntiles ? sums[i] : 0;
    uint64_t tot;
    uint64_t ex = block_exclusive_scan(v, lds, tot);
    if (i < ntiles) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ void __launch_bounds__(FSIG_BLOCK) k_scan_tiles(const uint32_t* __restrict__ in, uint64_t n,
                                                           const uint64_t* __restrict__ tile_off,
                                                           uint64_t* __restrict__ out) {
  __shared__ uint64_t lds[4];
  const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x * 4ull;
  uint32_t v[4];
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    v[j] = base + j < n ? in[base + j] : 0;
    s += v[j];
  }
  uint64_t tot;
  uint64_t ex = block_exclusive_scan(s, lds, tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (base + j < n) out[base + j] = ex;
    ex += v[j];
  }
}

// Zeroes a few words (a kernel rather than a memset, so it may sit in a
// captured graph).
__global__ void k_fill_words(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Plain copy, 16 bytes per thread per step (the STREAM "copy" kernel).
__global__ void __launch_bounds__(FSIG_BLOCK) k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                            uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

int scan_u32_to_u64(fsig_ctx* c, const uint32_t* in, uint64_t* out, uint64_t n) {
  KTimer t(c, FSIG_K_SCAN);
  const uint64_t ntiles = (n + kTile - 1) / kTile;
  size_t mk = c->arena.mark();
  uint64_t* sums = (uint64_t*)c->arena.alloc(sizeof(uint64_t) * std::max<uint64_t>(ntiles, 1));
  if (!sums) return set_err(c, FSIG_ENOMEM, "scan: arena exhausted");
  if (ntiles > 0) {
    k_scan_tile_sums<<<(uint32_t)ntiles, FSIG_BLOCK, 0, c->stream>>>(in, n, sums);
  }
  k_scan_tile_offsets<<<1, FSIG_BLOCK, 0, c->stream>>>(sums, ntiles, out + n);
  if (ntiles > 0) {
    k_scan_tiles<<<(uint32_t)ntiles, FSIG_BLOCK, 0, c->stream>>>(in, n, sums, out);
  }
  FSIG_HIP(c, hipGetLastError());
  c->arena.release(mk);
  t.done(n * 4 + n * 8 + 8, n);
  return FSIG_OK;
}

int resolve_timings(fsig_ctx* c) {
  for (size_t i = 0; i < c->timed_used; i++) {
    float ms = 0;
    FSIG_HIP(c, hipEventElapsedTime(&ms, c->timed[i].a, c->timed[i].b));
    c->kstat[c->timed[i].fam].ms += ms;
  }
  c->timed_used = 0;
  uint64_t fam[4 * FSIG_K_NUM];
  int rc = read_u64(c, c->d_counters + 16, fam, 4 * FSIG_K_NUM);
  if (rc) return rc;
  for (int f = 0; f < FSIG_K_NUM; f++) {
    c->kstat[f].alg_bytes += fam[4 * f];
    c->kstat[f].units += fam[4 * f + 1];
  }
  FSIG_HIP(c, hipMemsetAsync(c->d_counters + 16, 0, 8 * 4 * FSIG_K_NUM, c->stream));
  return FSIG_OK;
}

int read_u64(fsig_ctx* c, const void* dev, uint64_t* host, size_t count) {
  if (count > kPinnedWords) return set_err(c, FSIG_EINVAL, "read_u64: too many words");
  FSIG_HIP(c, hipMemcpyAsync(c->h_pinned, dev, count * 8, hipMemcpyDeviceToHost, c->stream));
  FSIG_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(host, c->h_pinned, count * 8);
  return FSIG_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// Neighbour-label profile: one wavefront per vertex counts its neighbours'
// labels in LDS and stores them saturated to a byte.
__global__ void __launch_bounds__(FSIG_BLOCK) k_nlf_build(const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ adj,
                                                          const int32_t* __restrict__ lab, uint32_t n, uint32_t L,
                                                          uint8_t* __restrict__ nlf) {
  extern __shared__ uint32_t hist[];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t* h = hist + wave * L;
  for (uint64_t v = (uint64_t)blockIdx.x * (FSIG_BLOCK / 64) + wave; v < n;
       v += (uint64_t)gridDim.x * (FSIG_BLOCK / 64)) {
    for (uint32_t l = lane; l < L; l += 64) h[l] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t j = off[v] + lane; j < off[v + 1]; j += 64) atomicAdd(&h[lab[adj[j]]], 1u);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t l = lane; l < L; l += 64) nlf[v * L + l] = (uint8_t)(h[l] > 255u ? 255u : h[l]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

extern "C" {

int fsig_ctx_create(int device, uint64_t budget, fsig_ctx** out) {
  if (!out) return FSIG_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FSIG_EHIP;
  if (device < 0 || device >= ndev) return FSIG_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return FSIG_EHIP;
  fsig_ctx* c = new fsig_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) { delete c; return FSIG_EHIP; }
  if (budget == 0) budget = (uint64_t)freeb / 4 * 3;
  if (budget > freeb) { delete c; return FSIG_ENOMEM; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev_a) != hipSuccess || hipEventCreate(&c->ev_b) != hipSuccess ||
      hipMalloc((void**)&c->d_counters, 64 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_pinned, kPinnedWords * 8, hipHostMallocDefault) != hipSuccess) {
    fsig_ctx_destroy(c);
    return FSIG_EHIP;
  }
  if (hipMalloc((void**)&c->arena.base, budget) != hipSuccess) {
    fsig_ctx_destroy(c);
    return FSIG_ENOMEM;
  }
  c->arena.cap = budget;
  if (hipMemset(c->d_counters, 0, 64 * sizeof(unsigned long long)) != hipSuccess) {
    fsig_ctx_destroy(c);
    return FSIG_EHIP;
  }
  *out = c;
  return FSIG_OK;
}

void fsig_ctx_destroy(fsig_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->arena.base) (void)hipFree(c->arena.base);
  if (c->d_counters) (void)hipFree(c->d_counters);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  for (auto& t : c->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  if (c->ev_a) (void)hipEventDestroy(c->ev_a);
  if (c->ev_b) (void)hipEventDestroy(c->ev_b);
  for (auto& e : c->batch_exec) (void)hipGraphExecDestroy(e.second);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* fsig_last_error(const fsig_ctx* c) { return c ? c->err.c_str() : "no context"; }

int fsig_ctx_synchronize(fsig_ctx* c) {
  if (!c) return FSIG_EINVAL;
  FSIG_HIP(c, hipSetDevice(c->device));
  FSIG_HIP(c, hipDeviceSynchronize());
  return FSIG_OK;
}

int fsig_graph_upload_csr(fsig_ctx* c, uint32_t n, const uint32_t* off, const uint32_t* adj,
                          const int32_t* labels, uint32_t num_labels, const uint32_t* inv_index,
                          const uint32_t* inv_list, fsig_graph** out) {
  if (!c || !out || !off || !labels || !inv_index || !inv_list || (off[n] > 0 && !adj))
    return set_err(c, FSIG_EINVAL, "upload_csr: null argument");
  *out = nullptr;
  if (num_labels == 0 || num_labels > (1u << 20)) return set_err(c, FSIG_EINVAL, "upload_csr: bad label count");
  // Host validation: the join treats adjacency lists as sets.
  if (off[0] != 0) return set_err(c, FSIG_EINVAL, "upload_csr: off[0] != 0");
  uint32_t maxdeg = 0;
  for (uint32_t v = 0; v < n; v++) {
    if (off[v + 1] < off[v]) return set_err(c, FSIG_EINVAL, "upload_csr: offsets decrease");
    maxdeg = std::max(maxdeg, off[v + 1] - off[v]);
    if (labels[v] < 0 || (uint32_t)labels[v] >= num_labels)
      return set_err(c, FSIG_EINVAL, "upload_csr: label out of range");
    for (uint32_t i = off[v]; i < off[v + 1]; i++) {
      if (adj[i] >= n) return set_err(c, FSIG_EINVAL, "upload_csr: neighbour id out of range");
      if (adj[i] == v) return set_err(c, FSIG_EINVAL, "upload_csr: self-loop");
      if (i > off[v] && adj[i] <= adj[i - 1])
        return set_err(c, FSIG_EINVAL, "upload_csr: adjacency not strictly ascending");
    }
  }
  if (inv_index[0] != 0 || inv_index[num_labels] != n)
    return set_err(c, FSIG_EINVAL, "upload_csr: inverted index does not cover the vertices");
  for (uint32_t l = 0; l < num_labels; l++) {
    if (inv_index[l + 1] < inv_index[l]) return set_err(c, FSIG_EINVAL, "upload_csr: inverted index decreases");
    for (uint32_t i = inv_index[l]; i < inv_index[l + 1]; i++) {
      uint32_t v = inv_list[i];
      if (v >= n || labels[v] != (int32_t)l || (i > inv_index[l] && v <= inv_list[i - 1]))
        return set_err(c, FSIG_EINVAL, "upload_csr: inverted index inconsistent with labels");
    }
  }
  FSIG_HIP(c, hipSetDevice(c->device));
  fsig_graph* g = new fsig_graph();
  g->n = n;
  g->num_labels = num_labels;
  g->nnz = off[n];
  g->max_degree = maxdeg;
  g->h_inv_index.assign(inv_index, inv_index + num_labels + 1);
  auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
    if (hipMalloc(dst, std::max<size_t>(bytes, 4)) != hipSuccess) return false;
    return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up((void**)&g->d_off, off, sizeof(uint32_t) * (n + 1ull)) &&
            up((void**)&g->d_adj, adj, sizeof(uint32_t) * g->nnz) &&
            up((void**)&g->d_labels, labels, sizeof(int32_t) * n) &&
            up((void**)&g->d_inv_index, inv_index, sizeof(uint32_t) * (num_labels + 1ull)) &&
            up((void**)&g->d_inv_list, inv_list, sizeof(uint32_t) * n);
  if (!ok) {
    fsig_graph_free(c, g);
    return set_err(c, FSIG_ENOMEM, "upload_csr: device allocation or copy failed");
  }
  // neighbour-label profile (a filter of alg 1), when labels are few enough
  // for one byte row per vertex
  if (num_labels >= 2 && num_labels <= 255 && (uint64_t)n * num_labels <= (16ull << 30)) {
    if (hipMalloc((void**)&g->d_nlf, std::max<uint64_t>((uint64_t)n * num_labels, 4)) != hipSuccess) {
      fsig_graph_free(c, g);
      return set_err(c, FSIG_ENOMEM, "upload_csr: no memory for the neighbour-label profile");
    }
    if (n) {
      k_nlf_build<<<(uint32_t)std::min<uint64_t>((n + 3) / 4, (uint64_t)c->num_cus * 64), FSIG_BLOCK,
                    (FSIG_BLOCK / 64) * num_labels * sizeof(uint32_t), c->stream>>>(g->d_off, g->d_adj, g->d_labels,
                                                                                   n, num_labels, g->d_nlf);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        fsig_graph_free(c, g);
        return set_err(c, FSIG_EHIP, "upload_csr: neighbour-label profile kernel failed");
      }
    }
  }
  *out = g;
  return FSIG_OK;
}

int fsig_graph_upload_pcsr(fsig_ctx* c, fsig_graph* g, uint32_t num_units, const uint32_t* unit_keys,
                           const uint32_t* const* unit_offset, const uint32_t* unit_nedges,
                           const uint32_t* const* unit_edges, const uint32_t* sig, int ref_ids) {
  if (!c || !g || !unit_keys || !unit_offset || !unit_nedges || !unit_edges || !sig)
    return set_err(c, FSIG_EINVAL, "upload_pcsr: null argument");
  const uint64_t L = g->num_labels;
  if (num_units != L * (L + 1) / 2) return set_err(c, FSIG_EINVAL, "upload_pcsr: unit count != L(L+1)/2");
  std::vector<UnitMeta> meta(num_units);
  uint64_t gw = 0, ew = 0;
  for (uint32_t u = 0; u < num_units; u++) {
    meta[u].group_base = gw;
    meta[u].edge_base = ew;
    meta[u].keys = unit_keys[u];
    meta[u].pad = 0;
    gw += 32ull * unit_keys[u];
    ew += unit_nedges[u];
    const uint32_t* o = unit_offset[u];
    // every stored range must lie inside the unit's edge array
    for (uint64_t j = 0; j < unit_keys[u]; j++) {
      for (int s = 0; s < 15; s++) {
        if (o[32 * j + 2 * s] == FSIG_INVALID) break;
        uint32_t b = o[32 * j + 2 * s + 1], e = o[32 * j + 2 * s + 3];
        if (o[32 * j + 2 * s] >= g->n || b > e || e > unit_nedges[u])
          return set_err(c, FSIG_EINVAL, "upload_pcsr: bucket slot out of range");
      }
      uint32_t nxt = o[32 * j + 30];
      if (nxt != FSIG_INVALID && nxt >= unit_keys[u]) return set_err(c, FSIG_EINVAL, "upload_pcsr: bad overflow id");
    }
    for (uint32_t i = 0; i < unit_nedges[u]; i++)
      if (unit_edges[u][i] >= g->n) return set_err(c, FSIG_EINVAL, "upload_pcsr: neighbour id out of range");
  }
  FSIG_HIP(c, hipSetDevice(c->device));
  // a failed re-upload leaves the graph without a PCSR, never with freed buffers
  g->has_pcsr = false;
  g->num_units = 0;
  if (g->d_groups) { (void)hipFree(g->d_groups); g->d_groups = nullptr; }
  if (g->d_pedges) { (void)hipFree(g->d_pedges); g->d_pedges = nullptr; }
  if (g->d_units) { (void)hipFree(g->d_units); g->d_units = nullptr; }
  if (g->d_sig) { (void)hipFree(g->d_sig); g->d_sig = nullptr; }
  if (hipMalloc((void**)&g->d_groups, std::max<uint64_t>(gw, 1) * 4) != hipSuccess ||
      hipMalloc((void**)&g->d_pedges, std::max<uint64_t>(ew, 1) * 4) != hipSuccess ||
      hipMalloc((void**)&g->d_units, sizeof(UnitMeta) * std::max<uint32_t>(num_units, 1)) != hipSuccess ||
      hipMalloc((void**)&g->d_sig, 64ull * std::max<uint32_t>(g->n, 1)) != hipSuccess)
    return set_err(c, FSIG_ENOMEM, "upload_pcsr: device allocation failed");
  for (uint32_t u = 0; u < num_units; u++) {
    if (unit_keys[u])
      FSIG_HIP(c, hipMemcpy(g->d_groups + meta[u].group_base, unit_offset[u], 128ull * unit_keys[u],
                            hipMemcpyHostToDevice));
    if (unit_nedges[u])
      FSIG_HIP(c, hipMemcpy(g->d_pedges + meta[u].edge_base, unit_edges[u], 4ull * unit_nedges[u],
                            hipMemcpyHostToDevice));
  }
  FSIG_HIP(c, hipMemcpy(g->d_units, meta.data(), sizeof(UnitMeta) * num_units, hipMemcpyHostToDevice));
  FSIG_HIP(c, hipMemcpy(g->d_sig, sig, 64ull * g->n, hipMemcpyHostToDevice));
  g->h_units = meta;
  g->num_units = num_units;
  g->ref_unit_ids = ref_ids;
  g->has_pcsr = true;
  return FSIG_OK;
}

void fsig_graph_free(fsig_ctx* c, fsig_graph* g) {
  if (!g) return;
  if (c) (void)hipSetDevice(c->device);
  void* ptrs[] = {g->d_off, g->d_adj, g->d_labels, g->d_inv_index, g->d_inv_list,
                  g->d_groups, g->d_pedges, g->d_units, g->d_sig, g->d_nlf};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete g;
}

}  // extern "C"

// One match; with digest_out the embeddings are enumerated into the set
// digest instead of being copied out.
static int match_impl(fsig_ctx* c, const fsig_graph* g, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                      const uint32_t* qadj, const uint32_t* qsig, int alg, int mode, uint64_t max_emb,
                      uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap, uint64_t* digest_out) {
  if (!c) return FSIG_EINVAL;
  if (!g || !qlabels || !qoff || !count_out) return set_err(c, FSIG_EINVAL, "match: null argument");
  if (k > 0 && qoff[k] > 0 && !qadj) return set_err(c, FSIG_EINVAL, "match: null query adjacency");
  if (mode != FSIG_MODE_COUNT && mode != FSIG_MODE_ENUMERATE) return set_err(c, FSIG_EINVAL, "match: bad mode");
  if (alg != FSIG_ALG_MMATCH && alg != FSIG_ALG_GSI) return set_err(c, FSIG_EINVAL, "match: bad algorithm");
  if (alg == FSIG_ALG_GSI && !g->has_pcsr) return set_err(c, FSIG_ENOTFOUND, "match: graph has no PCSR upload");
  *count_out = 0;
  if (digest_out) digest_out[0] = digest_out[1] = 0;
  FSIG_HIP(c, hipSetDevice(c->device));
  QueryPlan P;
  int rc = plan_classes(P, k, qlabels, qoff, qadj, qsig, alg, c->err);
  if (rc) return rc;
  c->cand_sets[0].clear();
  c->cand_sets[1].clear();
  for (uint32_t u = 0; u < k; u++)
    if ((uint32_t)qlabels[u] >= g->num_labels) {
      // a label absent from the data graph has no candidates
      c->last_cand.assign(k, 0);
      if (c->keep_cand) c->cand_sets[0].assign(k, {}), c->cand_sets[1].assign(k, {});
      return FSIG_OK;
    }
  const size_t mk = c->arena.mark();
  c->arena.peak = mk;
  unsigned long long* acc = nullptr;
  if (digest_out) {
    acc = (unsigned long long*)c->arena.alloc(2 * sizeof(unsigned long long));
    if (!acc) return set_err(c, FSIG_ENOMEM, "digest: arena exhausted");
    k_fill_words<<<1, 64, 0, c->stream>>>((uint32_t*)acc, 4);
    FSIG_HIP(c, hipGetLastError());
  }
  // A count counts one embedding per orbit of the query's automorphisms
  // (symmetry conditions on the join) and multiplies by their number; a
  // bound b then stops the search at ceil(b / |Aut|) orbits.
  uint64_t run_bound = max_emb;
  if (mode == FSIG_MODE_COUNT && !digest_out && c->tune.symmetry == 0 && plan_symmetry(P) && P.aut > 1 && max_emb)
    run_bound = max_emb / P.aut + (max_emb % P.aut != 0);
  rc = alg == FSIG_ALG_MMATCH ? run_mmatch(c, g, P, mode, run_bound, count_out, emb_out, emb_cap, acc)
                              : run_gsi(c, g, P, mode, run_bound, count_out, emb_out, emb_cap, acc);
  if (rc == FSIG_OK && P.aut > 1) {
    const uint64_t orbits = *count_out;
    *count_out = orbits > UINT64_MAX / P.aut ? UINT64_MAX : orbits * P.aut;
    if (max_emb) *count_out = std::min(*count_out, max_emb);
  }
  if (rc == FSIG_OK && acc) rc = read_u64(c, acc, digest_out, 2);
  c->last_peak = c->arena.peak - mk;
  c->arena.release(mk);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(c, FSIG_EHIP, std::string("match: ") + hipGetErrorString(e));
  int rt = resolve_timings(c);
  return rc ? rc : rt;
}

extern "C" {

int fsig_match(fsig_ctx* c, const fsig_graph* g, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
               const uint32_t* qadj, const uint32_t* qsig, int alg, int mode, uint64_t max_emb,
               uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap) {
  return match_impl(c, g, k, qlabels, qoff, qadj, qsig, alg, mode, max_emb, count_out, emb_out, emb_cap, nullptr);
}

int fsig_match_digest(fsig_ctx* c, const fsig_graph* g, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                      const uint32_t* qadj, const uint32_t* qsig, int alg, uint64_t* count_out,
                      uint64_t* digest_out) {
  if (c && !digest_out) return set_err(c, FSIG_EINVAL, "digest: null digest");
  return match_impl(c, g, k, qlabels, qoff, qadj, qsig, alg, FSIG_MODE_ENUMERATE, 0, count_out, nullptr, 0,
                    digest_out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Static scratch of fsig_match_noalloc: the device counters (first 512
// bytes) and the arena of one match.  Each device that loads the library
// holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory: its device address is taken
// once per device with hipGetSymbolAddress and only that address is used.
constexpr size_t kScratchCounters = 64 * sizeof(unsigned long long);
static __device__ __align__(256) char scratch[FSIG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {nullptr};

// Device address of `scratch` on the context's device (looked up once per
// device, with that device current).
int scratch_of(fsig_ctx* c, char** out) {
  if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, FSIG_EUNSUPPORTED, "noalloc: device id too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    FSIG_HIP(c, hipSetDevice(c->device));
    void* p = nullptr;
    FSIG_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(c, FSIG_EHIP, "noalloc: no device address for the scratch array");
    g_scratch_ptr[c->device] = (char*)p;
  }
  *out = g_scratch_ptr[c->device];
  return FSIG_OK;
}
}  // namespace

extern "C" {

int fsig_match_noalloc(fsig_ctx* c, const fsig_graph* g, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                       const uint32_t* qadj, const uint32_t* qsig, int alg, int mode, uint64_t max_emb,
                       uint64_t* count_out, uint32_t* emb_out, uint64_t emb_cap) {
  if (!c) return FSIG_EINVAL;
  if (!g || !count_out) return set_err(c, FSIG_EINVAL, "noalloc: null argument");
  // the candidate mask (4 bytes a vertex) and the candidate arrays (at most
  // 4 bytes a vertex of the query's labels, twice) must fit with room left
  const uint64_t need = 12ull * g->n + (1u << 20);
  if (need > FSIG_NOALLOC_SCRATCH_BYTES - kScratchCounters)
    return set_err(c, FSIG_EINVAL,
                   "noalloc: a graph of " + std::to_string(g->n) + " vertices needs at least " +
                       std::to_string(need) + " bytes of scratch, more than the static " +
                       std::to_string((unsigned long long)FSIG_NOALLOC_SCRATCH_BYTES) + " bytes");
  char* base = nullptr;
  int rc = scratch_of(c, &base);
  if (rc) return rc;
  // swap the context's counters and arena for the scratch for this call; no
  // event is recorded (profiling off), nothing is allocated on the device
  Arena saved_arena = c->arena;
  unsigned long long* saved_counters = c->d_counters;
  const bool saved_prof = c->profiling;
  c->d_counters = (unsigned long long*)base;
  c->arena = Arena();
  c->arena.base = base + kScratchCounters;
  c->arena.cap = FSIG_NOALLOC_SCRATCH_BYTES - kScratchCounters;
  c->profiling = false;
  rc = FSIG_OK;
  // the counters must start at zero: earlier calls leave theirs behind
  if (hipSetDevice(c->device) != hipSuccess) rc = set_err(c, FSIG_EHIP, "noalloc: hipSetDevice failed");
  if (rc == FSIG_OK) {
    k_fill_words<<<1, 64, 0, c->stream>>>((uint32_t*)base, (uint32_t)(kScratchCounters / 4));
    if (hipGetLastError() != hipSuccess) rc = set_err(c, FSIG_EHIP, "noalloc: clearing the scratch counters failed");
  }
  if (rc == FSIG_OK)
    rc = match_impl(c, g, k, qlabels, qoff, qadj, qsig, alg, mode, max_emb, count_out, emb_out, emb_cap, nullptr);
  if (rc == FSIG_ENOMEM) c->err = "noalloc: the static scratch is exhausted (" + c->err + ")";
  c->arena = saved_arena;
  c->d_counters = saved_counters;
  c->profiling = saved_prof;
  return rc;
}

int fsig_set_keep_candidates(fsig_ctx* c, int on) {
  if (!c) return FSIG_EINVAL;
  c->keep_cand = on != 0;
  if (!c->keep_cand) {
    c->cand_sets[0].clear();
    c->cand_sets[1].clear();
  }
  return FSIG_OK;
}

int fsig_last_candidate_set(const fsig_ctx* c, uint32_t u, int stage, uint32_t* out, uint32_t cap,
                            uint32_t* size) {
  if (!c || !size || stage < 0 || stage > 1) return FSIG_EINVAL;
  const auto& sets = c->cand_sets[stage];
  if (u >= sets.size()) return FSIG_ENOTFOUND;
  const auto& v = sets[u];
  *size = (uint32_t)v.size();
  if (out) std::memcpy(out, v.data(), 4ull * std::min<size_t>(cap, v.size()));
  return FSIG_OK;
}

int fsig_stream_copy(fsig_ctx* c, uint64_t bytes, uint32_t iters, double* gbps) {
  if (!c || !gbps || bytes < 16 || iters == 0) return set_err(c, FSIG_EINVAL, "stream_copy: bad argument");
  FSIG_HIP(c, hipSetDevice(c->device));
  const uint64_t n = bytes / 16;
  const size_t mk = c->arena.mark();
  uint4* a = (uint4*)c->arena.alloc(n * 16);
  uint4* b = (uint4*)c->arena.alloc(n * 16);
  if (!a || !b) {
    c->arena.release(mk);
    return set_err(c, FSIG_ENOMEM, "stream_copy: arena too small");
  }
  FSIG_HIP(c, hipMemsetAsync(a, 1, n * 16, c->stream));
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + FSIG_BLOCK - 1) / FSIG_BLOCK, (uint64_t)c->num_cus * 32);
  k_stream_copy<<<grid, FSIG_BLOCK, 0, c->stream>>>(a, b, n);  // warm
  FSIG_HIP(c, hipEventRecord(c->ev_a, c->stream));
  for (uint32_t i = 0; i < iters; i++) k_stream_copy<<<grid, FSIG_BLOCK, 0, c->stream>>>(a, b, n);
  FSIG_HIP(c, hipEventRecord(c->ev_b, c->stream));
  FSIG_HIP(c, hipEventSynchronize(c->ev_b));
  float ms = 0;
  FSIG_HIP(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
  c->arena.release(mk);
  *gbps = 2.0 * (double)(n * 16) * iters / (ms * 1e-3) / 1e9;
  return FSIG_OK;
}

int fsig_set_work_limit(fsig_ctx* c, uint64_t max_rows) {
  if (!c) return FSIG_EINVAL;
  c->work_limit = max_rows;
  return FSIG_OK;
}

int fsig_set_profiling(fsig_ctx* c, int on) {
  if (!c) return FSIG_EINVAL;
  c->profiling = on != 0;
  return FSIG_OK;
}

int fsig_stats_reset(fsig_ctx* c) {
  if (!c) return FSIG_EINVAL;
  for (auto& s : c->kstat) s = KStat();
  return FSIG_OK;
}

int fsig_kernel_stats(const fsig_ctx* c, int fam, uint64_t* launches, double* ms, uint64_t* bytes, uint64_t* units) {
  if (!c || fam < 0 || fam >= FSIG_K_NUM) return FSIG_EINVAL;
  const KStat& s = c->kstat[fam];
  if (launches) *launches = s.launches;
  if (ms) *ms = s.ms;
  if (bytes) *bytes = s.alg_bytes;
  if (units) *units = s.units;
  return FSIG_OK;
}

int fsig_last_candidates(const fsig_ctx* c, uint32_t k, uint32_t* sizes) {
  if (!c || !sizes) return FSIG_EINVAL;
  for (uint32_t i = 0; i < k; i++) sizes[i] = i < c->last_cand.size() ? c->last_cand[i] : 0;
  return FSIG_OK;
}

uint64_t fsig_last_peak_bytes(const fsig_ctx* c) { return c ? c->last_peak : 0; }

int fsig_last_batch_tune(const fsig_ctx* c, uint32_t* lanes, uint32_t* tile, uint32_t* engine) {
  if (!c) return FSIG_EINVAL;
  if (lanes) *lanes = c->last_tune_G;
  if (tile) *tile = c->last_tune_lds;
  if (engine) *engine = c->last_tune_engine;
  return FSIG_OK;
}

int fsig_set_tune(fsig_ctx* c, const fsig_tune* t) {
  if (!c || !t) return FSIG_EINVAL;
  const uint32_t g = t->lanes_per_row;
  if (g != 0 && g != 4 && g != 8 && g != 16 && g != 32 && g != 64) return set_err(c, FSIG_EINVAL, "tune: lanes_per_row");
  if (t->strategy > 3) return set_err(c, FSIG_EINVAL, "tune: strategy");
  if (t->fused_tail > 1) return set_err(c, FSIG_EINVAL, "tune: fused_tail");
  if (t->nlf_filter > 1) return set_err(c, FSIG_EINVAL, "tune: nlf_filter");
  if (t->beam_log2 > 40) return set_err(c, FSIG_EINVAL, "tune: beam_log2");
  if (t->symmetry > 1) return set_err(c, FSIG_EINVAL, "tune: symmetry");
  c->tune = *t;
  return FSIG_OK;
}

int fsig_get_tune(const fsig_ctx* c, fsig_tune* t) {
  if (!c || !t) return FSIG_EINVAL;
  *t = c->tune;
  return FSIG_OK;
}

}  // extern "C"
