// Internal state of the C ABI (include/onionpir_gpu.h), shared by the
// K = 1 composite path (opir_capi.hip) and the K = 2 path (opir_capi_k2.hip).
#pragma once
#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/onionpir_gpu.h"
#include "opir_k2.h"
#include "opir_kernels.h"

using namespace opir;
typedef unsigned __int128 u128;

namespace capi {

struct ClientKeys {
  void *bv = nullptr;   // K=1: u32 [h][l_ks][2 (b,a)][2 limbs][N]; K=2: u64, same order
  size_t h = 0;
  void *gsw = nullptr;  // [2 l_key][2][2 limbs][N]
};

// Small ring of pinned staging slots for per-batch key-pointer tables.
constexpr int kRing = 8;

// One captured answer: the launch sequence for B queries with these key
// table, query and response pointers and this scan shape.
struct GraphEntry {
  size_t B = 0;
  const void *table = nullptr, *dq = nullptr, *dr = nullptr;
  int scan = -1;
  hipGraphExec_t exec = nullptr;
  uint64_t last_use = 0;
};
constexpr size_t kMaxGraphs = 16;  // per database

}  // namespace capi

struct opir_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  opir_scheme sch{};
  bool k2 = false;          // two 60-bit primes at N = 4096 (CONFIG_N4096_K2_MP, opir_capi_k2.hip)
  int K = 1;                // RNS limbs of a polynomial at the ABI (2: K2 schemes)
  double noise_sigma = 2.55;  // error standard deviation of device-generated keys and queries
  size_t N = kN;            // ring degree
  SchemeConst S{};          // K = 1 constants
  GadgetCfg G{};
  u32 *d_tables = nullptr;
  k2::Scheme64 S64{};       // K = 2 constants
  k2::Gadget64 G64{};
  u64 *d_tables64 = nullptr;
  // client keys; a clone (opir_ctx_clone) reads its parent's
  std::unordered_map<uint64_t, capi::ClientKeys> own_clients;
  opir_ctx *key_owner = nullptr;
  std::unordered_map<uint64_t, capi::ClientKeys> &clients() { return key_owner ? key_owner->own_clients : own_clients; }
  const std::unordered_map<uint64_t, capi::ClientKeys> &clients() const {
    return key_owner ? key_owner->own_clients : own_clients;
  }
  std::string err;
  bool profiling = false;
  int scan_variant = -1;    // first-dimension scan tile shape: -1 table default, -2 autotune, >= 0 forced
  bool graphs = false;      // replay each answer's launch sequence as a captured HIP graph
  hipEvent_t ev[8] = {};
  // auxiliary stream (default priority): selector completion runs on it
  // concurrently with the first-dimension scan
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // pointer-table ring
  const void **pin_ptrs[capi::kRing] = {};
  const void **dev_ptrs[capi::kRing] = {};
  hipEvent_t ring_ev[capi::kRing] = {};
  size_t ring_cap = 0;
  int ring_pos = 0;
  // host staging for synchronous calls
  void *pin = nullptr;
  size_t pin_bytes = 0;
  // this device's copy of the allocation-free entry point's scratch array
  void *scratch = nullptr;
  std::vector<uint32_t> cu_mask;  // compute-unit mask of the own streams (empty: all CUs)
  int dev_cus = 0;                // compute units of the device
};

struct opir_db {
  opir_ctx *ctx = nullptr;
  bool owns_db = true;     // false for a view (opir_db_view): d_db belongs to another handle
  size_t N0 = 0, Nrest = 0, dims = 0, h = 0, num_pt = 0, u = 0;
  DbGeom geom;
  int8_t *d_db = nullptr;  // digit planes in MFMA fragment order (opir_scan.hip), P = 4 (K = 1) or 8 (K = 2)
  // workspace for up to ws_B queries (K = 1 u32 RNS cts / K = 2 u64 cts)
  size_t ws_B = 0;
  u32 *exp_a = nullptr, *exp_b = nullptr, *qntt = nullptr, *sel = nullptr, *mid_a = nullptr, *mid_b = nullptr;
  u32 *part = nullptr;  // partial external products of split tree levels (kMuxSplitMaxBlocks cts)
  int8_t *qfrag = nullptr;
  u64 *e64_a = nullptr, *e64_b = nullptr, *q64 = nullptr, *sel64 = nullptr, *m64_a = nullptr, *m64_b = nullptr;
  u64 *q_u64 = nullptr, *r_u64 = nullptr;
  u64 *part64 = nullptr;  // partial products of split two-prime tree levels (kSplitTargetBlocks64 cts)
  size_t exp_cts = 0;
  // autotuned scan shape per batch-size class (ceil(log2(2B))), -1 = not tuned
  int tuned_scan[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  // captured answers (opir_ctx_set_graphs), least recently used dropped first
  std::vector<capi::GraphEntry> graphs;
  uint64_t graph_tick = 0;
};

namespace capi {
inline int scan_class(size_t B) {
  int c = 0;
  while (c < 15 && (size_t{1} << c) < 2 * B) c++;
  return c;
}

// The scan shape for B queries. In autotune mode (-2) the first call per
// batch-size class times every applicable shape on the live operands (two
// runs each, HIP events on the context stream; every shape computes the same
// bytes) and keeps the fastest. launch(v) enqueues the scan with shape v.
template <typename Launch>
int pick_scan_variant(opir_db *db, size_t B, int P, Launch launch) {
  opir_ctx *ctx = db->ctx;
  if (ctx->scan_variant != -2) return ctx->scan_variant;
  int &slot = db->tuned_scan[scan_class(B)];
  if (slot >= 0 && scan_variant_ok(P, slot, (int)B, db->geom)) return slot;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  float best = 1e30f;
  int pick = -1;
  for (int v = 0; v < scan_variant_count(P); v++) {
    if (!scan_variant_ok(P, v, (int)B, db->geom)) continue;
    launch(v);  // warm-up
    (void)hipEventRecord(e0, ctx->stream);
    launch(v);
    launch(v);
    (void)hipEventRecord(e1, ctx->stream);
    float ms = 0;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
    if (ms < best) {
      best = ms;
      pick = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  slot = pick;
  return pick;
}
}  // namespace capi

namespace capi {

extern thread_local std::string g_create_error;

inline int set_err(opir_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return capi::set_err((ctx), OPIR_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_CHECK(ctx)                                                                          \
  do {                                                                                             \
    hipError_t e_ = hipGetLastError();                                                             \
    if (e_ != hipSuccess) return capi::set_err((ctx), OPIR_EHIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// Fork: work queued on ctx->aux_stream after this runs after everything
// queued on ctx->stream so far. Join: ctx->stream waits for the aux work.
// Compute units the context's streams run on.
inline int ctx_cus(const opir_ctx *ctx) {
  if (ctx->cu_mask.empty()) return ctx->dev_cus > 0 ? ctx->dev_cus : 1;
  int n = 0;
  for (uint32_t w : ctx->cu_mask) n += __builtin_popcount(w);
  return n;
}

// Whether the selector completion runs beside the first-dimension scan on
// the auxiliary stream, or ahead of it on the answer stream. Measured on
// MI355X (256 x 1024 layout): for one query the completion alone takes
// 0.048 ms ahead of the scan, and the scan then takes 0.641-0.645 ms; beside
// it the completion's workgroups hold CUs that the scan's one round of
// workgroups needs, and the scan takes 0.669-0.672 ms, so the answer is
// 1.177-1.179 ms ahead against 1.195-1.197 ms beside. At 8 queries ahead is
// also faster (2.63 against 2.68 ms); from about 32 queries on the 32-bit
// completion (two 512-thread workgroups per CU) beside the scan ties or wins
// (64 queries: 13.36 ms either way; earlier builds 14.60 beside against
// 14.70 ahead). The 64-bit limbs' completion (whole_cu: each workgroup holds
// most of a CU's LDS) runs ahead at every batch size: beside the scan it
// waits for whole CUs that the scan's workgroups keep taking (27.02 against
// 27.18 ms for one n = 4096 query; 75 ms stretched beside a 32-query scan).
// The single-prime 64-bit completion (256-thread workgroups) takes 0.099 ms
// alone for one query, and ahead of the scan the answer took 1.540 against
// 1.503 ms beside it, so it stays beside (ahead_below = 0).
// ahead_below: row-queries (rows x B) below which the completion runs ahead.
inline bool complete_beside_scan(size_t rows, size_t B, bool whole_cu, size_t ahead_below) {
  return !whole_cu && rows * B >= ahead_below;
}

inline void fork_aux(opir_ctx *ctx) {
  (void)hipEventRecord(ctx->ev_fork, ctx->stream);
  (void)hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fork, 0);
}
inline void join_aux(opir_ctx *ctx) {
  (void)hipEventRecord(ctx->ev_join, ctx->aux_stream);
  (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
}

inline void drop_graphs(opir_db *db) {
  for (auto &g : db->graphs) (void)hipGraphExecDestroy(g.exec);
  db->graphs.clear();
}

// Runs launch() -- an answer's whole launch sequence on ctx->stream, which
// may fork to the auxiliary stream and joins it back -- directly, or with
// graphs on as a captured HIP graph: captured on the first call for the same
// (B, key table, query, response pointers, scan shape), replayed after. The
// sequence holds no memset nodes and no host synchronisation; an autotuned
// scan shape that is not settled yet (its timing runs synchronise) and
// profiling (its event records) run directly. Workspace reallocation drops
// the cached graphs.
template <typename Launch>
int run_answer(opir_db *db, size_t B, int P, const void *table, const void *dq, const void *dr, Launch launch) {
  opir_ctx *ctx = db->ctx;
  if (!ctx->graphs || ctx->profiling) return launch();
  int scan = ctx->scan_variant;
  if (scan == -2) {
    scan = db->tuned_scan[scan_class(B)];
    if (scan < 0 || !scan_variant_ok(P, scan, (int)B, db->geom)) return launch();
  }
  GraphEntry *hit = nullptr;
  for (auto &g : db->graphs)
    if (g.B == B && g.table == table && g.dq == dq && g.dr == dr && g.scan == scan) hit = &g;
  if (!hit) {
    if (db->graphs.size() >= kMaxGraphs) {
      auto lru = std::min_element(db->graphs.begin(), db->graphs.end(),
                                  [](const GraphEntry &a, const GraphEntry &b) { return a.last_use < b.last_use; });
      (void)hipGraphExecDestroy(lru->exec);
      db->graphs.erase(lru);
    }
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess) return set_err(ctx, OPIR_EHIP, std::string("graph capture: ") + hipGetErrorString(ec));
    GraphEntry e;
    const hipError_t ei = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return set_err(ctx, OPIR_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    e.B = B; e.table = table; e.dq = dq; e.dr = dr; e.scan = scan;
    db->graphs.push_back(e);
    hit = &db->graphs.back();
  }
  hit->last_use = ++db->graph_tick;
  HIP_TRY(ctx, hipGraphLaunch(hit->exec, ctx->stream));
  return OPIR_OK;
}

inline u64 mulmod_h(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
inline u64 powmod_h(u64 b, u64 e, u64 m) {
  u64 r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod_h(r, b, m);
    b = mulmod_h(b, b, m);
    e >>= 1;
  }
  return r;
}
inline bool inv_h(u64 a, u64 m, u64 &out) {
  __int128 s0 = 1, s1 = 0, r0 = (__int128)(a % m), r1 = (__int128)m;
  if (r0 == 0) return false;
  while (r1) {
    __int128 q = r0 / r1, t = r1;
    r1 = r0 - q * r1; r0 = t;
    t = s1; s1 = s0 - q * s1; s0 = t;
  }
  if (r0 != 1) return false;
  out = (u64)(s0 < 0 ? s0 + (__int128)m : s0);
  return true;
}
inline u32 shoup_h(u32 w, u32 q) { return (u32)(((u64)w << 32) / q); }
inline u64 shoup64_h(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }
inline size_t brv(size_t x, int logn) {
  size_t y = 0;
  for (int i = 0; i < logn; i++) { y = (y << 1) | (x & 1); x >>= 1; }
  return y;
}

inline int ensure_pin(opir_ctx *ctx, size_t bytes) {
  if (ctx->pin_bytes >= bytes) return OPIR_OK;
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  ctx->pin = nullptr;
  ctx->pin_bytes = 0;
  HIP_TRY(ctx, hipHostMalloc(&ctx->pin, bytes, hipHostMallocDefault));
  ctx->pin_bytes = bytes;
  return OPIR_OK;
}

// Upload per-query key pointers for a batch through the pinned ring:
// table[0..B) = BV keys, table[B..2B) = RGSW(s) keys.
// Grow the pinned key-pointer ring to hold batches of B queries.
inline int ensure_ring(opir_ctx *ctx, size_t B) {
  if (ctx->ring_cap < B) {
    for (int i = 0; i < kRing; i++) {
      if (ctx->ring_ev[i]) (void)hipEventSynchronize(ctx->ring_ev[i]);
      if (ctx->pin_ptrs[i]) (void)hipHostFree(ctx->pin_ptrs[i]);
      if (ctx->dev_ptrs[i]) (void)hipFree(ctx->dev_ptrs[i]);
      ctx->pin_ptrs[i] = nullptr;
      ctx->dev_ptrs[i] = nullptr;
    }
    size_t cap = 64;
    while (cap < B) cap *= 2;
    for (int i = 0; i < kRing; i++) {
      HIP_TRY(ctx, hipHostMalloc((void **)&ctx->pin_ptrs[i], 2 * cap * sizeof(void *), hipHostMallocDefault));
      HIP_TRY(ctx, hipMalloc((void **)&ctx->dev_ptrs[i], 2 * cap * sizeof(void *)));
      if (!ctx->ring_ev[i]) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ring_ev[i], hipEventDisableTiming));
    }
    ctx->ring_cap = cap;
  }
  return OPIR_OK;
}

inline int upload_key_ptrs(opir_ctx *ctx, const uint64_t *ids, size_t B, const void *const **table) {
  if (int rc = ensure_ring(ctx, B)) return rc;
  const int slot = ctx->ring_pos;
  ctx->ring_pos = (ctx->ring_pos + 1) % kRing;
  HIP_TRY(ctx, hipEventSynchronize(ctx->ring_ev[slot]));
  const void **hp = ctx->pin_ptrs[slot];
  for (size_t i = 0; i < B; i++) {
    auto it = ctx->clients().find(ids[i]);
    if (it == ctx->clients().end())
      return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(ids[i]));
    hp[i] = it->second.bv;
    hp[B + i] = it->second.gsw;
  }
  HIP_TRY(ctx, hipMemcpyAsync(ctx->dev_ptrs[slot], hp, 2 * B * sizeof(void *), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ring_ev[slot], ctx->stream));
  *table = ctx->dev_ptrs[slot];
  return OPIR_OK;
}

// Nodes of expansion level L (server.cpp:731-816): ceil(u * 2^L / 2^h), at most 2^L.
inline size_t nodes_at_level(size_t u, size_t h, size_t L) {
  const size_t cap = size_t{1} << h;
  const size_t n = (u * (size_t{1} << L) + cap - 1) / cap;
  const size_t full = size_t{1} << L;
  return n < full ? n : full;
}

inline int validate_mod(opir_ctx *ctx, const uint64_t *v, size_t n, uint64_t mod, const char *what) {
  for (size_t i = 0; i < n; i++)
    if (v[i] >= mod) return set_err(ctx, OPIR_EINVAL, std::string(what) + " value not reduced mod Q");
  return OPIR_OK;
}

// server.cpp:613-622: the selector tree must not be taller than the candidates.
inline int check_answerable(opir_db *db) {
  if (db->dims > 1 && db->Nrest < (size_t{1} << (db->dims - 2)))
    return set_err(db->ctx, OPIR_EINVAL,
                   "evaluate_other_dim: layout has fewer candidates than its selector tree; query-shape views "
                   "cannot answer make_query");
  return OPIR_OK;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};
#define DEV_ALLOC(ctx, buf, bytes) HIP_TRY(ctx, hipMalloc(&(buf).p, (bytes)))

// Composite / MP scheme answer-path pieces shared with the tree PIR entry
// points (opir_capi.hip): key-pointer upload for B queries, the answer
// workspace of db for B queries, and the expansion of the root cts in
// db->exp_a (returns the buffer holding the leaves).
int upload_k1(opir_ctx *ctx, const uint64_t *ids, size_t B, BatchPtrs &P);
size_t ws_bytes(const opir_db *db, size_t B);
int ensure_ws(opir_db *db, size_t B);
void release_ws(opir_db *db);
u32 *run_expand(opir_db *db, size_t B, const BatchPtrs &P);

// Resolves ctx->scratch for ctx->device (the device must be current); once
// per device, at context creation (opir_noalloc.hip).
int noalloc_init(opir_ctx *ctx);

// K = 2 implementations behind the same entry points (opir_capi_k2.hip).
int k2_validate_scheme(const opir_scheme *sch);  // host-only checks, before any device call
int k2_ctx_init(opir_ctx *ctx, const opir_scheme *sch);
int k2_validate_rns(opir_ctx *ctx, const uint64_t *v, size_t npoly, const char *what);
void k2_ctx_free(opir_ctx *ctx);
int k2_db_init(opir_db *db);
size_t k2_ws_bytes(const opir_db *db, size_t B);
int k2_ensure_ws(opir_db *db, size_t B);
void k2_release_ws(opir_db *db);
int k2_tune_scan(opir_db *db, size_t B);  // settle the autotuned scan shape for B queries
void k2_db_free(opir_db *db);
int k2_ingest(opir_db *db, const u64 *d_pts, u64 *d_tmp, size_t first_pt, size_t n);
int k2_load_plain(opir_db *db, const uint64_t *coeffs, size_t first_pt, size_t count);
int k2_set_keys(opir_ctx *ctx, const uint64_t *bv, size_t h, const uint64_t *gsw, ClientKeys &k);
int k2_answer_device(opir_db *db, size_t B, const uint64_t *ids, const u64 *d_q, u64 *d_r);
int k2_ntt(opir_ctx *ctx, uint64_t *data, size_t count, bool inverse);
int k2_expand(opir_db *db, uint64_t id, const uint64_t *query, uint64_t *out);
int k2_complete(opir_db *db, uint64_t id, const uint64_t *expanded, uint64_t *sel);
int k2_first_dim(opir_db *db, const uint64_t *query, uint64_t *out);
int k2_other_dim(opir_db *db, const uint64_t *mid, const uint64_t *sel, uint64_t *out);
int k2_ext_prod(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t b, const uint64_t *ct, uint64_t *out);
int k2_ext_prod_mux(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t b, const uint64_t *x, const uint64_t *y,
                    uint64_t *out);
int k2_mod_switch(opir_ctx *ctx, const uint64_t *ct, uint64_t *out);
int k2_bv_ks(opir_ctx *ctx, const uint64_t *ct, int kinv, const uint64_t *key, uint64_t *out);

}  // namespace capi
