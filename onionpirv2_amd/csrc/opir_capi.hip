// C ABI of the OnionPIRv2 MI355X server library (include/onionpir_gpu.h).
// Orchestrates the kernels of opir_kernels.hip along PirServer::make_query
// (src/server.cpp:895-947).
#include "opir_capi_internal.h"

using namespace capi;

thread_local std::string capi::g_create_error;

namespace {

// u64 values per polynomial at the ABI: N (composite, mod Q) or 2N (the
// two-limb MP scheme, [limb][N] residues).
inline size_t abi_kn(const opir_ctx *ctx) { return (size_t)ctx->K * kN; }

// npoly ABI polynomials must hold reduced values: < Q (composite) or each
// limb below its prime (MP).
int validate_polys(opir_ctx *ctx, const uint64_t *v, size_t npoly, const char *what) {
  if (!ctx->S.mp) return validate_mod(ctx, v, npoly * kN, ctx->S.Q, what);
  const u64 q[2] = {ctx->S.limb[0].q, ctx->S.limb[1].q};
  for (size_t i = 0; i < npoly * 2 * kN; i++)
    if (v[i] >= q[(i / kN) & 1]) return set_err(ctx, OPIR_EINVAL, std::string(what) + " value not reduced mod its RNS prime");
  return OPIR_OK;
}

// One ABI polynomial -> its two u32 limbs (composite: residues of the value
// mod Q; MP: the limbs as given).
inline void poly_to_rns(const opir_ctx *ctx, const uint64_t *src, u32 *dst) {
  const u64 q1 = ctx->S.limb[0].q, q2 = ctx->S.limb[1].q;
  for (size_t j = 0; j < (size_t)kN; j++) {
    dst[j] = (u32)(ctx->S.mp ? src[j] : src[j] % q1);
    dst[kN + j] = (u32)(ctx->S.mp ? src[kN + j] : src[j] % q2);
  }
}

}  // namespace

namespace capi {

int upload_k1(opir_ctx *ctx, const uint64_t *ids, size_t B, BatchPtrs &P) {
  const void *const *t = nullptr;
  const int rc = upload_key_ptrs(ctx, ids, B, &t);
  if (rc) return rc;
  P.bv_keys = reinterpret_cast<const u32 *const *>(t);
  P.gsw_keys = reinterpret_cast<const u32 *const *>(t + B);
  return OPIR_OK;
}

// Device bytes of the answer workspace for batches of up to B queries (the
// buffers ensure_ws allocates).
size_t ws_bytes(const opir_db *db, size_t B) {
  const size_t ct = kCtWords * sizeof(u32);
  const size_t nsel = db->dims - 1;
  return qfrag_bytes(make_qfrag_geom(db->N0, (int)B)) + 2 * B * db->exp_cts * ct + B * db->N0 * ct +
         B * (nsel ? nsel : 1) * 2 * db->ctx->G.l_ep * ct + 2 * B * db->Nrest * ct + B * 2 * (abi_kn(db->ctx) + kN) * sizeof(u64) +
         kMuxSplitMaxBlocks * ct;
}

// Frees the answer workspace (the caller has drained the stream).
void release_ws(opir_db *db) {
  u32 **bufs[] = {&db->exp_a, &db->exp_b, &db->qntt, &db->sel, &db->mid_a, &db->mid_b, &db->part};
  for (auto b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  if (db->q_u64) (void)hipFree(db->q_u64);
  if (db->r_u64) (void)hipFree(db->r_u64);
  if (db->qfrag) (void)hipFree(db->qfrag);
  db->q_u64 = db->r_u64 = nullptr;
  db->qfrag = nullptr;
  db->ws_B = 0;
}

int ensure_ws(opir_db *db, size_t B) {
  opir_ctx *ctx = db->ctx;
  if (db->ws_B >= B) return OPIR_OK;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // queued work may still use the old buffers
  drop_graphs(db);
  release_ws(db);
  HIP_TRY(ctx, hipMalloc((void **)&db->qfrag, qfrag_bytes(make_qfrag_geom(db->N0, (int)B))));
  const size_t ct = kCtWords * sizeof(u32);
  const size_t nsel = db->dims - 1;
  HIP_TRY(ctx, hipMalloc((void **)&db->exp_a, B * db->exp_cts * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->exp_b, B * db->exp_cts * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->qntt, B * db->N0 * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->sel, B * (nsel ? nsel : 1) * 2 * ctx->G.l_ep * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->mid_a, B * db->Nrest * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->mid_b, B * db->Nrest * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->part, kMuxSplitMaxBlocks * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->q_u64, B * 2 * abi_kn(ctx) * sizeof(u64)));
  HIP_TRY(ctx, hipMalloc((void **)&db->r_u64, B * 2 * kN * sizeof(u64)));
  db->ws_B = B;
  return OPIR_OK;
}

void rec(opir_ctx *ctx, int i, bool aux = false) {
  if (ctx->profiling) (void)hipEventRecord(ctx->ev[i], aux ? ctx->aux_stream : ctx->stream);
}

// Expansion (server.cpp:731-816) for B queries whose root cts are already in
// exp_a[q][0] (RNS coefficient form). Returns the buffer holding the leaves.
u32 *run_expand(opir_db *db, size_t B, const BatchPtrs &P) {
  opir_ctx *ctx = db->ctx;
  const size_t stride = db->exp_cts * kCtWords;
  u32 *in = db->exp_a, *out = db->exp_b;
  for (size_t L = 0; L < db->h; L++) {
    const size_t n = nodes_at_level(db->u, db->h, L);
    launch_expand_level(in, out, stride, stride, (int)n, (int)B, (int)L, P, ctx->G, ctx->S, ctx->stream,
                        db->part);
    std::swap(in, out);
  }
  return in;
}

}  // namespace capi

namespace {

// MUX tree (server.cpp:599-649) over NTT-form candidates in mid_a.
// Returns the buffer whose slot 0 (per query) holds the result.
u32 *run_other_dim(opir_db *db, size_t B, const u32 *sel) {
  opir_ctx *ctx = db->ctx;
  const size_t stride = db->Nrest * kCtWords;
  if (db->dims == 1) return db->mid_a;
  const size_t hs = db->dims - 1;
  const size_t sel_words = 2 * (size_t)ctx->G.l_ep * kCtWords;  // one selector
  const size_t sel_stride = (db->dims - 1) * sel_words;
  const size_t last = 2 * db->Nrest - (size_t{1} << hs);
  const size_t offset = db->Nrest - last;
  const size_t perfect = size_t{1} << (hs - 1);
  u32 *in = db->mid_a, *out = db->mid_b;
  // levels with few nodes spread each node over several workgroups
  auto level = [&](int nslots, int off, int half, const u32 *s) {
    const int Gs = mux_split_groups(nslots, off, (int)B, ctx->G);
    if (Gs)
      launch_mux_split(in, out, stride, nslots, off, half, (int)B, s, sel_stride, db->part, Gs, ctx->G, ctx->S,
                       ctx->stream);
    else
      launch_mux(in, out, stride, nslots, off, half, (int)B, s, sel_stride, ctx->G, ctx->S, ctx->stream);
    std::swap(in, out);
  };
  level((int)perfect, (int)offset, 0, sel);
  for (size_t a = 1; a < hs; a++) {
    const size_t half = (size_t{1} << (hs - a)) >> 1;
    level((int)half, -1, (int)half, sel + a * sel_words);
  }
  return in;
}

// Full answer pipeline for B queries already in db->q_u64 (device).
int answer_device(opir_db *db, size_t B, const uint64_t *ids, const u64 *d_q, u64 *d_r) {
  opir_ctx *ctx = db->ctx;
  int rc = check_answerable(db);
  if (rc) return rc;
  rc = ensure_ws(db, B);
  if (rc) return rc;
  BatchPtrs P;
  rc = upload_k1(ctx, ids, B, P);
  if (rc) return rc;
  for (size_t i = 0; i < B; i++) {
    const ClientKeys &k = ctx->clients()[ids[i]];
    if (k.h < db->h) return set_err(ctx, OPIR_EINVAL, "PirServer session lacks keys for its expansion height");
  }
  const size_t exp_stride = db->exp_cts * kCtWords;
  return run_answer(db, B, 4, P.bv_keys, d_q, d_r, [&]() -> int {
    rec(ctx, 0);
    // root ct per query -> exp_a[q][0]
    launch_split_u64_batch(d_q, 2 * abi_kn(ctx), db->exp_a, exp_stride, 2, (int)B, ctx->S, ctx->stream);
    LAUNCH_CHECK(ctx);
    u32 *leaves = run_expand(db, B, P);
    LAUNCH_CHECK(ctx);
    rec(ctx, 1);
    // selector completion on the auxiliary stream, concurrent with the scan
    // (two of its workgroups share a CU)
    const size_t nsel = db->dims - 1;
    const bool beside = complete_beside_scan(nsel * ctx->G.l_ep, B, false, 1024);
    if (beside) fork_aux(ctx);
    const size_t sel_stride = nsel * 2 * (size_t)ctx->G.l_ep * kCtWords;
    launch_complete(leaves + db->N0 * kCtWords, db->sel, exp_stride, sel_stride, (int)(nsel * ctx->G.l_ep), (int)B, P,
                    ctx->G, ctx->S, beside ? ctx->aux_stream : ctx->stream, beside ? nullptr : db->part);
    LAUNCH_CHECK(ctx);
    rec(ctx, 2, beside);
    launch_ct_ntt(leaves, db->qntt, exp_stride, db->N0 * kCtWords, (int)db->N0, (int)B, ctx->S, ctx->stream);
    launch_qfrag(db->qntt, db->qfrag, db->N0 * kCtWords, (int)B, db->geom, ctx->stream);
    rec(ctx, 5);
    const int sv = pick_scan_variant(db, B, 4, [&](int v) {
      launch_scan(db->d_db, db->qfrag, db->mid_a, db->Nrest * kCtWords, (int)B, db->geom, ctx->S, ctx->stream, v);
  });
  if (launch_scan(db->d_db, db->qfrag, db->mid_a, db->Nrest * kCtWords, (int)B, db->geom, ctx->S, ctx->stream, sv))
    return set_err(ctx, OPIR_EINVAL, "scan variant cannot take this batch size");
  LAUNCH_CHECK(ctx);
  rec(ctx, 6);
  rec(ctx, 3);
  if (beside) join_aux(ctx);
  u32 *res = run_other_dim(db, B, db->sel);
  LAUNCH_CHECK(ctx);
  rec(ctx, 4);
  launch_final(res, d_r, db->Nrest * kCtWords, (int)B, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  rec(ctx, 7);
  return OPIR_OK;
  });
}

int tune_scan(opir_db *db, size_t B) {
  opir_ctx *ctx = db->ctx;
  const int sv = pick_scan_variant(db, B, 4, [&](int v) {
    launch_scan(db->d_db, db->qfrag, db->mid_a, db->Nrest * kCtWords, (int)B, db->geom, ctx->S, ctx->stream, v);
  });
  return sv < 0 ? set_err(ctx, OPIR_EHIP, "scan autotune failed") : OPIR_OK;
}

}  // namespace

extern "C" {

const char *opir_last_error(const opir_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

static int ctx_create_k2(const opir_scheme *sch, int device, opir_ctx **out) {
  if (int rc = k2_validate_scheme(sch)) return rc;
  opir_ctx *ctx = new (std::nothrow) opir_ctx();
  if (!ctx) return set_err(nullptr, OPIR_ENOMEM, "out of host memory");
  ctx->sch = *sch;
  ctx->device = device;
  if (sch->noise_std_dev > 0) ctx->noise_sigma = sch->noise_std_dev;
  auto fail = [&](int code, const std::string &m) {
    opir_ctx_destroy(ctx);
    return set_err(nullptr, code, m);
  };
  if (hipSetDevice(device) != hipSuccess) return fail(OPIR_EHIP, "hipSetDevice failed");
  if (hipDeviceGetAttribute(&ctx->dev_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return fail(OPIR_EHIP, "device attribute query failed");
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) return fail(OPIR_EHIP, "stream create");
  ctx->stream = ctx->own_stream;
  if (hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking) != hipSuccess) return fail(OPIR_EHIP, "stream create");
  for (auto &e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(OPIR_EHIP, "event create");
  if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(OPIR_EHIP, "event create");
  int rc = k2_ctx_init(ctx, sch);
  if (!rc) rc = noalloc_init(ctx);
  if (rc) {
    const std::string m = g_create_error;
    return fail(rc, m);
  }
  *out = ctx;
  return OPIR_OK;
}

int opir_ctx_create(const opir_scheme *sch, int device, opir_ctx **out) {
  if (!sch || !out) return set_err(nullptr, OPIR_EINVAL, "null argument");
  *out = nullptr;
  // two RNS primes below 2^30 at N = 2048 (CONFIG_N2048_K2_MP) run on the
  // composite scheme's 32-bit kernels with the MP gadgets and mod switch
  // (which take primes below 2^29; larger ones are rejected there)
  const bool mp = !sch->composite && sch->q2 && sch->poly_degree == (uint32_t)kN && sch->q1 < (1ULL << 30) &&
                  sch->q2 < (1ULL << 30);
  if (!sch->composite && !mp) return ctx_create_k2(sch, device, out);
  if (sch->poly_degree != (uint32_t)kN)
    return set_err(nullptr, OPIR_EINVAL, "the composite scheme runs at N=2048 (CONFIG_N2048_K1_COMP)");
  const u64 q1 = sch->q1, q2 = sch->q2;
  // below 2^29: the forward transforms let values grow to 8q < 2^32 between
  // reductions (ntt_fwd_2048)
  if (q1 >= (1ULL << 29) || q2 >= (1ULL << 29) || q1 < 3 || q2 < 3 || q1 >= 2 * q2 || q2 >= 2 * q1 ||
      (q1 - 1) % (2 * kN) || (q2 - 1) % (2 * kN))
    return set_err(nullptr, OPIR_EINVAL, "limb primes must be < 2^29, within a factor 2, and 1 mod 2N");
  if (sch->l_ep == 0 || sch->l_key == 0 || sch->l_ks == 0 || sch->base_log2 == 0 || sch->base_log2 > 32 ||
      sch->base_log2_key == 0 || sch->base_log2_key > 32 || sch->bv_base_log2 == 0 || sch->bv_base_log2 > 32)
    return set_err(nullptr, OPIR_EINVAL, "invalid gadget parameters");
  {
    // Digits must be residues (|r| <= 2^(b-1) < q), their shifts must fit an
    // int64, and each lane block's 64-bit digit x key sums (NTT outputs kept
    // in [0, 4q)) must not overflow: terms * (4q - 1) * (q - 1) < 2^64.
    const uint64_t qmin = q1 < q2 ? q1 : q2, qmax = q1 < q2 ? q2 : q1;
    const uint64_t bmax = std::max({sch->base_log2, sch->base_log2_key, sch->bv_base_log2});
    const uint64_t terms = std::max({(uint64_t)sch->l_ep, (uint64_t)sch->l_key, (uint64_t)(sch->l_ks + 1) / 2});
    const unsigned __int128 worst = (unsigned __int128)terms * (4 * qmax - 1) * (qmax - 1);
    if ((1ULL << (bmax - 1)) >= qmin || (uint64_t)sch->base_log2 * (sch->l_ep - 1) > 62 ||
        (uint64_t)sch->base_log2_key * (sch->l_key - 1) > 62 || (uint64_t)sch->bv_base_log2 * (sch->l_ks - 1) > 62 ||
        worst >> 64)
      return set_err(nullptr, OPIR_EINVAL, "gadget too long or too wide for exact 64-bit digit sums");
  }
  if (sch->small_q < 2 || sch->small_q >= (1ULL << 32))
    return set_err(nullptr, OPIR_EINVAL, "small_q must fit 32 bits");
  if (mp && sch->small_q >= std::min(q1, q2))
    return set_err(nullptr, OPIR_EINVAL, "small_q must be below the RNS primes");
  opir_ctx *ctx = new (std::nothrow) opir_ctx();
  if (!ctx) return set_err(nullptr, OPIR_ENOMEM, "out of host memory");
  ctx->sch = *sch;
  ctx->device = device;
  ctx->K = mp ? 2 : 1;
  ctx->S.mp = mp ? 1 : 0;
  if (sch->noise_std_dev > 0) ctx->noise_sigma = sch->noise_std_dev;
  auto fail = [&](const std::string &m) {
    g_create_error = m;
    opir_ctx_destroy(ctx);
    return OPIR_EHIP;
  };
  if (hipSetDevice(device) != hipSuccess) return fail("hipSetDevice failed");
  if (hipDeviceGetAttribute(&ctx->dev_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return fail("device attribute query failed");
  if (noalloc_init(ctx)) return fail(g_create_error);
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) return fail("stream create");
  ctx->stream = ctx->own_stream;
  if (hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking) != hipSuccess) return fail("stream create");
  for (auto &e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail("event create");
  if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail("event create");
  // twiddle tables per limb
  std::vector<u32> tab(2 * 4 * kN);
  const u64 Q = q1 * q2;
  for (int l = 0; l < 2; l++) {
    const u64 q = l == 0 ? q1 : q2;
    // composite: one root of Z_Q seen mod each prime; MP: a root per prime
    const u64 w = (mp && l == 1 ? sch->ntt_root2 : sch->ntt_root) % q;
    if (powmod_h(w, kN, q) != q - 1) {
      opir_ctx_destroy(ctx);
      return set_err(nullptr, OPIR_EINVAL, "ntt_root is not a primitive 2N-th root of unity");
    }
    u64 wi;
    inv_h(w, q, wi);
    u32 *psi = tab.data() + (size_t)l * 4 * kN, *psi_s = psi + kN, *ipsi = psi + 2 * kN, *ipsi_s = psi + 3 * kN;
    u64 p = 1, pi = 1;
    for (size_t i = 0; i < (size_t)kN; i++) {
      const size_t r = brv(i, kLogN);
      psi[r] = (u32)p; psi_s[r] = shoup_h((u32)p, (u32)q);
      ipsi[r] = (u32)pi; ipsi_s[r] = shoup_h((u32)pi, (u32)q);
      p = mulmod_h(p, w, q);
      pi = mulmod_h(pi, wi, q);
    }
    LimbConst &L = ctx->S.limb[l];
    L.q = (u32)q; L.q2 = (u32)(2 * q);
    L.r32 = (u32)((1ULL << 32) % q); L.r32s = shoup_h(L.r32, (u32)q);
    L.one_s = (u32)((1ULL << 32) / q);
    u64 ni;
    inv_h(kN, q, ni);
    L.ninv = (u32)ni; L.ninv_s = shoup_h((u32)ni, (u32)q);
  }
  if (hipMalloc((void **)&ctx->d_tables, tab.size() * sizeof(u32)) != hipSuccess) return fail("hipMalloc tables");
  if (hipMemcpy(ctx->d_tables, tab.data(), tab.size() * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess)
    return fail("hipMemcpy tables");
  for (int l = 0; l < 2; l++) {
    LimbConst &L = ctx->S.limb[l];
    const u32 *base = ctx->d_tables + (size_t)l * 4 * kN;
    L.psi = base; L.psi_s = base + kN; L.ipsi = base + 2 * kN; L.ipsi_s = base + 3 * kN;
  }
  ctx->S.Q = Q;
  ctx->S.half_Q = Q >> 1;
  ctx->S.q1_mod_q2 = (u32)(q1 % q2);
  u64 q1inv;
  if (!inv_h(q1 % q2, q2, q1inv)) {
    opir_ctx_destroy(ctx);
    return set_err(nullptr, OPIR_EINVAL, "q1, q2 not coprime");
  }
  ctx->S.q1inv = (u32)q1inv;
  ctx->S.q1inv_s = shoup_h((u32)q1inv, (u32)q2);
  ctx->S.q2_mod_q1 = (u32)(q2 % q1);
  ctx->S.small_q = sch->small_q;
  {
    int bits = 0;
    while (bits < 64 && (Q >> bits)) bits++;
    ctx->S.q_recip = (u64)(((u128)1 << (63 + bits)) / Q);
    ctx->S.q_shift = bits - 1;
  }
  ctx->G = GadgetCfg{(int)sch->l_ep, (int)sch->base_log2, (int)sch->l_key, (int)sch->base_log2_key, (int)sch->l_ks,
                     (int)sch->bv_base_log2};
  *out = ctx;
  return OPIR_OK;
}

void opir_ctx_destroy(opir_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->aux_stream) (void)hipStreamSynchronize(ctx->aux_stream);
  for (auto &kv : ctx->own_clients) {
    (void)hipFree(kv.second.bv);
    (void)hipFree(kv.second.gsw);
  }
  for (int i = 0; i < kRing; i++) {
    if (ctx->ring_ev[i]) (void)hipEventDestroy(ctx->ring_ev[i]);
    if (ctx->pin_ptrs[i]) (void)hipHostFree(ctx->pin_ptrs[i]);
    if (ctx->dev_ptrs[i]) (void)hipFree(ctx->dev_ptrs[i]);
  }
  for (auto &e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  if (ctx->d_tables) (void)hipFree(ctx->d_tables);
  k2_ctx_free(ctx);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

int opir_ctx_clone(opir_ctx *parent, opir_ctx **out) {
  if (!parent || !out) return OPIR_EINVAL;
  *out = nullptr;
  opir_ctx *ctx = new (std::nothrow) opir_ctx();
  if (!ctx) return set_err(parent, OPIR_ENOMEM, "out of host memory");
  // scheme constants and device tables are the parent's (it must outlive the clone)
  ctx->device = parent->device;
  ctx->sch = parent->sch;
  ctx->k2 = parent->k2;
  ctx->K = parent->K;
  ctx->noise_sigma = parent->noise_sigma;
  ctx->N = parent->N;
  ctx->S = parent->S;
  ctx->G = parent->G;
  ctx->S64 = parent->S64;
  ctx->G64 = parent->G64;
  ctx->scan_variant = parent->scan_variant;
  ctx->scratch = parent->scratch;
  ctx->dev_cus = parent->dev_cus;
  ctx->key_owner = parent->key_owner ? parent->key_owner : parent;
  auto fail = [&](const char *m) {
    opir_ctx_destroy(ctx);
    return set_err(parent, OPIR_EHIP, m);
  };
  if (hipSetDevice(ctx->device) != hipSuccess) return fail("hipSetDevice failed");
  // the parent's compute-unit mask carries over to the clone's streams
  ctx->cu_mask = parent->cu_mask;
  auto make = [&](hipStream_t *st) {
    return ctx->cu_mask.empty() ? hipStreamCreateWithFlags(st, hipStreamNonBlocking)
                                : hipExtStreamCreateWithCUMask(st, (uint32_t)ctx->cu_mask.size(), ctx->cu_mask.data());
  };
  if (make(&ctx->own_stream) != hipSuccess) return fail("stream create");
  ctx->stream = ctx->own_stream;
  if (make(&ctx->aux_stream) != hipSuccess) return fail("stream create");
  for (auto &e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail("event create");
  if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail("event create");
  *out = ctx;
  return OPIR_OK;
}

int opir_db_view(opir_db *src, opir_ctx *ctx, opir_db **out) {
  if (!src || !ctx || !out) return OPIR_EINVAL;
  *out = nullptr;
  const opir_ctx *a = src->ctx->key_owner ? src->ctx->key_owner : src->ctx;
  const opir_ctx *b = ctx->key_owner ? ctx->key_owner : ctx;
  if (a != b) return set_err(ctx, OPIR_EINVAL, "a database view needs a context of the same key family (opir_ctx_clone)");
  opir_db *db = new (std::nothrow) opir_db();
  if (!db) return set_err(ctx, OPIR_ENOMEM, "out of host memory");
  db->ctx = ctx;
  db->owns_db = false;
  db->N0 = src->N0; db->Nrest = src->Nrest; db->dims = src->dims; db->h = src->h;
  db->num_pt = src->num_pt; db->u = src->u; db->exp_cts = src->exp_cts;
  db->geom = src->geom;
  db->d_db = src->d_db;
  for (int i = 0; i < 16; i++) db->tuned_scan[i] = src->tuned_scan[i];
  *out = db;
  return OPIR_OK;
}

int opir_ctx_set_stream(opir_ctx *ctx, void *s) {
  if (!ctx) return OPIR_EINVAL;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  return OPIR_OK;
}

int opir_ctx_set_cu_mask(opir_ctx *ctx, const uint32_t *mask, size_t words) {
  if (!ctx || (words && !mask) || words > 64) return OPIR_EINVAL;
  if (words) {
    bool any = false;
    for (size_t i = 0; i < words; i++) any = any || mask[i];
    if (!any) return set_err(ctx, OPIR_EINVAL, "CU mask selects no compute unit");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->aux_stream));
  // captured answer graphs hold no stream: they stay valid on the new ones
  hipStream_t own = nullptr, aux = nullptr;
  auto make = [&](hipStream_t *s) {
    return words ? hipExtStreamCreateWithCUMask(s, (uint32_t)words, mask)
                 : hipStreamCreateWithFlags(s, hipStreamNonBlocking);
  };
  if (make(&own) != hipSuccess) return set_err(ctx, OPIR_EHIP, "stream create");
  if (make(&aux) != hipSuccess) {
    (void)hipStreamDestroy(own);
    return set_err(ctx, OPIR_EHIP, "stream create");
  }
  const bool own_current = ctx->stream == ctx->own_stream;
  (void)hipStreamDestroy(ctx->own_stream);
  (void)hipStreamDestroy(ctx->aux_stream);
  ctx->own_stream = own;
  ctx->aux_stream = aux;
  if (own_current) ctx->stream = own;
  ctx->cu_mask.assign(mask, mask + words);
  return OPIR_OK;
}

int opir_ctx_synchronize(opir_ctx *ctx) {
  if (!ctx) return OPIR_EINVAL;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return OPIR_OK;
}

int opir_ctx_set_profiling(opir_ctx *ctx, int on) {
  if (!ctx) return OPIR_EINVAL;
  ctx->profiling = on != 0;
  return OPIR_OK;
}

int opir_ctx_set_scan_variant(opir_ctx *ctx, int variant) {
  if (!ctx) return OPIR_EINVAL;
  if (variant < -2 || variant >= scan_variant_count(ctx->k2 ? 8 : 4))
    return set_err(ctx, OPIR_EINVAL, "scan variant out of range");
  ctx->scan_variant = variant;
  return OPIR_OK;
}

int opir_ctx_set_graphs(opir_ctx *ctx, int on) {
  if (!ctx) return OPIR_EINVAL;
  ctx->graphs = on != 0;
  return OPIR_OK;
}

int opir_db_tuned_scan_variant(const opir_db *db, size_t B, int *out) {
  if (!db || !out || B == 0) return OPIR_EINVAL;
  *out = db->tuned_scan[scan_class(B)];
  return OPIR_OK;
}

int opir_scan_variant_count(const opir_ctx *ctx, int *out) {
  if (!ctx || !out) return OPIR_EINVAL;
  *out = scan_variant_count(ctx->k2 ? 8 : 4);
  return OPIR_OK;
}

int opir_last_stage_ms(opir_ctx *ctx, float *o) {
  if (!ctx || !o) return OPIR_EINVAL;
  if (!ctx->profiling) return set_err(ctx, OPIR_EINVAL, "profiling is off");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipEventElapsedTime(&o[0], ctx->ev[0], ctx->ev[1]));  // Expand
  HIP_TRY(ctx, hipEventElapsedTime(&o[1], ctx->ev[1], ctx->ev[2]));  // Convert (selector completion)
  // First dim (query NTT + scan): concurrent with Convert, or after it when
  // the completion runs ahead of the scan (complete_beside_scan)
  HIP_TRY(ctx, hipEventElapsedTime(&o[2], ctx->ev[1], ctx->ev[3]));
  HIP_TRY(ctx, hipEventElapsedTime(&o[3], ctx->ev[3], ctx->ev[4]));  // Other dim
  HIP_TRY(ctx, hipEventElapsedTime(&o[4], ctx->ev[4], ctx->ev[7]));  // Modulus switching
  HIP_TRY(ctx, hipEventElapsedTime(&o[5], ctx->ev[5], ctx->ev[6]));  // scan kernel
  return OPIR_OK;
}

int opir_db_create(opir_ctx *ctx, size_t N0, size_t Nrest, size_t dims, size_t h, opir_db **out) {
  if (!ctx || !out) return OPIR_EINVAL;
  *out = nullptr;
  if (N0 == 0 || Nrest == 0 || dims == 0 || h > 11) return set_err(ctx, OPIR_EINVAL, "invalid layout");
  const size_t cap = size_t{1} << h;
  const size_t l_ep = ctx->k2 ? (size_t)ctx->G64.l_ep : (size_t)ctx->G.l_ep;
  const size_t u = N0 + l_ep * (dims - 1);
  if (u > cap) return set_err(ctx, OPIR_EINVAL, "layout does not fit the expansion capacity");
  if (dims > 1 && (dims - 1 > 40 || Nrest > (size_t{1} << (dims - 1))))
    return set_err(ctx, OPIR_EINVAL, "more candidates than the selector tree addresses");
  opir_db *db = new (std::nothrow) opir_db();
  if (!db) return set_err(ctx, OPIR_ENOMEM, "out of host memory");
  db->ctx = ctx;
  db->N0 = N0; db->Nrest = Nrest; db->dims = dims; db->h = h;
  db->geom = make_db_geom(N0, Nrest);
  db->num_pt = N0 * Nrest;
  db->u = u;
  db->exp_cts = u + 2;
  if (ctx->k2) {
    const int rc = k2_db_init(db);
    if (rc) {
      delete db;
      return rc;
    }
    *out = db;
    return OPIR_OK;
  }
  const size_t bytes = db_bytes(db->geom);
  if (hipMalloc((void **)&db->d_db, bytes) != hipSuccess) {
    delete db;
    return set_err(ctx, OPIR_ENOMEM, "hipMalloc database (" + std::to_string(bytes) + " bytes)");
  }
  if (hipMemsetAsync(db->d_db, 0, bytes, ctx->stream) != hipSuccess) {
    (void)hipFree(db->d_db);
    delete db;
    return set_err(ctx, OPIR_EHIP, "hipMemset database");
  }
  *out = db;
  return OPIR_OK;
}

void opir_db_destroy(opir_db *db) {
  if (!db) return;
  (void)hipStreamSynchronize(db->ctx->stream);
  drop_graphs(db);
  k2_db_free(db);
  for (u32 *p : {db->exp_a, db->exp_b, db->qntt, db->sel, db->mid_a, db->mid_b, db->part})
    if (p) (void)hipFree(p);
  if (db->d_db && db->owns_db) (void)hipFree(db->d_db);
  if (db->qfrag) (void)hipFree(db->qfrag);
  if (db->q_u64) (void)hipFree(db->q_u64);
  if (db->r_u64) (void)hipFree(db->r_u64);
  delete db;
}

int opir_db_load_plain(opir_db *db, const uint16_t *coeffs, size_t first_pt, size_t count) {
  if (!db || (!coeffs && count)) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2) {
    std::vector<uint64_t> w(coeffs, coeffs + count * ctx->N);
    return k2_load_plain(db, w.data(), first_pt, count);
  }
  if (first_pt > db->num_pt || count > db->num_pt - first_pt)
    return set_err(ctx, OPIR_EINVAL, "PirServer::load_data logical plaintext count exceeds its layout");
  // server.cpp:129-137: every coefficient must be below t
  {
    const u64 t = ctx->sch.plain_mod;
    const size_t nv = count * kN;
    for (size_t i = 0; i < nv; i++)
      if (coeffs[i] >= t) return set_err(ctx, OPIR_EINVAL, "plaintext coefficient not reduced mod t");
  }
  const size_t chunk = 8192;
  uint16_t *d_pts = nullptr;
  u32 *d_tmp = nullptr;
  HIP_TRY(ctx, hipMalloc((void **)&d_pts, chunk * kN * sizeof(uint16_t)));
  if (hipMalloc((void **)&d_tmp, 2 * chunk * kN * sizeof(u32)) != hipSuccess) {
    (void)hipFree(d_pts);
    return set_err(ctx, OPIR_ENOMEM, "hipMalloc staging");
  }
  int rc = OPIR_OK;
  for (size_t off = 0; off < count && rc == OPIR_OK; off += chunk) {
    const size_t n = std::min(chunk, count - off);
    hipError_t e = hipMemcpyAsync(d_pts, coeffs + off * kN, n * kN * sizeof(uint16_t), hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e != hipSuccess) { rc = set_err(ctx, OPIR_EHIP, "upload plaintexts"); break; }
    launch_db_ntt_plain(d_pts, d_tmp, n, ctx->S, ctx->stream);
    launch_db_scatter(d_tmp, db->d_db, first_pt + off, n, db->geom, ctx->stream);
    e = hipGetLastError();
    if (e != hipSuccess) { rc = set_err(ctx, OPIR_EHIP, std::string("db ingest: ") + hipGetErrorString(e)); break; }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { rc = set_err(ctx, OPIR_EHIP, std::string("db ingest: ") + hipGetErrorString(e)); break; }
  }
  (void)hipFree(d_pts);
  (void)hipFree(d_tmp);
  return rc;
}

int opir_db_load_plain_u64(opir_db *db, const uint64_t *coeffs, size_t first_pt, size_t count) {
  if (!db || (!coeffs && count)) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2) return k2_load_plain(db, coeffs, first_pt, count);
  std::vector<uint16_t> w(count * kN);
  for (size_t i = 0; i < w.size(); i++) {
    if (coeffs[i] >= ctx->sch.plain_mod) return set_err(ctx, OPIR_EINVAL, "plaintext coefficient not reduced mod t");
    w[i] = (uint16_t)coeffs[i];
  }
  return opir_db_load_plain(db, w.data(), first_pt, count);
}

int opir_db_gen_random(opir_db *db, uint64_t seed) {
  if (!db) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  const u64 t = ctx->sch.plain_mod;
  const size_t N = ctx->N, chunk = ctx->k2 ? 4096 : 8192;
  DevBuf pts, tmp;
  DEV_ALLOC(ctx, pts, chunk * N * (ctx->k2 ? sizeof(u64) : sizeof(uint16_t)));
  DEV_ALLOC(ctx, tmp, 2 * chunk * N * (ctx->k2 ? sizeof(u64) : sizeof(u32)));
  for (size_t off = 0; off < db->num_pt; off += chunk) {
    const size_t n = std::min(chunk, db->num_pt - off);
    if (ctx->k2) {
      launch_gen_plain_u64((u64 *)pts.p, off, n, (int)N, seed, t, ctx->stream);
      if (int rc = k2_ingest(db, (const u64 *)pts.p, (u64 *)tmp.p, off, n)) return rc;
    } else {
      launch_gen_plain_u16((uint16_t *)pts.p, off, n, (int)N, seed, t, ctx->stream);
      launch_db_ntt_plain((const uint16_t *)pts.p, (u32 *)tmp.p, n, ctx->S, ctx->stream);
      launch_db_scatter((const u32 *)tmp.p, db->d_db, off, n, db->geom, ctx->stream);
    }
    LAUNCH_CHECK(ctx);
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return OPIR_OK;
}

int opir_db_load_ntt(opir_db *db, const uint32_t *lo, const uint32_t *hi) {
  if (!db || !lo || !hi) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2 || ctx->S.mp) return set_err(ctx, OPIR_EINVAL, "the composite (lo, hi) database layout exists only for K=1");
  const size_t n = (size_t)kN * db->num_pt;
  // residues must be reduced: the digit planes assume values below 2^30
  const u32 q1 = ctx->S.limb[0].q, q2 = ctx->S.limb[1].q;
  for (size_t i = 0; i < n; i++)
    if (lo[i] >= q1 || hi[i] >= q2)
      return set_err(ctx, OPIR_EINVAL, "database residue not reduced mod its limb prime");
  u32 *d_ref = nullptr;
  HIP_TRY(ctx, hipMalloc((void **)&d_ref, 2 * n * sizeof(u32)));
  int rc = OPIR_OK;
  if (hipMemcpy(d_ref, lo, n * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_ref + n, hi, n * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess)
    rc = set_err(ctx, OPIR_EHIP, "upload reference-layout database");
  if (!rc) {
    launch_db_from_ref(d_ref, db->d_db, db->geom, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = set_err(ctx, OPIR_EHIP, std::string("db transpose: ") + hipGetErrorString(e));
  }
  (void)hipFree(d_ref);
  return rc;
}

static void replace_client(opir_ctx *ctx, uint64_t id, const ClientKeys &k) {
  auto it = ctx->clients().find(id);
  if (it != ctx->clients().end()) {
    (void)hipStreamSynchronize(ctx->stream);  // queued answers may still read the old keys
    (void)hipFree(it->second.bv);
    (void)hipFree(it->second.gsw);
  }
  ctx->clients()[id] = k;
}

int opir_client_set_keys(opir_ctx *ctx, uint64_t id, const uint64_t *bv, size_t h, const uint64_t *gsw) {
  if (!ctx) return OPIR_EINVAL;
  if (ctx->key_owner) return set_err(ctx, OPIR_EINVAL, "a cloned context reads its parent's client keys");
  if ((!bv && h) || !gsw) return set_err(ctx, OPIR_EINVAL, "PirServer session key bundle cannot be null");
  if (h > 12) return set_err(ctx, OPIR_EINVAL, "BV key count out of range");
  if (ctx->k2) {
    ClientKeys k;
    const int rc = k2_set_keys(ctx, bv, h, gsw, k);
    if (rc) return rc;
    replace_client(ctx, id, k);
    return OPIR_OK;
  }
  if (h > 11) return set_err(ctx, OPIR_EINVAL, "BV key count out of range");
  const size_t L = ctx->G.l_ks, lk = ctx->G.l_key, KN = abi_kn(ctx);
  int rc = h ? validate_polys(ctx, bv, h * L * 2, "BV key") : OPIR_OK;
  if (rc) return rc;
  rc = validate_polys(ctx, gsw, 2 * lk * 2, "RGSW key");
  if (rc) return rc;
  std::vector<u32> hbv(h * L * kCtWords), hg(2 * lk * kCtWords);
  for (size_t r = 0; r < h * L; r++)
    for (int ab = 0; ab < 2; ab++) {      // input order a, b
      const int out = ab == 0 ? 1 : 0;    // device order b (-> c0), a (-> c1)
      poly_to_rns(ctx, bv + (r * 2 + ab) * KN, hbv.data() + r * kCtWords + (size_t)out * 2 * kN);
    }
  for (size_t r = 0; r < 2 * lk; r++)
    for (int p = 0; p < 2; p++) poly_to_rns(ctx, gsw + (r * 2 + p) * KN, hg.data() + r * kCtWords + (size_t)p * 2 * kN);
  ClientKeys k;
  k.h = h;
  if (h) HIP_TRY(ctx, hipMalloc((void **)&k.bv, hbv.size() * sizeof(u32)));
  if (hipMalloc((void **)&k.gsw, hg.size() * sizeof(u32)) != hipSuccess) {
    (void)hipFree(k.bv);
    return set_err(ctx, OPIR_ENOMEM, "hipMalloc keys");
  }
  if ((h && hipMemcpy(k.bv, hbv.data(), hbv.size() * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(k.gsw, hg.data(), hg.size() * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(k.bv);
    (void)hipFree(k.gsw);
    return set_err(ctx, OPIR_EHIP, "upload keys");
  }
  replace_client(ctx, id, k);
  return OPIR_OK;
}

int opir_client_remove(opir_ctx *ctx, uint64_t id) {
  if (!ctx) return OPIR_EINVAL;
  if (ctx->key_owner) return set_err(ctx, OPIR_EINVAL, "a cloned context reads its parent's client keys");
  auto it = ctx->clients().find(id);
  if (it == ctx->clients().end()) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  (void)hipFree(it->second.bv);
  (void)hipFree(it->second.gsw);
  ctx->clients().erase(it);
  return OPIR_OK;
}

int opir_db_workspace_bytes(const opir_db *db, size_t B, size_t *out) {
  if (!db || !out) return OPIR_EINVAL;
  *out = db->ctx->k2 ? k2_ws_bytes(db, B) : ws_bytes(db, B);
  return OPIR_OK;
}

int opir_db_trim(opir_db *db, size_t B) {
  if (!db) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (db->ws_B <= B) return OPIR_OK;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // queued work may still use the buffers
  drop_graphs(db);
  if (ctx->k2)
    k2_release_ws(db);
  else
    release_ws(db);
  return B ? (ctx->k2 ? k2_ensure_ws(db, B) : ensure_ws(db, B)) : OPIR_OK;
}

int opir_db_reserve(opir_db *db, size_t B) {
  if (!db) return OPIR_EINVAL;
  if (!B) return OPIR_OK;
  opir_ctx *ctx = db->ctx;
  int rc = ctx->k2 ? k2_ensure_ws(db, B) : ensure_ws(db, B);
  if (rc) return rc;
  // the host staging of opir_answer_batch and the key-pointer ring
  const size_t vals = 2 * (size_t)ctx->K * ctx->N;
  if ((rc = ensure_pin(ctx, B * vals * sizeof(u64)))) return rc;
  if ((rc = ensure_ring(ctx, B))) return rc;
  if (ctx->scan_variant == -2) {
    // settle the scan shape of every batch-size class up to B (the tuner
    // times the shapes on the reserved, not yet meaningful, operands)
    for (size_t b = 1;; b = std::min(2 * b, B)) {
      rc = ctx->k2 ? k2_tune_scan(db, b) : tune_scan(db, b);
      if (rc) return rc;
      if (b == B) break;
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return OPIR_OK;
}

int opir_client_key_bytes(const opir_ctx *ctx, size_t h, size_t *out) {
  if (!ctx || !out) return OPIR_EINVAL;
  *out = ctx->k2 ? (h * ctx->G64.l_ks + 2 * (size_t)ctx->G64.l_key) * k2::ct_values(ctx->S64) * sizeof(u64)
                 : (h * ctx->G.l_ks + 2 * (size_t)ctx->G.l_key) * kCtWords * sizeof(u32);
  return OPIR_OK;
}

int opir_ctx_resident_key_bytes(const opir_ctx *ctx, size_t *out) {
  if (!ctx || !out) return OPIR_EINVAL;
  size_t tot = 0;
  for (const auto &kv : ctx->clients()) {
    size_t kb = 0;
    opir_client_key_bytes(ctx, kv.second.h, &kb);
    tot += kb;
  }
  *out = tot;
  return OPIR_OK;
}

int opir_answer_batch_device(opir_db *db, size_t B, const uint64_t *ids, const uint64_t *d_q, uint64_t *d_r) {
  if (!db || !ids || !d_q || !d_r) return OPIR_EINVAL;
  if (B == 0) return OPIR_OK;
  return db->ctx->k2 ? k2_answer_device(db, B, ids, d_q, d_r) : answer_device(db, B, ids, d_q, d_r);
}

int opir_answer_batch(opir_db *db, size_t B, const uint64_t *ids, const uint64_t *queries, uint64_t *resps) {
  if (!db || !ids || !queries || !resps) return OPIR_EINVAL;
  if (B == 0) return OPIR_OK;
  opir_ctx *ctx = db->ctx;
  const size_t N = ctx->N, qvals = 2 * (size_t)ctx->K * N, rvals = 2 * N;
  int rc = ctx->k2 ? k2_validate_rns(ctx, queries, B * 2, "query") : validate_polys(ctx, queries, B * 2, "query");
  if (rc) return rc;
  for (size_t i = 0; i < B; i++)
    if (!ctx->clients().count(ids[i])) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(ids[i]));
  rc = ctx->k2 ? k2_ensure_ws(db, B) : ensure_ws(db, B);
  if (rc) return rc;
  const size_t qbytes = B * qvals * sizeof(u64), rbytes = B * rvals * sizeof(u64);
  rc = ensure_pin(ctx, std::max(qbytes, rbytes));
  if (rc) return rc;
  std::memcpy(ctx->pin, queries, qbytes);
  HIP_TRY(ctx, hipMemcpyAsync(db->q_u64, ctx->pin, qbytes, hipMemcpyHostToDevice, ctx->stream));
  rc = ctx->k2 ? k2_answer_device(db, B, ids, db->q_u64, db->r_u64) : answer_device(db, B, ids, db->q_u64, db->r_u64);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->pin, db->r_u64, rbytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(resps, ctx->pin, rbytes);
  return OPIR_OK;
}

int opir_answer(opir_db *db, uint64_t id, const uint64_t *query, uint64_t *resp) {
  return opir_answer_batch(db, 1, &id, query, resp);
}

// ---------------------------------------------------------------------------
// Stage entry points (host buffers)
// ---------------------------------------------------------------------------

static int ntt_entry(opir_ctx *ctx, uint64_t *data, size_t count, bool inverse) {
  if (!ctx || (!data && count)) return OPIR_EINVAL;
  if (!count) return OPIR_OK;
  if (ctx->k2) return k2_ntt(ctx, data, count, inverse);
  int rc = validate_polys(ctx, data, count, "NTT input");
  if (rc) return rc;
  DevBuf d;
  DEV_ALLOC(ctx, d, count * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(d.p, data, count * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_ntt_u64((u64 *)d.p, count, inverse, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(data, d.p, count * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_ntt_fwd(opir_ctx *ctx, uint64_t *data, size_t count) { return ntt_entry(ctx, data, count, false); }
int opir_ntt_inv(opir_ctx *ctx, uint64_t *data, size_t count) { return ntt_entry(ctx, data, count, true); }

int opir_expand(opir_db *db, uint64_t id, const uint64_t *query, uint64_t *out) {
  if (!db || !query || !out) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2) return k2_expand(db, id, query, out);
  int rc = validate_polys(ctx, query, 2, "query");
  if (rc) return rc;
  if (!ctx->clients().count(id)) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  if (ctx->clients()[id].h < db->h) return set_err(ctx, OPIR_EINVAL, "PirServer session lacks keys for its expansion height");
  rc = ensure_ws(db, 1);
  if (rc) return rc;
  BatchPtrs P;
  rc = upload_k1(ctx, &id, 1, P);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpy(db->q_u64, query, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_split_u64(db->q_u64, db->exp_a, 2, ctx->S, ctx->stream);
  u32 *leaves = run_expand(db, 1, P);
  LAUNCH_CHECK(ctx);
  DevBuf o;
  DEV_ALLOC(ctx, o, db->u * 2 * abi_kn(ctx) * sizeof(u64));
  launch_join_u64(leaves, (u64 *)o.p, db->u * 2, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, db->u * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_complete(opir_db *db, uint64_t id, const uint64_t *expanded, uint64_t *sel) {
  if (!db || !expanded || !sel) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2) return k2_complete(db, id, expanded, sel);
  const size_t nsel = db->dims - 1;
  if (!nsel) return OPIR_OK;
  if (!ctx->clients().count(id)) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  int rc = validate_polys(ctx, expanded, db->u * 2, "expanded");
  if (rc) return rc;
  rc = ensure_ws(db, 1);
  if (rc) return rc;
  BatchPtrs P;
  rc = upload_k1(ctx, &id, 1, P);
  if (rc) return rc;
  DevBuf e;
  DEV_ALLOC(ctx, e, db->u * 2 * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(e.p, expanded, db->u * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_split_u64((u64 *)e.p, db->exp_a, db->u * 2, ctx->S, ctx->stream);
  const size_t sel_stride = nsel * 2 * (size_t)ctx->G.l_ep * kCtWords;
  launch_complete(db->exp_a + db->N0 * kCtWords, db->sel, 0, sel_stride, (int)(nsel * ctx->G.l_ep), 1, P, ctx->G,
                  ctx->S, ctx->stream);
  const size_t npoly = nsel * 2 * ctx->G.l_ep * 2;
  DevBuf o;
  DEV_ALLOC(ctx, o, npoly * abi_kn(ctx) * sizeof(u64));
  launch_join_u64(db->sel, (u64 *)o.p, npoly, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(sel, o.p, npoly * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_first_dim(opir_db *db, const uint64_t *query, uint64_t *out) {
  if (!db || !query || !out) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2) return k2_first_dim(db, query, out);
  int rc = validate_polys(ctx, query, db->N0 * 2, "query");
  if (rc) return rc;
  rc = ensure_ws(db, 1);
  if (rc) return rc;
  DevBuf e;
  DEV_ALLOC(ctx, e, db->N0 * 2 * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(e.p, query, db->N0 * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_split_u64((u64 *)e.p, db->exp_a, db->N0 * 2, ctx->S, ctx->stream);
  launch_ct_ntt(db->exp_a, db->qntt, 0, 0, (int)db->N0, 1, ctx->S, ctx->stream);
  launch_qfrag(db->qntt, db->qfrag, 0, 1, db->geom, ctx->stream);
  if (launch_scan(db->d_db, db->qfrag, db->mid_a, 0, 1, db->geom, ctx->S, ctx->stream,
                  ctx->scan_variant == -2 ? -1 : ctx->scan_variant))
    return set_err(ctx, OPIR_EINVAL, "scan variant cannot take this batch size");
  DevBuf o;
  DEV_ALLOC(ctx, o, db->Nrest * 2 * abi_kn(ctx) * sizeof(u64));
  launch_ct_to_coef_u64(db->mid_a, (u64 *)o.p, db->Nrest, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, db->Nrest * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_other_dim(opir_db *db, const uint64_t *mid, const uint64_t *sel, uint64_t *out) {
  if (!db || !mid || !out || (db->dims > 1 && !sel)) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (ctx->k2) return k2_other_dim(db, mid, sel, out);
  if (int e = check_answerable(db)) return e;
  const size_t nsel = db->dims - 1;
  const size_t sel_polys = nsel * 2 * ctx->G.l_ep * 2;
  int rc = validate_polys(ctx, mid, db->Nrest * 2, "candidates");
  if (rc) return rc;
  if (nsel) rc = validate_polys(ctx, sel, sel_polys, "selectors");
  if (rc) return rc;
  rc = ensure_ws(db, 1);
  if (rc) return rc;
  DevBuf m;
  DEV_ALLOC(ctx, m, db->Nrest * 2 * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(m.p, mid, db->Nrest * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  // stage coefficient-form candidates in mid_b (Nrest cts), NTT into mid_a
  launch_split_u64((u64 *)m.p, db->mid_b, db->Nrest * 2, ctx->S, ctx->stream);
  launch_ct_ntt(db->mid_b, db->mid_a, 0, 0, (int)db->Nrest, 1, ctx->S, ctx->stream);
  DevBuf s;
  if (nsel) {
    DEV_ALLOC(ctx, s, sel_polys * abi_kn(ctx) * sizeof(u64));
    HIP_TRY(ctx, hipMemcpy(s.p, sel, sel_polys * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
    launch_split_u64((u64 *)s.p, db->sel, sel_polys, ctx->S, ctx->stream);
  }
  u32 *res = run_other_dim(db, 1, db->sel);
  DevBuf o;
  DEV_ALLOC(ctx, o, 2 * abi_kn(ctx) * sizeof(u64));
  launch_ct_to_coef_u64(res, (u64 *)o.p, 1, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_ext_prod(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t b, const uint64_t *ct, uint64_t *out) {
  if (!ctx || !gsw || !ct || !out) return OPIR_EINVAL;
  if (ctx->k2) return k2_ext_prod(ctx, gsw, l, b, ct, out);
  if (l == 0 || l > 64 || b == 0 || b > 32) return set_err(ctx, OPIR_EINVAL, "invalid gadget");
  int rc = validate_polys(ctx, gsw, 2 * l * 2, "RGSW");
  if (rc) return rc;
  rc = validate_polys(ctx, ct, 2, "ciphertext");
  if (rc) return rc;
  DevBuf g64, g32, c64, c32, o32, o64;
  DEV_ALLOC(ctx, g64, 2 * l * 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, g32, 2 * l * kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, c64, 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, c32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o64, 2 * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(g64.p, gsw, 2 * l * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(c64.p, ct, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_split_u64((u64 *)g64.p, (u32 *)g32.p, 2 * l * 2, ctx->S, ctx->stream);
  launch_split_u64((u64 *)c64.p, (u32 *)c32.p, 2, ctx->S, ctx->stream);
  launch_ext_prod((u32 *)g32.p, (int)l, (int)b, (u32 *)c32.p, (u32 *)o32.p, ctx->S, ctx->stream);
  launch_join_u64((u32 *)o32.p, (u64 *)o64.p, 2, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o64.p, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_ext_prod_mux(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t b, const uint64_t *x, const uint64_t *y,
                      uint64_t *out) {
  if (!ctx || !gsw || !x || !y || !out) return OPIR_EINVAL;
  if (ctx->k2) return k2_ext_prod_mux(ctx, gsw, l, b, x, y, out);
  if (l == 0 || l > 64 || b == 0 || b > 32) return set_err(ctx, OPIR_EINVAL, "invalid gadget");
  int rc = validate_polys(ctx, gsw, 2 * l * 2, "RGSW");
  if (rc) return rc;
  if ((rc = validate_polys(ctx, x, 2, "ciphertext x"))) return rc;
  if ((rc = validate_polys(ctx, y, 2, "ciphertext y"))) return rc;
  DevBuf g64, g32, xd, yd, c32, o32, o64;
  DEV_ALLOC(ctx, g64, 2 * l * 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, g32, 2 * l * kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, xd, 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, yd, 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, c32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o64, 2 * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(g64.p, gsw, 2 * l * 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(xd.p, x, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(yd.p, y, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  // per ABI limb: Q (composite) or the two primes (MP)
  const u64 m0 = ctx->S.mp ? ctx->S.limb[0].q : ctx->S.Q, m1 = ctx->S.mp ? ctx->S.limb[1].q : ctx->S.Q;
  // y - x, external product (NTT-form output), INTT, + x (server.cpp:696-721)
  launch_addsub_mod((u64 *)yd.p, (u64 *)xd.p, (u64 *)yd.p, 2 * abi_kn(ctx), kN, ctx->K, m0, m1, true, ctx->stream);
  launch_split_u64((u64 *)g64.p, (u32 *)g32.p, 2 * l * 2, ctx->S, ctx->stream);
  launch_split_u64((u64 *)yd.p, (u32 *)c32.p, 2, ctx->S, ctx->stream);
  launch_ext_prod((u32 *)g32.p, (int)l, (int)b, (u32 *)c32.p, (u32 *)o32.p, ctx->S, ctx->stream);
  launch_join_u64((u32 *)o32.p, (u64 *)o64.p, 2, ctx->S, ctx->stream);
  launch_ntt_u64((u64 *)o64.p, 2, true, ctx->S, ctx->stream);
  launch_addsub_mod((u64 *)xd.p, (u64 *)o64.p, (u64 *)o64.p, 2 * abi_kn(ctx), kN, ctx->K, m0, m1, false, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o64.p, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_bv_ks(opir_ctx *ctx, const uint64_t *ct, uint32_t galois_k, const uint64_t *key, uint64_t *out) {
  if (!ctx || !ct || !key || !out) return OPIR_EINVAL;
  const size_t N = ctx->N;
  if (!(galois_k & 1) || galois_k >= 2 * N) return set_err(ctx, OPIR_EINVAL, "galois element must be odd and < 2N");
  // k^-1 mod 2N: the odd residues mod 2N = 2^(log N + 1) form a group of order N
  unsigned kinv = 1;
  for (size_t e = 0; e + 1 < N; e++) kinv = (unsigned)((kinv * (u64)galois_k) & (2 * N - 1));
  if (ctx->k2) return k2_bv_ks(ctx, ct, (int)kinv, key, out);
  const size_t L = ctx->G.l_ks;
  int rc = validate_polys(ctx, ct, 2, "ciphertext");
  if (rc) return rc;
  if ((rc = validate_polys(ctx, key, L * 2, "BV key"))) return rc;
  // key in: [l_ks][a, b][K*N]; device: [l_ks][b (-> c0), a (-> c1)][2 limbs][N]
  std::vector<u32> hk(L * kCtWords);
  for (size_t r = 0; r < L; r++)
    for (int ab = 0; ab < 2; ab++)
      poly_to_rns(ctx, key + (r * 2 + ab) * abi_kn(ctx), hk.data() + r * kCtWords + (size_t)(ab == 0 ? 1 : 0) * 2 * kN);
  DevBuf kd, kp, c64, c32, o32, o64;
  DEV_ALLOC(ctx, kd, hk.size() * sizeof(u32));
  DEV_ALLOC(ctx, kp, sizeof(void *));
  DEV_ALLOC(ctx, c64, 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, c32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o64, 2 * abi_kn(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(kd.p, hk.data(), hk.size() * sizeof(u32), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(kp.p, &kd.p, sizeof(void *), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(c64.p, ct, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_split_u64((u64 *)c64.p, (u32 *)c32.p, 2, ctx->S, ctx->stream);
  launch_bv_apply((const u32 *)c32.p, (u32 *)o32.p, (const u32 *const *)kp.p, (int)kinv, ctx->G, ctx->S, ctx->stream);
  launch_join_u64((u32 *)o32.p, (u64 *)o64.p, 2, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o64.p, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_mod_switch(opir_ctx *ctx, const uint64_t *ct, uint64_t *out) {
  if (!ctx || !ct || !out) return OPIR_EINVAL;
  if (ctx->k2) return k2_mod_switch(ctx, ct, out);
  int rc = validate_polys(ctx, ct, 2, "ciphertext");
  if (rc) return rc;
  DevBuf c64, c32, n32, o64;
  DEV_ALLOC(ctx, c64, 2 * abi_kn(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, c32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, n32, kCtWords * sizeof(u32));
  DEV_ALLOC(ctx, o64, 2 * kN * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(c64.p, ct, 2 * abi_kn(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  launch_split_u64((u64 *)c64.p, (u32 *)c32.p, 2, ctx->S, ctx->stream);
  launch_ct_ntt((u32 *)c32.p, (u32 *)n32.p, 0, 0, 1, 1, ctx->S, ctx->stream);
  launch_final((u32 *)n32.p, (u64 *)o64.p, 0, 1, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o64.p, 2 * kN * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

}  // extern "C"
