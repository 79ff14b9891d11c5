// C ABI entry points for the 64-bit-limb schemes: CONFIG_N4096_K2_MP (N =
// 4096, two ~60-bit RNS limbs) and CONFIG_N2048_K1 (N = 2048, one ~60-bit
// prime). The same PirServer::make_query pipeline (src/server.cpp:895-947)
// runs on the kernels of opir_k2.hip. Ciphertexts cross the ABI in the
// reference's layout: c0 = [limb0 N][limb1 N], then c1 (rlwe.h:17-31).
#include "opir_capi_internal.h"

namespace capi {

// u64 values of one ciphertext of ctx's scheme.
static inline size_t ctv(const opir_ctx *ctx) { return k2::ct_values(ctx->S64); }

// Every polynomial of v ([npoly][K limbs][N]) reduced mod its limb prime.
int k2_validate_rns(opir_ctx *ctx, const uint64_t *v, size_t npoly, const char *what) {
  const size_t N = ctx->N, K = ctx->K;
  for (size_t p = 0; p < npoly; p++)
    for (size_t m = 0; m < K; m++) {
      const u64 q = ctx->S64.limb[m].q;
      const uint64_t *x = v + (p * K + m) * N;
      for (size_t j = 0; j < N; j++)
        if (x[j] >= q) return set_err(ctx, OPIR_EINVAL, std::string(what) + " value not reduced mod its RNS prime");
    }
  return OPIR_OK;
}

namespace {

int validate_rns(opir_ctx *ctx, const uint64_t *v, size_t npoly, const char *what) {
  return k2_validate_rns(ctx, v, npoly, what);
}

int upload_k2(opir_ctx *ctx, const uint64_t *ids, size_t B, k2::KeyPtrs64 &P) {
  const void *const *t = nullptr;
  const int rc = upload_key_ptrs(ctx, ids, B, &t);
  if (rc) return rc;
  P.bv = reinterpret_cast<const u64 *const *>(t);
  P.gsw = reinterpret_cast<const u64 *const *>(t + B);
  return OPIR_OK;
}

void rec(opir_ctx *ctx, int i, bool aux = false) {
  if (ctx->profiling) (void)hipEventRecord(ctx->ev[i], aux ? ctx->aux_stream : ctx->stream);
}

// Expansion levels over e64_a / e64_b; returns the buffer holding the leaves.
u64 *run_expand(opir_db *db, size_t B, const k2::KeyPtrs64 &P) {
  opir_ctx *ctx = db->ctx;
  const size_t stride = db->exp_cts * ctv(ctx);
  u64 *in = db->e64_a, *out = db->e64_b;
  for (size_t L = 0; L < db->h; L++) {
    k2::launch_k2_expand_level(in, out, stride, (int)nodes_at_level(db->u, db->h, L), (int)B, (int)L, P,
                               (size_t)ctx->G64.l_ks * ctv(ctx), ctx->G64, ctx->S64, ctx->stream, db->part64);
    std::swap(in, out);
  }
  return in;
}

// MUX tree (server.cpp:599-649) over NTT-form candidates in m64_a.
u64 *run_other_dim(opir_db *db, size_t B, const u64 *sel) {
  opir_ctx *ctx = db->ctx;
  const size_t stride = db->Nrest * ctv(ctx);
  if (db->dims == 1) return db->m64_a;
  const size_t hs = db->dims - 1;
  const size_t sel_vals = 2 * (size_t)ctx->G64.l_ep * ctv(ctx);
  const size_t sel_stride = hs * sel_vals;
  const size_t last = 2 * db->Nrest - (size_t{1} << hs);
  const size_t offset = db->Nrest - last;
  const size_t perfect = size_t{1} << (hs - 1);
  u64 *in = db->m64_a, *out = db->m64_b;
  k2::launch_k2_mux(in, out, stride, (int)perfect, (int)offset, 0, (int)B, sel, sel_stride, ctx->G64, ctx->S64,
                    ctx->stream, db->part64);
  std::swap(in, out);
  for (size_t a = 1; a < hs; a++) {
    const size_t half = (size_t{1} << (hs - a)) >> 1;
    k2::launch_k2_mux(in, out, stride, (int)half, -1, (int)half, (int)B, sel + a * sel_vals, sel_stride, ctx->G64,
                      ctx->S64, ctx->stream, db->part64);
    std::swap(in, out);
  }
  return in;
}

}  // namespace

// A scheme for this path: two primes at N = 4096 (q2 != 0) or one prime at
// N = 2048 (q2 = 0), each prime < 2^60 (so 4q < 2^62 for the lazy butterflies)
// and 1 mod 2N.
int k2_validate_scheme(const opir_scheme *sch) {
  const int K = sch->q2 ? 2 : 1;
  const u64 N = sch->poly_degree;
  if (!((K == 2 && N == 4096) || (K == 1 && N == 2048)))
    return set_err(nullptr, OPIR_EINVAL,
                   "the 64-bit-limb schemes are two primes at N=4096 (CONFIG_N4096_K2_MP) or one prime at N=2048 "
                   "(CONFIG_N2048_K1)");
  const u64 qs[2] = {sch->q1, sch->q2};
  const u64 roots[2] = {sch->ntt_root % qs[0], K == 2 ? sch->ntt_root2 % qs[1] : 0};
  for (int m = 0; m < K; m++)
    if (qs[m] < 3 || qs[m] >= (1ULL << 60) || (qs[m] - 1) % (2 * N))
      return set_err(nullptr, OPIR_EINVAL, "limb primes must be < 2^60 and 1 mod 2N");
  if (K == 2 && (qs[0] == qs[1] || qs[0] >= 2 * qs[1] || qs[1] >= 2 * qs[0]))
    return set_err(nullptr, OPIR_EINVAL, "K=2 limb primes must be distinct and within a factor 2");
  for (int m = 0; m < K; m++)
    if (powmod_h(roots[m], N, qs[m]) != qs[m] - 1)
      return set_err(nullptr, OPIR_EINVAL, "ntt_root / ntt_root2 is not a primitive 2N-th root of its prime");
  // K = 2: unsigned digits stay below both primes, the signed BV digits' last
  // shift b*(l_ks-1) fits a 128-bit value with its sign. K = 1: every gadget
  // is balanced over the centred int64 value, so its last offset 2^(b(l-1)-1)
  // plus |D| < 2^59 must fit an int64. Both: a 128-bit multiply-accumulate
  // sum holds at most 64 products (2 l_ep, 2 l_key, l_ks).
  const u64 wmax = K == 2 ? 126 : 62;
  if (sch->l_ep == 0 || sch->l_key == 0 || sch->l_ks == 0 || sch->base_log2 < 2 || sch->base_log2 > 32 ||
      sch->base_log2_key < 2 || sch->base_log2_key > 32 || sch->bv_base_log2 < 2 || sch->bv_base_log2 > 32 ||
      (u64)sch->bv_base_log2 * (sch->l_ks - 1) > wmax || sch->l_ep > 32 || sch->l_key > 32 || sch->l_ks > 64 ||
      (K == 1 && ((u64)sch->base_log2 * (sch->l_ep - 1) > wmax || (u64)sch->base_log2_key * (sch->l_key - 1) > wmax)))
    return set_err(nullptr, OPIR_EINVAL, "invalid gadget parameters");
  if (sch->small_q < 2 || sch->small_q >= qs[0]) return set_err(nullptr, OPIR_EINVAL, "small_q must be below q0");
  return OPIR_OK;
}

int k2_ctx_init(opir_ctx *ctx, const opir_scheme *sch) {
  const int K = sch->q2 ? 2 : 1;
  const size_t N = sch->poly_degree;
  const u64 qs[2] = {sch->q1, sch->q2};
  const u64 roots[2] = {sch->ntt_root % qs[0], K == 2 ? sch->ntt_root2 % qs[1] : 0};
  const int logn = 63 - __builtin_clzll(N);
  ctx->S64.N = (int)N;
  ctx->S64.K = K;
  std::vector<u64> tab((size_t)K * 4 * N);
  for (int m = 0; m < K; m++) {
    const u64 q = qs[m], w = roots[m];
    u64 wi = 0, ni = 0;
    if (!inv_h(w, q, wi) || !inv_h(N % q, q, ni)) return set_err(nullptr, OPIR_EINVAL, "root or N not invertible");
    u64 *psi = tab.data() + (size_t)m * 4 * N, *psi_s = psi + N, *ipsi = psi + 2 * N, *ipsi_s = psi + 3 * N;
    u64 p = 1, pi = 1;
    for (size_t i = 0; i < N; i++) {
      const size_t r = brv(i, logn);
      psi[r] = p; psi_s[r] = shoup64_h(p, q);
      ipsi[r] = pi; ipsi_s[r] = shoup64_h(pi, q);
      p = mulmod_h(p, w, q);
      pi = mulmod_h(pi, wi, q);
    }
    k2::Limb64 &L = ctx->S64.limb[m];
    L.q = q; L.q2 = 2 * q;
    L.r64 = (u64)(((u128)1 << 64) % q); L.r64s = shoup64_h(L.r64, q);
    L.one_s = (u64)(((u128)1 << 64) / q);
    L.ninv = ni; L.ninv_s = shoup64_h(ni, q);
  }
  HIP_TRY(nullptr, hipMalloc((void **)&ctx->d_tables64, tab.size() * sizeof(u64)));
  HIP_TRY(nullptr, hipMemcpy(ctx->d_tables64, tab.data(), tab.size() * sizeof(u64), hipMemcpyHostToDevice));
  for (int m = 0; m < K; m++) {
    k2::Limb64 &L = ctx->S64.limb[m];
    const u64 *base = ctx->d_tables64 + (size_t)m * 4 * N;
    L.psi = base; L.psi_s = base + N; L.ipsi = base + 2 * N; L.ipsi_s = base + 3 * N;
  }
  if (K == 2) {
    const u64 q0 = qs[0], q1 = qs[1];
    u64 q0inv = 0;
    if (!inv_h(q0 % q1, q1, q0inv)) return set_err(nullptr, OPIR_EINVAL, "q0, q1 not coprime");
    ctx->S64.q0inv = q0inv;
    ctx->S64.q0inv_s = shoup64_h(q0inv, q1);
    const u128 Q = (u128)q0 * q1;
    ctx->S64.Q_lo = (u64)Q; ctx->S64.Q_hi = (u64)(Q >> 64);
    ctx->S64.hQ_lo = (u64)(Q >> 1); ctx->S64.hQ_hi = (u64)(Q >> 65);
  }
  ctx->S64.small_q = sch->small_q;
  for (int i = 0; i < K; i++) {  // k = 63 + bit length of q: floor(2^k / q) in [2^63, 2^64)
    const u64 q = qs[i];
    const int k = 63 + (64 - __builtin_clzll(q));
    ctx->S64.rm[i] = (u64)(((u128)1 << k) / q);
    ctx->S64.rk[i] = k;
  }
  ctx->G64 = k2::Gadget64{(int)sch->l_ep, (int)sch->base_log2, (int)sch->l_key, (int)sch->base_log2_key,
                          (int)sch->l_ks, (int)sch->bv_base_log2};
  ctx->k2 = true;
  ctx->N = N;
  ctx->K = K;
  return OPIR_OK;
}

void k2_ctx_free(opir_ctx *ctx) {
  if (ctx->d_tables64) (void)hipFree(ctx->d_tables64);
  ctx->d_tables64 = nullptr;
}

int k2_db_init(opir_db *db) {
  opir_ctx *ctx = db->ctx;
  // 8 balanced int8 digit planes per 60-bit residue, MFMA fragment order
  db->geom = make_db_geom(db->N0, db->Nrest, 8, (int)ctx->N, ctx->K);
  const size_t bytes = db_bytes(db->geom);
  if (hipMalloc((void **)&db->d_db, bytes) != hipSuccess)
    return set_err(ctx, OPIR_ENOMEM, "hipMalloc database (" + std::to_string(bytes) + " bytes)");
  HIP_TRY(ctx, hipMemsetAsync(db->d_db, 0, bytes, ctx->stream));
  return OPIR_OK;
}

size_t k2_ws_bytes(const opir_db *db, size_t B) {
  const opir_ctx *ctx = db->ctx;
  const size_t ct = ctv(ctx) * sizeof(u64);
  const size_t nsel = db->dims - 1;
  return qfrag_bytes(make_qfrag_geom(db->N0, (int)B, 8, (int)ctx->N, ctx->K)) + 2 * B * db->exp_cts * ct +
         B * db->N0 * ct + B * (nsel ? nsel : 1) * 2 * ctx->G64.l_ep * ct + 2 * B * db->Nrest * ct + B * ct +
         B * 2 * ctx->N * sizeof(u64) + k2::kSplitTargetBlocks64 * ct;
}

void k2_release_ws(opir_db *db) {
  for (u64 **b : {&db->e64_a, &db->e64_b, &db->q64, &db->sel64, &db->m64_a, &db->m64_b, &db->q_u64, &db->r_u64,
                  &db->part64}) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  if (db->qfrag) (void)hipFree(db->qfrag);
  db->qfrag = nullptr;
  db->ws_B = 0;
}

int k2_ensure_ws(opir_db *db, size_t B) {
  opir_ctx *ctx = db->ctx;
  if (db->ws_B >= B) return OPIR_OK;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // queued work may still use the old buffers
  drop_graphs(db);
  k2_release_ws(db);
  const size_t ct = ctv(ctx) * sizeof(u64), nsel = db->dims - 1;
  HIP_TRY(ctx, hipMalloc((void **)&db->e64_a, B * db->exp_cts * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->e64_b, B * db->exp_cts * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->q64, B * db->N0 * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->sel64, B * (nsel ? nsel : 1) * 2 * ctx->G64.l_ep * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->m64_a, B * db->Nrest * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->m64_b, B * db->Nrest * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->q_u64, B * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->r_u64, B * 2 * ctx->N * sizeof(u64)));
  HIP_TRY(ctx, hipMalloc((void **)&db->part64, k2::kSplitTargetBlocks64 * ct));
  HIP_TRY(ctx, hipMalloc((void **)&db->qfrag, qfrag_bytes(make_qfrag_geom(db->N0, (int)B, 8, (int)ctx->N, ctx->K))));
  db->ws_B = B;
  return OPIR_OK;
}

void k2_db_free(opir_db *db) {
  if (!db->ctx->k2) return;
  k2_release_ws(db);
  if (db->d_db && db->owns_db) (void)hipFree(db->d_db);
  db->d_db = nullptr;
}

// NTT `n` plaintexts (u64 coefficients < t, device) per limb into tmp and
// scatter them into the digit-plane database at first_pt.
int k2_ingest(opir_db *db, const u64 *d_pts, u64 *d_tmp, size_t first_pt, size_t n) {
  opir_ctx *ctx = db->ctx;
  k2::launch_k2_db_load(d_pts, d_tmp, n, 0, n, ctx->S64, ctx->stream);
  launch_db_scatter64(d_tmp, db->d_db, first_pt, n, db->geom, ctx->stream);
  LAUNCH_CHECK(ctx);
  return OPIR_OK;
}

int k2_load_plain(opir_db *db, const uint64_t *coeffs, size_t first_pt, size_t count) {
  opir_ctx *ctx = db->ctx;
  if (first_pt > db->num_pt || count > db->num_pt - first_pt)
    return set_err(ctx, OPIR_EINVAL, "PirServer::load_data logical plaintext count exceeds its layout");
  const u64 t = ctx->sch.plain_mod;
  const size_t N = ctx->N;
  for (size_t i = 0; i < count * N; i++)
    if (coeffs[i] >= t) return set_err(ctx, OPIR_EINVAL, "plaintext coefficient not reduced mod t");
  const size_t chunk = 4096, cn = std::min(chunk, std::max<size_t>(count, 1));
  DevBuf d, tmp;
  DEV_ALLOC(ctx, d, cn * N * sizeof(u64));
  DEV_ALLOC(ctx, tmp, (size_t)ctx->K * cn * N * sizeof(u64));
  for (size_t off = 0; off < count; off += chunk) {
    const size_t n = std::min(chunk, count - off);
    HIP_TRY(ctx, hipMemcpyAsync(d.p, coeffs + off * N, n * N * sizeof(u64), hipMemcpyHostToDevice, ctx->stream));
    if (int rc = k2_ingest(db, (const u64 *)d.p, (u64 *)tmp.p, first_pt + off, n)) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return OPIR_OK;
}

int k2_set_keys(opir_ctx *ctx, const uint64_t *bv, size_t h, const uint64_t *gsw, ClientKeys &k) {
  const size_t L = ctx->G64.l_ks, lk = ctx->G64.l_key, N = ctx->N, K = ctx->K, ct = ctv(ctx);
  // bv in: [h][l_ks][a, b][limb][N]; device: [h][l_ks][b (-> c0), a (-> c1)][limb][N]
  std::vector<u64> hbv(h * L * ct);
  for (size_t r = 0; r < h * L; r++)
    for (size_t ab = 0; ab < 2; ab++)
      for (size_t m = 0; m < K; m++) {
        const u64 q = ctx->S64.limb[m].q;
        const uint64_t *src = bv + ((r * 2 + ab) * K + m) * N;
        u64 *dst = hbv.data() + r * ct + ((ab == 0 ? 1 : 0) * K + m) * N;
        for (size_t j = 0; j < N; j++) {
          if (src[j] >= q) return set_err(ctx, OPIR_EINVAL, "BV key value not reduced mod its RNS prime");
          dst[j] = src[j];
        }
      }
  int rc = validate_rns(ctx, gsw, 2 * lk * 2, "RGSW key");
  if (rc) return rc;
  if (h) HIP_TRY(ctx, hipMalloc(&k.bv, hbv.size() * sizeof(u64)));
  if (hipMalloc(&k.gsw, 2 * lk * ctv(ctx) * sizeof(u64)) != hipSuccess) {
    (void)hipFree(k.bv);
    return set_err(ctx, OPIR_ENOMEM, "hipMalloc keys");
  }
  if ((h && hipMemcpy(k.bv, hbv.data(), hbv.size() * sizeof(u64), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(k.gsw, gsw, 2 * lk * ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(k.bv);
    (void)hipFree(k.gsw);
    return set_err(ctx, OPIR_EHIP, "upload keys");
  }
  k.h = h;
  return OPIR_OK;
}

int k2_tune_scan(opir_db *db, size_t B) {
  opir_ctx *ctx = db->ctx;
  const int sv = pick_scan_variant(db, B, 8, [&](int v) {
    launch_scan64(db->d_db, db->qfrag, db->m64_a, db->Nrest * ctv(ctx), (int)B, db->geom, ctx->S64, ctx->stream, v);
  });
  return sv < 0 ? set_err(ctx, OPIR_EHIP, "scan autotune failed") : OPIR_OK;
}

int k2_answer_device(opir_db *db, size_t B, const uint64_t *ids, const u64 *d_q, u64 *d_r) {
  opir_ctx *ctx = db->ctx;
  int rc = check_answerable(db);
  if (rc) return rc;
  rc = k2_ensure_ws(db, B);
  if (rc) return rc;
  k2::KeyPtrs64 P;
  rc = upload_k2(ctx, ids, B, P);
  if (rc) return rc;
  for (size_t i = 0; i < B; i++)
    if (ctx->clients()[ids[i]].h < db->h)
      return set_err(ctx, OPIR_EINVAL, "PirServer session lacks keys for its expansion height");
  const size_t exp_stride = db->exp_cts * ctv(ctx);
  return run_answer(db, B, 8, P.bv, d_q, d_r, [&]() -> int {
    rec(ctx, 0);
    // root ct of query i -> e64_a[i][0]
    HIP_TRY(ctx, hipMemcpy2DAsync(db->e64_a, exp_stride * sizeof(u64), d_q, ctv(ctx) * sizeof(u64), ctv(ctx) * sizeof(u64), B,
                                  hipMemcpyDeviceToDevice, ctx->stream));
    u64 *leaves = run_expand(db, B, P);
    LAUNCH_CHECK(ctx);
    rec(ctx, 1);
    // selector completion: beside the scan on the auxiliary stream, or ahead
    // of it (complete_beside_scan); at N = 4096 with two limbs its
    // 1024-thread workgroups hold 130 KiB of LDS, one per CU
    const size_t nsel = db->dims - 1;
    const bool beside = complete_beside_scan(nsel * ctx->G64.l_ep, B, ctx->K == 2, 0);
    if (beside) fork_aux(ctx);
    const size_t sel_stride = nsel * 2 * (size_t)ctx->G64.l_ep * ctv(ctx);
    k2::launch_k2_complete(leaves + db->N0 * ctv(ctx), db->sel64, exp_stride, sel_stride, (int)(nsel * ctx->G64.l_ep),
                           (int)B, P, ctx->G64, ctx->S64, beside ? ctx->aux_stream : ctx->stream);
    LAUNCH_CHECK(ctx);
    rec(ctx, 2, beside);
    k2::launch_k2_ct_ntt(leaves, db->q64, exp_stride, db->N0 * ctv(ctx), (int)db->N0, (int)B, ctx->S64, ctx->stream);
    launch_qfrag64(db->q64, db->qfrag, db->N0 * ctv(ctx), (int)B, db->geom, ctx->stream);
    rec(ctx, 5);
    const int sv = pick_scan_variant(db, B, 8, [&](int v) {
      launch_scan64(db->d_db, db->qfrag, db->m64_a, db->Nrest * ctv(ctx), (int)B, db->geom, ctx->S64, ctx->stream, v);
  });
  if (launch_scan64(db->d_db, db->qfrag, db->m64_a, db->Nrest * ctv(ctx), (int)B, db->geom, ctx->S64, ctx->stream, sv))
    return set_err(ctx, OPIR_EINVAL, "scan variant cannot take this batch size");
  LAUNCH_CHECK(ctx);
  rec(ctx, 6);
  rec(ctx, 3);
  if (beside) join_aux(ctx);
  u64 *res = run_other_dim(db, B, db->sel64);
  LAUNCH_CHECK(ctx);
  rec(ctx, 4);
  k2::launch_k2_final(res, d_r, db->Nrest * ctv(ctx), (int)B, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  rec(ctx, 7);
  return OPIR_OK;
  });
}

// ---------------------------------------------------------------------------
// Stage entry points (host buffers)
// ---------------------------------------------------------------------------
int k2_ntt(opir_ctx *ctx, uint64_t *data, size_t count, bool inverse) {
  int rc = validate_rns(ctx, data, count, "NTT input");
  if (rc) return rc;
  DevBuf d;
  const size_t bytes = count * ctx->K * ctx->N * sizeof(u64);
  DEV_ALLOC(ctx, d, bytes);
  HIP_TRY(ctx, hipMemcpy(d.p, data, bytes, hipMemcpyHostToDevice));
  k2::launch_k2_ntt((u64 *)d.p, count, inverse, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(data, d.p, bytes, hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_expand(opir_db *db, uint64_t id, const uint64_t *query, uint64_t *out) {
  opir_ctx *ctx = db->ctx;
  int rc = validate_rns(ctx, query, 2, "query");
  if (rc) return rc;
  if (!ctx->clients().count(id)) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  if (ctx->clients()[id].h < db->h) return set_err(ctx, OPIR_EINVAL, "PirServer session lacks keys for its expansion height");
  rc = k2_ensure_ws(db, 1);
  if (rc) return rc;
  k2::KeyPtrs64 P;
  rc = upload_k2(ctx, &id, 1, P);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpy(db->e64_a, query, ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  u64 *leaves = run_expand(db, 1, P);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, leaves, db->u * ctv(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_complete(opir_db *db, uint64_t id, const uint64_t *expanded, uint64_t *sel) {
  opir_ctx *ctx = db->ctx;
  const size_t nsel = db->dims - 1;
  if (!nsel) return OPIR_OK;
  if (!ctx->clients().count(id)) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  int rc = validate_rns(ctx, expanded, db->u * 2, "expanded");
  if (rc) return rc;
  rc = k2_ensure_ws(db, 1);
  if (rc) return rc;
  k2::KeyPtrs64 P;
  rc = upload_k2(ctx, &id, 1, P);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpy(db->e64_a, expanded, db->u * ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  const size_t sel_vals = nsel * 2 * (size_t)ctx->G64.l_ep * ctv(ctx);
  k2::launch_k2_complete(db->e64_a + db->N0 * ctv(ctx), db->sel64, 0, sel_vals, (int)(nsel * ctx->G64.l_ep), 1, P,
                         ctx->G64, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(sel, db->sel64, sel_vals * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_first_dim(opir_db *db, const uint64_t *query, uint64_t *out) {
  opir_ctx *ctx = db->ctx;
  int rc = validate_rns(ctx, query, db->N0 * 2, "query");
  if (rc) return rc;
  rc = k2_ensure_ws(db, 1);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpy(db->e64_a, query, db->N0 * ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  k2::launch_k2_ct_ntt(db->e64_a, db->q64, 0, 0, (int)db->N0, 1, ctx->S64, ctx->stream);
  launch_qfrag64(db->q64, db->qfrag, 0, 1, db->geom, ctx->stream);
  if (launch_scan64(db->d_db, db->qfrag, db->m64_a, 0, 1, db->geom, ctx->S64, ctx->stream,
                    ctx->scan_variant == -2 ? -1 : ctx->scan_variant))
    return set_err(ctx, OPIR_EINVAL, "scan variant cannot take this batch size");
  k2::launch_k2_ntt(db->m64_a, db->Nrest * 2, true, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, db->m64_a, db->Nrest * ctv(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_other_dim(opir_db *db, const uint64_t *mid, const uint64_t *sel, uint64_t *out) {
  opir_ctx *ctx = db->ctx;
  if (int e = check_answerable(db)) return e;
  const size_t nsel = db->dims - 1;
  const size_t sel_vals = nsel * 2 * (size_t)ctx->G64.l_ep * ctv(ctx);
  int rc = validate_rns(ctx, mid, db->Nrest * 2, "candidates");
  if (rc) return rc;
  if (nsel) rc = validate_rns(ctx, sel, sel_vals / (ctx->N * ctx->K), "selectors");
  if (rc) return rc;
  rc = k2_ensure_ws(db, 1);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpy(db->m64_a, mid, db->Nrest * ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  k2::launch_k2_ntt(db->m64_a, db->Nrest * 2, false, ctx->S64, ctx->stream);
  if (nsel) HIP_TRY(ctx, hipMemcpy(db->sel64, sel, sel_vals * sizeof(u64), hipMemcpyHostToDevice));
  u64 *res = run_other_dim(db, 1, db->sel64);
  k2::launch_k2_ntt(res, 2, true, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, res, ctv(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_ext_prod(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t b, const uint64_t *ct, uint64_t *out) {
  if (l == 0 || l > 64 || b == 0 || b > 32) return set_err(ctx, OPIR_EINVAL, "invalid gadget");
  int rc = validate_rns(ctx, gsw, 2 * l * 2, "RGSW");
  if (rc) return rc;
  rc = validate_rns(ctx, ct, 2, "ciphertext");
  if (rc) return rc;
  DevBuf g, c, o;
  DEV_ALLOC(ctx, g, 2 * l * ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, c, ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, o, ctv(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(g.p, gsw, 2 * l * ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(c.p, ct, ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  k2::launch_k2_ext_prod((const u64 *)g.p, (int)l, (int)b, (const u64 *)c.p, (u64 *)o.p, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, ctv(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_ext_prod_mux(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t b, const uint64_t *x, const uint64_t *y,
                    uint64_t *out) {
  if (l == 0 || l > 64 || b == 0 || b > 32) return set_err(ctx, OPIR_EINVAL, "invalid gadget");
  int rc = validate_rns(ctx, gsw, 2 * l * 2, "RGSW");
  if (rc) return rc;
  if ((rc = validate_rns(ctx, x, 2, "ciphertext x"))) return rc;
  if ((rc = validate_rns(ctx, y, 2, "ciphertext y"))) return rc;
  DevBuf g, xd, yd, o;
  DEV_ALLOC(ctx, g, 2 * l * ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, xd, ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, yd, ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, o, ctv(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(g.p, gsw, 2 * l * ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(xd.p, x, ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(yd.p, y, ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  const u64 q0 = ctx->S64.limb[0].q, q1 = ctx->K == 2 ? ctx->S64.limb[1].q : q0;
  opir::launch_addsub_mod((u64 *)yd.p, (u64 *)xd.p, (u64 *)yd.p, ctv(ctx), ctx->N, ctx->K, q0, q1, true, ctx->stream);
  k2::launch_k2_ext_prod((const u64 *)g.p, (int)l, (int)b, (const u64 *)yd.p, (u64 *)o.p, ctx->S64, ctx->stream);
  k2::launch_k2_ntt((u64 *)o.p, 2, true, ctx->S64, ctx->stream);
  opir::launch_addsub_mod((u64 *)xd.p, (u64 *)o.p, (u64 *)o.p, ctv(ctx), ctx->N, ctx->K, q0, q1, false, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, ctv(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_bv_ks(opir_ctx *ctx, const uint64_t *ct, int kinv, const uint64_t *key, uint64_t *out) {
  const size_t L = ctx->G64.l_ks, N = ctx->N, K = ctx->K;
  int rc = validate_rns(ctx, ct, 2, "ciphertext");
  if (rc) return rc;
  // key in: [l_ks][a, b][limb][N]; device: [l_ks][b (-> c0), a (-> c1)][limb][N]
  std::vector<u64> hk(L * ctv(ctx));
  for (size_t r = 0; r < L; r++)
    for (size_t ab = 0; ab < 2; ab++)
      for (size_t m = 0; m < K; m++) {
        const u64 q = ctx->S64.limb[m].q;
        const uint64_t *src = key + ((r * 2 + ab) * K + m) * N;
        u64 *dst = hk.data() + r * ctv(ctx) + ((ab == 0 ? 1 : 0) * K + m) * N;
        for (size_t j = 0; j < N; j++) {
          if (src[j] >= q) return set_err(ctx, OPIR_EINVAL, "BV key value not reduced mod its RNS prime");
          dst[j] = src[j];
        }
      }
  DevBuf kd, kp, c, o;
  DEV_ALLOC(ctx, kd, hk.size() * sizeof(u64));
  DEV_ALLOC(ctx, kp, sizeof(void *));
  DEV_ALLOC(ctx, c, ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, o, ctv(ctx) * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(kd.p, hk.data(), hk.size() * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(kp.p, &kd.p, sizeof(void *), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(c.p, ct, ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  k2::launch_k2_bv_apply((const u64 *)c.p, (u64 *)o.p, (const u64 *const *)kp.p, kinv, ctx->G64, ctx->S64,
                         ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, ctv(ctx) * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int k2_mod_switch(opir_ctx *ctx, const uint64_t *ct, uint64_t *out) {
  int rc = validate_rns(ctx, ct, 2, "ciphertext");
  if (rc) return rc;
  DevBuf c, o;
  DEV_ALLOC(ctx, c, ctv(ctx) * sizeof(u64));
  DEV_ALLOC(ctx, o, 2 * (size_t)ctx->N * sizeof(u64));
  HIP_TRY(ctx, hipMemcpy(c.p, ct, ctv(ctx) * sizeof(u64), hipMemcpyHostToDevice));
  k2::launch_k2_mod_switch((const u64 *)c.p, (u64 *)o.p, ctx->S64, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, o.p, 2 * (size_t)ctx->N * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

}  // namespace capi
