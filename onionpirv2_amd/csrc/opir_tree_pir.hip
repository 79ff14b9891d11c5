// Tree PIR on the answer path's kernels: the server side of the reference's
// full-binary-tree PIR layer (src/tree_query.cpp:198-240 unpack_tree_query,
// src/tree_select.cpp:121-238 build_alpha_pyramid, evaluate_alpha_dimension,
// fold_beta_dimension, select_level), for the 32-bit-limb schemes.
//
// One query selects one node per tree level. Its expansion and selector
// completion are the answer path's (a query-shape database of N0 first-dim
// slots and b + r selectors). Every level's plaintexts D_l[p] are resident
// in NTT form; the alpha pyramid (pairwise sums of the N0 one-hot cts) is
// formed in the NTT domain; every plaintext x ciphertext product of every
// level runs in one launch (k_ptct_mac: each output sums its terms
// pointwise); AlphaBeta levels fold their 2^d candidates MSB-first with the
// MUX kernels (NTT-form candidates, as in the answer path); one INTT per
// level gives the coefficient-form result. Products and sums in the NTT
// domain equal the reference's coefficient-domain sums of NTT-INTT products
// exactly (the transform is linear), so every output is bit-exact.
#include "opir_capi_internal.h"

using namespace capi;

struct opir_tree {
  opir_ctx *ctx = nullptr;
  opir_db *qdb = nullptr;       // query-shape database: expansion and completion workspace
  size_t L = 0, a = 0, N0 = 0, r = 0, b = 0, P = 0, B = 0, w = 0, h_q = 0, ell = 0;
  std::vector<size_t> R, pt_off;  // per level: plaintext count and first plaintext
  u32 *pts = nullptr;             // [sum R][2 limbs][N], NTT form
  std::vector<bool> loaded;
  u32 *pyr = nullptr, *cand = nullptr, *fold = nullptr;  // NTT-form RNS cts
  u64 *out = nullptr;             // [(L + 1)][2][K*N] coefficient form
  size_t ncand = 0, nfold = 0;
};

namespace {

struct Plan {
  size_t R = 0;
  int select_case = 0;  // 0 Single, 1 CoarsenedAlpha, 2 AlphaBeta (tree_index.cpp:172-215)
  size_t coarsen = 0, beta_begin = 0, beta_count = 0;
};

Plan plan_of(const opir_tree *t, size_t level) {
  Plan p;
  p.R = level >= t->r ? size_t{1} << (level - t->r) : 1;
  if (p.R == 1) {
    p.select_case = 0;
    p.coarsen = t->a;
  } else if (p.R < t->N0) {
    p.select_case = 1;
    p.coarsen = t->a - (level - t->r);
  } else {
    p.select_case = 2;
    p.beta_begin = t->L - level;
    p.beta_count = level - t->r - t->a;
  }
  if (p.select_case != 2) p.beta_begin = t->b;
  return p;
}

// Pyramid level c starts at ct offset sum_{c' < c} N0 / 2^c'.
size_t pyr_off(const opir_tree *t, size_t c) {
  size_t o = 0;
  for (size_t k = 0; k < c; k++) o += t->N0 >> k;
  return o;
}

// out[o] = sum over terms [off[o], off[o+1]) of pt[term.x] (.) ct[term.y],
// NTT form, both polys, both limbs. grid (outputs, 4 = poly * 2 + limb).
__global__ __launch_bounds__(256) void k_ptct_mac(const u32 *pts, const u32 *cts, const int *off, const int2 *terms,
                                                  u32 *out, SchemeConst S) {
  const int o = blockIdx.x, pl = blockIdx.y, limb = pl & 1;
  const LimbConst &L = S.limb[limb];
  u32 acc[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc[c] = 0;
  for (int k = off[o]; k < off[o + 1]; k++) {
    const int2 tm = terms[k];
    const u32 *pp = pts + ((size_t)tm.x * 2 + limb) * kN;
    const u32 *cp = cts + (size_t)tm.y * kCtWords + (size_t)pl * kN;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int j = threadIdx.x + 256 * c;
      acc[c] = add_mod(acc[c], reduce_u64((u64)pp[j] * cp[j], L), L.q);
    }
  }
  u32 *dst = out + (size_t)o * kCtWords + (size_t)pl * kN;
#pragma unroll
  for (int c = 0; c < 8; c++) dst[threadIdx.x + 256 * c] = acc[c];
}

// Pyramid step: out[j] = in[2j] + in[2j+1] (RNS cts, any form), j < n.
__global__ void k_pair_add(const u32 *in, u32 *out, size_t n, SchemeConst S) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kCtWords) return;
  const size_t j = i / kCtWords, e = i % kCtWords;
  const u32 q = S.limb[(e / kN) & 1].q;
  out[i] = add_mod(in[2 * j * kCtWords + e], in[(2 * j + 1) * kCtWords + e], q);
}

}  // namespace

extern "C" {

int opir_tree_create(opir_ctx *ctx, size_t L, size_t a, opir_tree **out) {
  if (!ctx || !out) return OPIR_EINVAL;
  *out = nullptr;
  if (ctx->k2) return set_err(ctx, OPIR_EINVAL, "tree PIR runs the 32-bit-limb schemes (N = 2048)");
  const size_t n = kN, r = 11, ell = (size_t)ctx->G.l_ep;
  // tree_index.cpp:53-118 with ell_beta = ell_gamma = L_EP (tree_query.cpp:44-61)
  if (L == 0 || L > 40 || a > 20 || L < r + a) return set_err(ctx, OPIR_EINVAL, "tree height is below r + a");
  const size_t b = L - r - a, N0 = size_t{1} << a;
  const size_t w = N0 + ell * (b + r);
  if (w > n) return set_err(ctx, OPIR_EINVAL, "packed width exceeds the ring degree");
  size_t W = 1, h = 0;
  while (W < w) W <<= 1, h++;
  if (ctx->sch.plain_mod % 2 == 0) return set_err(ctx, OPIR_EINVAL, "W must be invertible modulo t");
  opir_tree *t = new (std::nothrow) opir_tree();
  if (!t) return set_err(ctx, OPIR_ENOMEM, "out of host memory");
  t->ctx = ctx;
  t->L = L; t->a = a; t->N0 = N0; t->r = r; t->b = b; t->P = (size_t{1} << L) / n; t->B = size_t{1} << b;
  t->w = w; t->h_q = h; t->ell = ell;
  size_t total = 0, maxc = 1;
  for (size_t level = 0; level <= L; level++) {
    const Plan p = plan_of(t, level);
    t->R.push_back(p.R);
    t->pt_off.push_back(total);
    total += p.R;
    t->ncand += p.select_case == 2 ? size_t{1} << p.beta_count : 1;
    if (p.select_case == 2) maxc = std::max(maxc, size_t{1} << p.beta_count);
  }
  t->loaded.assign(L + 1, false);
  t->nfold = maxc;
  // query shape: N0 first-dimension slots, b + r selectors, one candidate row
  int rc = opir_db_create(ctx, N0, 1, 1 + b + r, h, &t->qdb);
  const size_t KN = (size_t)ctx->K * kN;
  if (!rc && (hipMalloc((void **)&t->pts, total * 2 * kN * sizeof(u32)) != hipSuccess ||
              hipMalloc((void **)&t->pyr, (2 * N0) * kCtWords * sizeof(u32)) != hipSuccess ||
              hipMalloc((void **)&t->cand, t->ncand * kCtWords * sizeof(u32)) != hipSuccess ||
              hipMalloc((void **)&t->fold, maxc * kCtWords * sizeof(u32)) != hipSuccess ||
              hipMalloc((void **)&t->out, (L + 1) * 2 * KN * sizeof(u64)) != hipSuccess))
    rc = set_err(ctx, OPIR_ENOMEM, "hipMalloc tree buffers");
  if (rc) {
    opir_tree_destroy(t);
    return rc;
  }
  *out = t;
  return OPIR_OK;
}

void opir_tree_destroy(opir_tree *t) {
  if (!t) return;
  if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
  for (void *p : {(void *)t->pts, (void *)t->pyr, (void *)t->cand, (void *)t->fold, (void *)t->out})
    if (p) (void)hipFree(p);
  if (t->qdb) opir_db_destroy(t->qdb);
  delete t;
}

int opir_tree_shape(const opir_tree *t, size_t *w, size_t *h_q, size_t *N0, size_t *nsel) {
  if (!t || !w || !h_q || !N0 || !nsel) return OPIR_EINVAL;
  *w = t->w; *h_q = t->h_q; *N0 = t->N0; *nsel = t->b + t->r;
  return OPIR_OK;
}

int opir_tree_load_level(opir_tree *t, size_t level, const uint64_t *coeffs) {
  if (!t || !coeffs) return OPIR_EINVAL;
  opir_ctx *ctx = t->ctx;
  if (level > t->L) return set_err(ctx, OPIR_EINVAL, "level is out of range");
  const size_t R = t->R[level];
  const u64 tmod = ctx->sch.plain_mod;
  std::vector<u32> h(R * 2 * kN);
  for (size_t p = 0; p < R; p++)
    for (size_t j = 0; j < (size_t)kN; j++) {
      const u64 v = coeffs[p * kN + j];
      if (v >= tmod) return set_err(ctx, OPIR_EINVAL, "plaintext coefficient not reduced mod t");
      h[(p * 2) * kN + j] = h[(p * 2 + 1) * kN + j] = (u32)v;  // < t < q: the same residue in both limbs
    }
  DevBuf d;
  DEV_ALLOC(ctx, d, h.size() * sizeof(u32));
  HIP_TRY(ctx, hipMemcpy(d.p, h.data(), h.size() * sizeof(u32), hipMemcpyHostToDevice));
  launch_ntt_rns((const u32 *)d.p, t->pts + t->pt_off[level] * 2 * kN, R, false, ctx->S, ctx->stream);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  t->loaded[level] = true;
  return OPIR_OK;
}

int opir_tree_select(opir_tree *t, uint64_t id, const uint64_t *query, uint64_t *out) {
  if (!t || !query || !out) return OPIR_EINVAL;
  opir_ctx *ctx = t->ctx;
  opir_db *db = t->qdb;
  for (size_t level = 0; level <= t->L; level++)
    if (!t->loaded[level]) return set_err(ctx, OPIR_EINVAL, "tree level " + std::to_string(level) + " not loaded");
  const size_t KN = (size_t)ctx->K * kN;
  {  // values reduced (per limb for the MP scheme)
    const u64 m0 = ctx->S.mp ? ctx->S.limb[0].q : ctx->S.Q, m1 = ctx->S.mp ? ctx->S.limb[1].q : ctx->S.Q;
    for (size_t i = 0; i < 2 * KN; i++)
      if (query[i] >= ((i / kN) % ctx->K ? m1 : m0)) return set_err(ctx, OPIR_EINVAL, "query value not reduced");
  }
  auto it = ctx->clients().find(id);
  if (it == ctx->clients().end()) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  if (it->second.h < t->h_q) return set_err(ctx, OPIR_EINVAL, "PirServer session lacks keys for its expansion height");
  int rc = ensure_ws(db, 1);
  if (rc) return rc;
  BatchPtrs P;
  if ((rc = upload_k1(ctx, &id, 1, P))) return rc;
  // terms of every output: (plaintext, pyramid ct)
  std::vector<int> off{0};
  std::vector<int2> terms;
  std::vector<size_t> first_out;
  for (size_t level = 0; level <= t->L; level++) {
    const Plan p = plan_of(t, level);
    const int base = (int)t->pt_off[level];
    first_out.push_back(off.size() - 1);
    if (p.select_case == 0) {
      terms.push_back(make_int2(base, (int)pyr_off(t, t->a)));
      off.push_back((int)terms.size());
    } else if (p.select_case == 1) {
      for (size_t k = 0; k < p.R; k++) terms.push_back(make_int2(base + (int)k, (int)(pyr_off(t, p.coarsen) + k)));
      off.push_back((int)terms.size());
    } else {
      const size_t stride = size_t{1} << p.beta_count;
      for (size_t delta = 0; delta < stride; delta++) {
        for (size_t k = 0; k < t->N0; k++) terms.push_back(make_int2(base + (int)(k * stride + delta), (int)k));
        off.push_back((int)terms.size());
      }
    }
  }
  DevBuf doff, dterms;
  DEV_ALLOC(ctx, doff, off.size() * sizeof(int));
  DEV_ALLOC(ctx, dterms, terms.size() * sizeof(int2));
  HIP_TRY(ctx, hipMemcpyAsync(doff.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(dterms.p, terms.data(), terms.size() * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(db->q_u64, query, 2 * KN * sizeof(u64), hipMemcpyHostToDevice, ctx->stream));
  // unpack_tree_query: expansion, then selector completion (b beta groups, then r gamma groups)
  launch_split_u64(db->q_u64, db->exp_a, 2, ctx->S, ctx->stream);
  u32 *leaves = run_expand(db, 1, P);
  const size_t nsel = t->b + t->r;
  const size_t sel_words = 2 * t->ell * kCtWords;
  launch_complete(leaves + t->N0 * kCtWords, db->sel, 0, nsel * sel_words, (int)(nsel * t->ell), 1, P, ctx->G, ctx->S,
                  ctx->stream);
  // alpha pyramid in the NTT domain (build_alpha_pyramid, tree_select.cpp:121-146)
  launch_ct_ntt(leaves, t->pyr, 0, 0, (int)t->N0, 1, ctx->S, ctx->stream);
  for (size_t c = 0; c < t->a; c++) {
    const size_t n = t->N0 >> (c + 1);
    k_pair_add<<<(unsigned)((n * kCtWords + 255) / 256), 256, 0, ctx->stream>>>(t->pyr + pyr_off(t, c) * kCtWords,
                                                                                t->pyr + pyr_off(t, c + 1) * kCtWords,
                                                                                n, ctx->S);
  }
  // every plaintext x ciphertext product of every level
  const size_t nout = off.size() - 1;
  k_ptct_mac<<<dim3((unsigned)nout, 4), 256, 0, ctx->stream>>>(t->pts, t->pyr, (const int *)doff.p,
                                                               (const int2 *)dterms.p, t->cand, ctx->S);
  LAUNCH_CHECK(ctx);
  // fold the AlphaBeta levels' candidates MSB-first (fold_beta_dimension,
  // tree_select.cpp:175-200), then one INTT per level
  for (size_t level = 0; level <= t->L; level++) {
    const Plan p = plan_of(t, level);
    u32 *res = t->cand + first_out[level] * kCtWords;
    if (p.select_case == 2 && p.beta_count) {
      u32 *in = res, *o = t->fold;
      size_t count = size_t{1} << p.beta_count;
      for (size_t step = 0; step < p.beta_count; step++) {
        const size_t u = p.beta_begin + p.beta_count - 1 - step, half = count / 2;
        launch_mux(in, o, 0, (int)half, -1, (int)half, 1, db->sel + u * sel_words, 0, ctx->G, ctx->S, ctx->stream);
        in = o;
        o = (o == t->fold) ? res : t->fold;
        count = half;
      }
      res = in;
    }
    launch_ct_to_coef_u64(res, t->out + level * 2 * KN, 1, ctx->S, ctx->stream);
  }
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipMemcpy(out, t->out, (t->L + 1) * 2 * KN * sizeof(u64), hipMemcpyDeviceToHost));
  return OPIR_OK;
}

}  // extern "C"
