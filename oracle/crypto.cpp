// OnionPIRv2 CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// RLWE sampling/encryption, RGSW external product, BV keyswitch, first-dim mat-mat.
#include <cassert>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "oracle.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Samplers. utils.cpp:407-443, rlwe.cpp:174-215, bv_keyswitch.cpp:166-185
// ---------------------------------------------------------------------------
void sample_uniform_poly(u64 *out, size_t N, u64 q, std::mt19937_64 &rng) {
  const u64 limit = (~u64{0}) - ((~u64{0}) % q);
  for (size_t i = 0; i < N; i++) {
    u64 r;
    do { r = rng(); } while (r >= limit);
    out[i] = r % q;
  }
}

void sample_gaussian_shared(u64 *out, size_t N, const std::vector<u64> &qs, double sigma,
                            std::mt19937_64 &rng) {
  std::normal_distribution<double> dist(0.0, sigma);
  std::vector<int64_t> e(N);
  for (size_t i = 0; i < N; i++) e[i] = std::llround(dist(rng));
  for (size_t k = 0; k < qs.size(); k++) {
    const u64 qk = qs[k];
    for (size_t i = 0; i < N; i++) {
      const int64_t v = e[i];
      if (v >= 0) out[k * N + i] = (u64)v % qk;
      else {
        const u64 a = (u64)(-v) % qk;
        out[k * N + i] = a == 0 ? 0 : qk - a;
      }
    }
  }
}

RlweSk gen_secret_key_rns(size_t N, const std::vector<u64> &qs, std::mt19937_64 &rng) {
  RlweSk sk;
  sk.data.resize(N * qs.size());
  std::uniform_int_distribution<int> dist(0, 2);
  std::vector<int8_t> tern(N);
  for (size_t i = 0; i < N; i++) {
    const int v = dist(rng);
    tern[i] = (int8_t)(v == 2 ? -1 : v);
  }
  for (size_t k = 0; k < qs.size(); k++)
    for (size_t i = 0; i < N; i++)
      sk.data[k * N + i] = tern[i] == 0 ? 0 : (tern[i] > 0 ? 1 : qs[k] - 1);
  for (size_t k = 0; k < qs.size(); k++) ntt_fwd(sk.data.data() + k * N, N, qs[k]);
  return sk;
}

void encrypt_zero_rns(const RlweSk &sk, size_t N, const std::vector<u64> &qs, double sigma,
                      std::mt19937_64 &rng, RlweCt &ct, bool ntt_form) {
  const size_t K = qs.size();
  ct.c0.assign(N * K, 0);
  ct.c1.assign(N * K, 0);
  for (size_t k = 0; k < K; k++) sample_uniform_poly(ct.c1.data() + k * N, N, qs[k], rng);
  std::vector<u64> e(N * K);
  sample_gaussian_shared(e.data(), N, qs, sigma, rng);
  std::vector<u64> a(N);
  for (size_t k = 0; k < K; k++) {
    const u64 q = qs[k];
    u64 *c0 = ct.c0.data() + k * N, *c1 = ct.c1.data() + k * N;
    std::memcpy(a.data(), c1, N * sizeof(u64));
    ntt_fwd(a.data(), N, q);
    for (size_t i = 0; i < N; i++) c0[i] = mulmod(a[i], sk.data[k * N + i], q);
    ntt_inv(c0, N, q);
    for (size_t i = 0; i < N; i++) {
      const u64 s = (c0[i] + e[k * N + i]) % q;
      c0[i] = (q - s) % q;
    }
    if (ntt_form) { ntt_fwd(c0, N, q); ntt_fwd(c1, N, q); }
  }
  ct.ntt_form = ntt_form;
}

void encrypt_bfv_rns(const std::vector<u64> &m, const RlweSk &sk, size_t N, const std::vector<u64> &qs,
                     u64 t, double sigma, std::mt19937_64 &rng, RlweCt &ct) {
  encrypt_zero_rns(sk, N, qs, sigma, rng, ct, false);
  u128 Q = 1;
  for (u64 q : qs) Q *= q;
  const u128 Delta = Q / t;
  const u64 r = (u64)(Q - Delta * t);
  for (size_t i = 0; i < N && i < m.size(); i++) {
    const u64 mi = m[i] % t;
    if (mi == 0) continue;
    const u64 rr = (u64)(((u128)r * mi + (t >> 1)) / t);
    const u128 scaled = Delta * mi + rr;
    for (size_t k = 0; k < qs.size(); k++) {
      u64 *c0 = ct.c0.data() + k * N;
      c0[i] = (c0[i] + (u64)(scaled % qs[k])) % qs[k];
    }
  }
}

void decrypt_rns(const RlweCt &ct, const RlweSk &sk, size_t N, const std::vector<u64> &qs, u64 t,
                 u64 q0_inv_mod_q1, std::vector<u64> &pt) {
  const size_t K = qs.size();
  std::vector<u64> phase(N * K), c1(N), c0(N);
  for (size_t k = 0; k < K; k++) {
    const u64 q = qs[k];
    std::memcpy(c1.data(), ct.c1.data() + k * N, N * sizeof(u64));
    if (!ct.ntt_form) ntt_fwd(c1.data(), N, q);
    u64 *ph = phase.data() + k * N;
    for (size_t i = 0; i < N; i++) ph[i] = mulmod(c1[i], sk.data[k * N + i], q);
    ntt_inv(ph, N, q);
    std::memcpy(c0.data(), ct.c0.data() + k * N, N * sizeof(u64));
    if (ct.ntt_form) ntt_inv(c0.data(), N, q);
    for (size_t i = 0; i < N; i++) ph[i] = (ph[i] + c0[i]) % q;
  }
  pt.assign(N, 0);
  if (K == 1) {
    for (size_t i = 0; i < N; i++) pt[i] = rescale(phase[i], qs[0], t);
    return;
  }
  const u64 q0 = qs[0], q1 = qs[1];
  for (size_t i = 0; i < N; i++) {
    const u64 p0 = phase[i], p1 = phase[N + i];
    const u64 diff = (p1 + q1 - p0 % q1) % q1;
    const u64 s = mulmod(diff, q0_inv_mod_q1, q1);
    const u128 mp = (u128)q0 * s + p0;
    u64 pp = (u64)((mp + ((u128)q1 >> 1)) / q1);
    if (pp >= q0) pp -= q0;
    pt[i] = rescale(pp, q0, t);
  }
}

int decrypt_and_budget(const RlweCt &ct, const RlweSk &sk, size_t N, u64 q, u64 t, std::vector<u64> &pt) {
  // rlwe.cpp:83-113
  std::vector<u64> c0(ct.c0.begin(), ct.c0.begin() + N), c1(ct.c1.begin(), ct.c1.begin() + N), ph(N);
  if (ct.ntt_form) ntt_inv(c0.data(), N, q);
  else ntt_fwd(c1.data(), N, q);
  for (size_t i = 0; i < N; i++) ph[i] = mulmod(c1[i], sk.data[i], q);
  ntt_inv(ph.data(), N, q);
  for (size_t i = 0; i < N; i++) ph[i] = (ph[i] + c0[i]) % q;
  pt.assign(N, 0);
  const u64 delta = q / t, half_q = q / 2;
  u64 max_noise = 0;
  for (size_t i = 0; i < N; i++) {
    const u64 m = rescale(ph[i], q, t);
    pt[i] = m;
    const u64 approx = (u64)((u128)delta * m % q);
    const u64 np = ph[i] >= approx ? ph[i] - approx : q - approx + ph[i];
    const u64 na = np > half_q ? q - np : np;
    if (na > max_noise) max_noise = na;
  }
  return max_noise > 0 ? (int)std::log2((double)q / (2.0 * t * max_noise)) : (int)std::log2((double)q / (2.0 * t));
}

// ---------------------------------------------------------------------------
// RGSW. src/gsw.cpp
// ---------------------------------------------------------------------------
void GSWEval::gsw_ntt_forward(GSWCt &g) const {  // gsw.cpp:105-117
  const size_t N = p->N(), K = p->K();
  for (auto &row : g)
    for (size_t i = 0; i < 2 * K; i++) ntt_fwd(row.data() + N * i, N, p->rns_mods[i % K]);
}

void GSWEval::decomp_single_mod(const RlweCt &ct, std::vector<std::vector<u64>> &out) const {
  // gsw.cpp:252-295: balanced digits, d <- (d - r) >> b, row (l-1-i).
  const size_t N = p->N();
  const u64 q = p->rns_mods[0];
  const int64_t qi = (int64_t)q;
  const u64 half_q = q >> 1;
  const int64_t native = 64 - (int64_t)base_log2;
  std::vector<int64_t> d(N);
  for (size_t poly = 0; poly < 2; poly++) {
    const u64 *src = ct.data(poly);
    for (size_t k = 0; k < N; k++) d[k] = src[k] > half_q ? (int64_t)src[k] - qi : (int64_t)src[k];
    for (size_t i = 0; i < l; i++) {
      u64 *row = out[poly * l + (l - 1 - i)].data();
      for (size_t k = 0; k < N; k++) {
        const int64_t r = (int64_t)((u64)d[k] << native) >> native;
        d[k] = (d[k] - r) >> (int64_t)base_log2;
        row[k] = r >= 0 ? (u64)r : (u64)(r + qi);
      }
    }
  }
}

void GSWEval::decomp_mp(const RlweCt &ct, std::vector<std::vector<u64>> &out) const {
  // gsw.cpp:184-250 with compose_rns_to_mp/decompose_mp_to_rns (gsw.cpp:22-85), K=2.
  const size_t N = p->N(), K = p->K();
  if (K != 2) throw std::runtime_error("decomp_mp: K=2 only");
  const u64 q0 = p->rns_mods[0], q1 = p->rns_mods[1];
  const u64 mask = (u64{1} << base_log2) - 1;
  for (size_t poly = 0; poly < 2; poly++) {
    const u64 *src = ct.data(poly);
    std::vector<u128> x(N);
    for (size_t k = 0; k < N; k++) {
      const u64 r0 = src[k], r1 = src[N + k];
      const u64 diff = (r1 + q1 - r0 % q1) % q1;
      const u64 s = mulmod(diff, p->q0_inv_mod_q1, q1);
      x[k] = (u128)q0 * s + r0;
    }
    for (size_t pp = l; pp-- > 0;) {
      std::vector<u64> &row = out[poly * l + (l - 1 - pp)];
      for (size_t k = 0; k < N; k++) {
        const u64 dig = (u64)(x[k] >> (pp * base_log2)) & mask;
        // decompose_mp_to_rns of (lo=dig, hi=0)
        row[k] = (u64)(((u128)(0 % q0) * p->r64_mod_q[0] + dig % q0) % q0);
        row[N + k] = (u64)(((u128)(0 % q1) * p->r64_mod_q[1] + dig % q1) % q1);
      }
    }
  }
}

void GSWEval::decomp_to_ntt(std::vector<std::vector<u64>> &rows) const {  // gsw.cpp:297-315
  const size_t N = p->N(), K = p->K();
  for (auto &r : rows)
    for (size_t i = 0; i < K; i++) ntt_fwd(r.data() + N * i, N, p->rns_mods[i]);
}

void GSWEval::external_product(const GSWCt &g, const RlweCt &bfv, RlweCt &res) const {
  // gsw.cpp:119-182: decompose, NTT rows, out_k = sum_j row_j * g[j][k].
  const size_t N = p->N(), K = p->K(), KN = N * K;
  std::vector<std::vector<u64>> rows(2 * l, std::vector<u64>(KN));
  if (K == 1) decomp_single_mod(bfv, rows);
  else decomp_mp(bfv, rows);
  decomp_to_ntt(rows);
  res.c0.resize(KN);
  res.c1.resize(KN);
  for (size_t k = 0; k < 2; k++) {
    u64 *out = res.data(k);
    for (size_t m = 0; m < K; m++) {
      const u64 q = p->rns_mods[m];
      const size_t off = m * N;
      for (size_t i = 0; i < N; i++) {
        u64 acc = mulmod(rows[0][off + i], g[0][k * KN + off + i], q);
        for (size_t j = 1; j < 2 * l; j++)
          acc = (acc + mulmod(rows[j][off + i], g[j][k * KN + off + i], q)) % q;
        out[off + i] = acc;
      }
    }
  }
  res.ntt_form = true;
}

void GSWEval::query_to_gsw(std::vector<RlweCt> query, const GSWCt &key, GSWCt &out) const {
  // gsw.cpp:317-357: top rows = NTT(query rows), bottom row i = EP(RGSW(s), query_i).
  const size_t n = query.size(), KN = p->N() * p->K();
  out.assign(n, {});
  for (size_t i = 0; i < n; i++) {
    out[i].insert(out[i].end(), query[i].c0.begin(), query[i].c0.begin() + KN);
    out[i].insert(out[i].end(), query[i].c1.begin(), query[i].c1.begin() + KN);
  }
  gsw_ntt_forward(out);
  out.resize(2 * n);
  for (size_t i = 0; i < n; i++) {
    external_product(key, query[i], query[i]);
    out[i + n].insert(out[i + n].end(), query[i].c0.begin(), query[i].c0.begin() + KN);
    out[i + n].insert(out[i + n].end(), query[i].c1.begin(), query[i].c1.begin() + KN);
  }
}

GSWCt GSWEval::plain_to_gsw(const std::vector<u64> &pt, const RlweSk &sk, std::mt19937_64 &rng) const {
  // gsw.cpp:359-422
  const size_t N = p->N(), K = p->K();
  const auto &qs = p->rns_mods;
  const auto gad = gsw_gadget(l, base_log2, qs);
  GSWCt out(2 * l, std::vector<u64>(2 * K * N));
  const u64 q0 = qs[0], half_q0 = q0 >> 1;
  auto canon = [&](size_t c, u64 qk) -> u64 {
    if (K == 1 || pt[c] <= half_q0) return pt[c] % qk;
    const u64 a = q0 - pt[c];
    return a >= qk ? (qk - (a % qk)) % qk : qk - a;
  };
  RlweCt ct;
  for (size_t half = 0; half < 2; half++) {
    for (size_t r = 0; r < l; r++) {
      encrypt_zero_rns(sk, N, qs, p->sigma(), rng, ct, false);
      for (size_t k = 0; k < K; k++) {
        const u64 qk = qs[k], gk = gad[k][r];
        u64 *tgt = ct.data(half) + k * N;
        for (size_t c = 0; c < N; c++) tgt[c] = (tgt[c] + (u64)((u128)canon(c, qk) * gk % qk)) % qk;
      }
      for (size_t k = 0; k < K; k++) {
        ntt_fwd(ct.c0.data() + k * N, N, qs[k]);
        ntt_fwd(ct.c1.data() + k * N, N, qs[k]);
      }
      const size_t row = half * l + r;
      std::memcpy(out[row].data(), ct.c0.data(), K * N * sizeof(u64));
      std::memcpy(out[row].data() + K * N, ct.c1.data(), K * N * sizeof(u64));
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// BV keyswitch. src/bv_keyswitch.cpp
// ---------------------------------------------------------------------------
const BvKey &BvGaloisKeys::get(uint32_t k) const {
  for (auto &x : keys)
    if (x.galois_k == k) return x;
  throw std::out_of_range("BvGaloisKeys::get: galois_k not found");
}

void signed_gadget_decompose(u64 val, size_t base_log2, u64 q, u64 *out, size_t nd) {
  const u64 half_q = q >> 1;
  const int64_t ns = 64 - (int64_t)base_log2;
  int64_t d = val > half_q ? (int64_t)val - (int64_t)q : (int64_t)val;
  for (size_t i = 0; i < nd; ++i) {
    const int64_t r = (int64_t)((u64)d << ns) >> ns;
    d -= r;
    d >>= base_log2;
    out[i] = r >= 0 ? (u64)r : (u64)(r + (int64_t)q);
  }
}

void signed_gadget_decompose_mp(u128 val, u128 Q, size_t base_log2, int64_t *out, size_t nd) {
  const u128 half_Q = Q >> 1;
  const u64 B = u64{1} << base_log2, half_B = B >> 1, mask = B - 1;
  i128 d = val > half_Q ? (i128)val - (i128)Q : (i128)val;
  for (size_t i = 0; i < nd; ++i) {
    const u64 low = (u64)d & mask;
    int64_t r;
    if (low > half_B) { r = (int64_t)low - (int64_t)B; d = (d >> base_log2) + 1; }
    else { r = (int64_t)low; d >>= base_log2; }
    out[i] = r;
  }
}

BvKey gen_bv_ks_key(const PirParams &p, const RlweSk &sk, uint32_t galois_k, std::mt19937_64 &rng) {
  const size_t N = p.N(), K = p.K();
  const auto &qs = p.rns_mods;
  std::vector<u64> sigma_s(N * K);
  for (size_t k = 0; k < K; k++)
    automorphism_ntt(sk.data.data() + k * N, N, galois_k, qs[k], sigma_s.data() + k * N);
  BvKey key;
  key.galois_k = galois_k;
  const size_t b = p.bv_base_log2();
  key.rows.resize(p.l_ks());
  std::vector<u64> msg(N * K);
  for (size_t i = 0; i < p.l_ks(); i++) {
    for (size_t k = 0; k < K; k++) {
      u64 Bi = 1 % qs[k];
      for (size_t t = 0; t < i * b; t++) Bi = (u64)(((u128)Bi << 1) % qs[k]);
      for (size_t c = 0; c < N; c++) msg[k * N + c] = mulmod(sigma_s[k * N + c], Bi, qs[k]);
    }
    // build_ksk_row_rns (bv_keyswitch.cpp:191-220)
    BvRow &row = key.rows[i];
    row.a.assign(N * K, 0);
    row.b.assign(N * K, 0);
    for (size_t k = 0; k < K; k++) sample_uniform_poly(row.a.data() + k * N, N, qs[k], rng);
    std::vector<u64> e(N * K);
    sample_gaussian_shared(e.data(), N, qs, p.sigma(), rng);
    for (size_t k = 0; k < K; k++) {
      const u64 q = qs[k];
      u64 *a = row.a.data() + k * N, *bb = row.b.data() + k * N, *ek = e.data() + k * N;
      ntt_fwd(a, N, q);
      ntt_fwd(ek, N, q);
      for (size_t c = 0; c < N; c++) {
        const u64 as = mulmod(a[c], sk.data[k * N + c], q);
        const u64 v = (msg[k * N + c] + q - as) % q;
        bb[c] = (v + ek[c]) % q;
      }
    }
  }
  return key;
}

BvGaloisKeys gen_bv_galois_keys(const PirParams &p, const RlweSk &sk, u64 seed) {
  // bv_keyswitch.cpp:261-276, with an explicit seed for the key RNG.
  BvGaloisKeys r;
  std::mt19937_64 rng(seed);
  for (size_t i = 0; i < p.expansion_height; ++i)
    r.keys.push_back(gen_bv_ks_key(p, sk, (uint32_t)((p.N() >> i) + 1), rng));
  return r;
}

static void bv_apply_k1(RlweCt &ct, uint32_t galois_k, const BvKey &key, const PirParams &p) {
  // bv_keyswitch.cpp:296-367
  const size_t N = p.N(), L = p.l_ks(), b = p.bv_base_log2();
  const u64 q = p.rns_mods[0];
  std::vector<u64> c0p(N), c1p(N), da(N), db(N), dig(L * N), tmp(L);
  automorphism_coeff(ct.c0.data(), N, galois_k, q, c0p.data());
  automorphism_coeff(ct.c1.data(), N, galois_k, q, c1p.data());
  for (size_t k = 0; k < N; k++) {
    signed_gadget_decompose(c1p[k], b, q, tmp.data(), L);
    for (size_t i = 0; i < L; i++) dig[i * N + k] = tmp[i];
  }
  for (size_t i = 0; i < L; i++) {
    u64 *d = dig.data() + i * N;
    ntt_fwd(d, N, q);
    for (size_t c = 0; c < N; c++) {
      const u64 pb = mulmod(d[c], key.rows[i].b[c], q), pa = mulmod(d[c], key.rows[i].a[c], q);
      db[c] = i == 0 ? pb : (db[c] + pb) % q;
      da[c] = i == 0 ? pa : (da[c] + pa) % q;
    }
  }
  ntt_inv(db.data(), N, q);
  ntt_inv(da.data(), N, q);
  for (size_t c = 0; c < N; c++) c0p[c] = (c0p[c] + db[c]) % q;
  ct.c0 = c0p;
  ct.c1 = da;
}

static void bv_apply_k2(RlweCt &ct, uint32_t galois_k, const BvKey &key, const PirParams &p) {
  // bv_keyswitch.cpp:373-481
  const size_t N = p.N(), L = p.l_ks(), b = p.bv_base_log2();
  const auto &qs = p.rns_mods;
  const u64 q0 = qs[0], q1 = qs[1];
  std::vector<u64> c0p(2 * N), c1p(2 * N);
  for (size_t k = 0; k < 2; k++) {
    automorphism_coeff(ct.c0.data() + k * N, N, galois_k, qs[k], c0p.data() + k * N);
    automorphism_coeff(ct.c1.data() + k * N, N, galois_k, qs[k], c1p.data() + k * N);
  }
  const u128 Q = (u128)q0 * q1;
  std::vector<int64_t> sd(L * N), buf(L);
  for (size_t j = 0; j < N; j++) {
    const u64 r0 = c1p[j], r1 = c1p[N + j];
    const u64 diff = (r1 + q1 - r0 % q1) % q1;
    const u64 s = mulmod(diff, p.q0_inv_mod_q1, q1);
    signed_gadget_decompose_mp((u128)q0 * s + r0, Q, b, buf.data(), L);
    for (size_t i = 0; i < L; i++) sd[i * N + j] = buf[i];
  }
  std::vector<u64> da(2 * N), db(2 * N), d(N);
  for (size_t k = 0; k < 2; k++) {
    const u64 qk = qs[k];
    for (size_t i = 0; i < L; i++) {
      for (size_t j = 0; j < N; j++) {
        const int64_t s = sd[i * N + j];
        d[j] = s >= 0 ? (u64)s : qk - (u64)(-s);
      }
      ntt_fwd(d.data(), N, qk);
      for (size_t c = 0; c < N; c++) {
        const u64 pb = mulmod(d[c], key.rows[i].b[k * N + c], qk);
        const u64 pa = mulmod(d[c], key.rows[i].a[k * N + c], qk);
        db[k * N + c] = i == 0 ? pb : (db[k * N + c] + pb) % qk;
        da[k * N + c] = i == 0 ? pa : (da[k * N + c] + pa) % qk;
      }
    }
    ntt_inv(db.data() + k * N, N, qk);
    ntt_inv(da.data() + k * N, N, qk);
    for (size_t c = 0; c < N; c++) c0p[k * N + c] = (c0p[k * N + c] + db[k * N + c]) % qk;
  }
  ct.c0 = c0p;
  ct.c1 = da;
}

void bv_apply_galois_inplace(RlweCt &ct, uint32_t galois_k, const BvKey &key, const PirParams &p) {
  assert(key.galois_k == galois_k && !ct.ntt_form);
  if (p.K() == 1) bv_apply_k1(ct, galois_k, key, p);
  else bv_apply_k2(ct, galois_k, key, p);
}

// ---------------------------------------------------------------------------
// First-dim mat-mat. src/matrix.cpp
// ---------------------------------------------------------------------------
void level_mat_mat_32(const uint32_t *A, const uint32_t *B, u64 *out, size_t m, size_t n, size_t levels,
                      u64 q) {
  // matrix.cpp:232-293,315-350: exact sum of u32*u32 products, Barrett-u128 reduced.
  const BarrettU128 br = barrett_u128_setup(q);
  for (size_t lv = 0; lv < levels; ++lv) {
    const uint32_t *Ap = A + lv * m * n, *Bp = B + lv * n * 2;
    u64 *Cp = out + lv * m * 2;
    for (size_t i = 0; i < m; ++i) {
      u128 t0 = 0, t1 = 0;
      for (size_t k = 0; k < n; ++k) {
        t0 += (u128)((u64)Ap[i * n + k] * Bp[2 * k]);
        t1 += (u128)((u64)Ap[i * n + k] * Bp[2 * k + 1]);
      }
      Cp[2 * i] = barrett_reduce_u128(t0, br);
      Cp[2 * i + 1] = barrett_reduce_u128(t1, br);
    }
  }
}

void level_mat_mat_u64(const u64 *A, const u64 *B, u128 *out, size_t m, size_t n, size_t levels,
                       const u64 *level_qs) {
  // matrix.cpp:24-151: chunked u128 accumulation, reduced between chunks.
  for (size_t lv = 0; lv < levels; ++lv) {
    const u64 q = level_qs[lv];
    const u128 qi = q, q2 = qi * qi, MAX = ~(u128)0;
    size_t chunk = n;
    if (q != 0 && q2 != 0) {
      const u128 r = (MAX - qi) / q2;
      chunk = r == 0 ? 1 : (r >= (u128)n ? n : (size_t)r);
    }
    const u64 *Ap = A + lv * m * n, *Bp = B + lv * n * 2;
    u128 *Cp = out + lv * m * 2;
    for (size_t i = 0; i < m; i++) {
      u128 a0 = 0, a1 = 0;
      for (size_t base = 0; base < n; base += chunk) {
        const size_t end = std::min(base + chunk, n);
        u128 t0 = 0, t1 = 0;
        for (size_t k = base; k < end; k++) {
          t0 += (u128)Ap[i * n + k] * Bp[2 * k];
          t1 += (u128)Ap[i * n + k] * Bp[2 * k + 1];
        }
        a0 = q ? (a0 + t0) % qi : a0 + t0;
        a1 = q ? (a1 + t1) % qi : a1 + t1;
      }
      Cp[2 * i] = a0;
      Cp[2 * i + 1] = a1;
    }
  }
}

}  // namespace oracle
