// OnionPIRv2 CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Client (query pack, keys, decrypt, response decode) and server answer path.
#include <algorithm>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_set>

#include "oracle.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Client. src/client.cpp
// ---------------------------------------------------------------------------
PirClient::PirClient(const PirParams &p, size_t id, u64 s)
    : client_id(id), params(p), rng(s), seed(s) {
  sk = gen_secret_key_rns(p.N(), p.rns_mods, rng);  // client.cpp:14-23
}

std::vector<size_t> PirClient::get_query_indices(const PirParams &qp, size_t pt_idx) const {
  // client.cpp:48-98
  const size_t col = pt_idx % qp.fst_dim_sz, row = pt_idx / qp.fst_dim_sz;
  const size_t other = qp.other_dim_sz(), d = qp.num_dims, h = d - 1;
  std::vector<size_t> idx = {col};
  if (d == 1) return idx;
  const size_t r = 2 * other - (size_t{1} << h), sl = other - r;
  size_t perfect;
  if (row < other - r) { idx.push_back(0); perfect = row; }
  else {
    const size_t c = row - sl;
    idx.push_back(c % 2);
    perfect = sl + c / 2;
  }
  if (h > 1)
    for (size_t k = h - 2; k + 1 > 0; k--) {
      idx.push_back((perfect >> k) & 1);
      if (k == 0) break;
    }
  return idx;
}

RlweCt PirClient::fast_generate_query(const PirParams &qp, size_t pt_idx) {
  // client.cpp:107-166
  if (!params.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  if (pt_idx >= qp.num_pt) throw std::out_of_range("PIR plaintext index exceeds the runtime layout");
  const size_t N = qp.N(), K = qp.K();
  const auto &qs = qp.rns_mods;
  const u64 t = qp.plain_mod;
  auto idx = get_query_indices(qp, pt_idx);
  const size_t h = qp.expansion_height, cap = size_t{1} << h;
  u64 inv = 0;
  try_invert_uint_mod(cap, t, inv);
  const size_t rev = bit_reverse(idx[0], h);
  RlweCt q;
  encrypt_zero_rns(sk, N, qs, qp.sigma(), rng, q, false);
  if (K == 1) {
    const u64 Q = qs[0];
    const u64 scaled = (u64)(((u128)Q * inv + (t >> 1)) / t) % Q;
    q.c0[rev] = (u64)(((u128)q.c0[rev] + scaled) % Q);
  } else {
    const u128 Q = (u128)qs[0] * qs[1];
    const u128 Delta = Q / t;
    const u64 r = (u64)(Q - Delta * t);
    const u64 rr = (u64)(((u128)r * inv + (t >> 1)) / t);
    const u128 sm = Delta * inv + rr;
    for (size_t k = 0; k < K; ++k) {
      const size_t j = k * N + rev;
      q.c0[j] = (u64)(((u128)q.c0[j] + (u64)(sm % qs[k])) % qs[k]);
    }
  }
  add_gsw_to_query(qp, q, idx);
  return q;
}

void PirClient::add_gsw_to_query(const PirParams &qp, RlweCt &query, const std::vector<size_t> &idx) {
  // client.cpp:174-225
  if (!params.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  if (idx.size() == 1) return;
  const size_t h = qp.expansion_height, cap = size_t{1} << h, l = qp.l_ep(), K = qp.K(), N = qp.N();
  const auto &mods = qp.rns_mods;
  std::vector<u64> inv(K);
  for (size_t k = 0; k < K; k++) try_invert_uint_mod(cap, mods[k], inv[k]);
  const auto gad = gsw_gadget(l, qp.base_log2, mods);
  for (size_t i = 1; i < idx.size(); i++) {
    if (idx[i] != 1) continue;
    for (size_t k = 0; k < l; k++) {
      const size_t pos = qp.fst_dim_sz + (i - 1) * l + k;
      const size_t rev = bit_reverse(pos, h);
      for (size_t m = 0; m < K; m++) {
        const u64 coef = (u64)((u128)gad[m][k] * inv[m] % mods[m]);
        u64 &v = query.c0[rev + m * N];
        v = (u64)(((u128)v + coef) % mods[m]);
      }
    }
  }
}

BvGaloisKeys PirClient::create_bv_galois_keys() {
  // client.h:38-40; the reference seeds this RNG from random_device
  // (bv_keyswitch.cpp:267). Here it is derived from the client seed.
  u64 z = seed + 0x9E3779B97F4A7C15ULL;  // splitmix64 step
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return gen_bv_galois_keys(params, sk, z);
}

GSWCt PirClient::generate_gsw_from_key() {
  // client.cpp:25-38
  const size_t N = params.N();
  const u64 q0 = params.rns_mods[0];
  std::vector<u64> skc(sk.data.begin(), sk.data.begin() + N);
  ntt_inv(skc.data(), N, q0);
  GSWEval g(params, params.l_key(), params.base_log2_key);
  return g.plain_to_gsw(skc, sk, rng);
}

std::vector<u64> PirClient::decrypt_ct(const RlweCt &ct) const {
  std::vector<u64> pt;
  decrypt_rns(ct, sk, params.N(), params.rns_mods, params.plain_mod, params.q0_inv_mod_q1, pt);
  return pt;
}

RlweCt PirClient::load_resp_from_stream(const std::string &bytes) const {
  // client.cpp:315-361
  const size_t w = small_q_width(params.small_q), N = params.N();
  RlweCt r;
  r.c0.assign(N, 0);
  r.c1.assign(N, 0);
  size_t pos = 0, left = 0;
  uint8_t cur = 0;
  auto next_bit = [&]() -> u64 {
    if (left == 0) {
      if (pos >= bytes.size()) throw std::runtime_error("unexpected end of response stream");
      cur = (uint8_t)bytes[pos++];
      left = 8;
    }
    const u64 b = cur & 1;
    cur >>= 1;
    --left;
    return b;
  };
  for (size_t p = 0; p < 2; p++)
    for (size_t i = 0; i < N; i++) {
      u64 v = 0;
      for (size_t j = 0; j < w; j++) v |= next_bit() << j;
      r.data(p)[i] = v;
    }
  return r;
}

std::vector<u64> PirClient::decrypt_mod_q(const RlweCt &ct, int *budget) const {
  // client.cpp:364-430
  const size_t N = params.N();
  const u64 q = params.small_q, t = params.plain_mod, old_q = params.rns_mods[0];
  std::vector<u64> c0(N), c1(N), ph(N);
  for (size_t i = 0; i < N; i++) { c0[i] = ct.c0[i] % q; c1[i] = ct.c1[i] % q; }
  std::vector<u64> skc(sk.data.begin(), sk.data.begin() + N);
  ntt_inv(skc.data(), N, old_q);
  std::vector<u64> sks(N);
  for (size_t i = 0; i < N; i++) sks[i] = skc[i] > 1 ? q - 1 : skc[i];
  ntt_fwd(sks.data(), N, q);
  ntt_fwd(c1.data(), N, q);
  for (size_t i = 0; i < N; i++) ph[i] = mulmod(c1[i], sks[i], q);
  ntt_inv(ph.data(), N, q);
  for (size_t i = 0; i < N; i++) ph[i] = (ph[i] + c0[i]) % q;
  std::vector<u64> m(N);
  const u64 half_q = q / 2;
  u64 max_noise = 0;
  for (size_t i = 0; i < N; i++) {
    m[i] = (u64)(((u128)ph[i] * t + (q >> 1)) / q) % t;
    const u64 approx = (u64)(((u128)q * m[i] + (t >> 1)) / t) % q;
    const u64 np = ph[i] >= approx ? ph[i] - approx : q - approx + ph[i];
    const u64 na = np > half_q ? q - np : np;
    if (na > max_noise) max_noise = na;
  }
  if (budget)
    *budget = max_noise > 0 ? (int)std::log2((double)q / (2.0 * t * max_noise))
                            : (int)std::log2((double)q / (2.0 * t));
  return m;
}

// ---------------------------------------------------------------------------
// Server. src/server.cpp
// ---------------------------------------------------------------------------
PirServer::PirServer(const PirParams &p) : params(p), num_pt(p.num_pt) {
  const size_t cnt = num_pt * p.N() * p.K();
  if (p.composite) { db_lo.assign(cnt, 0); db_hi.assign(cnt, 0); }
  else db_u64.assign(cnt, 0);
}

void PirServer::load_data(size_t logical, const PlaintextSource &src, const std::vector<size_t> &record) {
  // server.cpp:80-203
  if (!src) throw std::invalid_argument("PirServer::load_data requires a source");
  if (logical > num_pt) throw std::invalid_argument("PirServer::load_data logical plaintext count exceeds its layout");
  for (size_t r : record)
    if (r >= num_pt) throw std::invalid_argument("PirServer::load_data record index exceeds its layout");
  recorded.clear();
  std::unordered_set<size_t> rs(record.begin(), record.end());
  const size_t N = params.N(), K = params.K();
  const u64 t = params.plain_mod;
  std::vector<u64> pt(N), buf;
  for (size_t id = 0; id < num_pt; id++) {
    if (id < logical) {
      buf.clear();
      src(id, buf);
      if (buf.size() != N) throw std::invalid_argument("Plaintext source returned the wrong coefficient count");
      for (size_t i = 0; i < N; i++) {
        if (buf[i] >= t) throw std::invalid_argument("Plaintext source coefficient exceeds the plaintext modulus");
        pt[i] = buf[i];
      }
    } else {
      std::fill(pt.begin(), pt.end(), 0);
    }
    if (rs.count(id)) recorded[id] = pt;
    if (params.composite) {
      std::vector<u64> x = pt;
      ntt_fwd(x.data(), N, params.rns_mods[0]);
      for (size_t c = 0; c < N; c++) {
        db_lo[c * num_pt + id] = (uint32_t)(x[c] % params.q1);
        db_hi[c * num_pt + id] = (uint32_t)(x[c] % params.q2);
      }
    } else {
      for (size_t k = 0; k < K; k++) {
        std::vector<u64> x(N);
        for (size_t i = 0; i < N; i++) x[i] = pt[i] % params.rns_mods[k];
        ntt_fwd(x.data(), N, params.rns_mods[k]);
        for (size_t c = 0; c < N; c++) db_u64[(k * N + c) * num_pt + id] = x[c];
      }
    }
  }
}

void PirServer::set_client_keys(size_t id, BvGaloisKeys bv, GSWCt gsw) {
  bv_keys[id] = std::move(bv);
  gsw_keys[id] = std::move(gsw);
}

std::vector<RlweCt> PirServer::fast_expand_qry(size_t id, const RlweCt &query) const {
  // server.cpp:731-816
  const size_t useful = params.fst_dim_sz + params.l_ep() * (params.num_dims - 1);
  const size_t h = params.expansion_height, cap = size_t{1} << h;
  const BvGaloisKeys &keys = bv_keys.at(id);
  const size_t N = params.N(), K = params.K();
  const auto &qs = params.rns_mods;
  std::vector<RlweCt> cts(2 * cap);
  cts[1] = query;
  for (size_t i = 1; i < cap; ++i) {
    const size_t k = size_t{1} << (std::bit_width(i) - 1);
    const size_t left_leaf = i * cap / k - cap;
    if (left_leaf >= useful) continue;
    RlweCt cp = cts[i];
    const uint32_t gk = (uint32_t)(N / k + 1);
    bv_apply_galois_inplace(cp, gk, keys.get(gk), params);
    RlweCt &a = cts[i], &l = cts[2 * i], &r = cts[2 * i + 1];
    l.c0.resize(K * N); l.c1.resize(K * N);
    for (size_t m = 0; m < K; m++) {
      const u64 q = qs[m];
      for (size_t c = 0; c < N; c++) {
        const size_t j = m * N + c;
        l.c0[j] = (a.c0[j] + cp.c0[j]) % q;
        l.c1[j] = (a.c1[j] + cp.c1[j]) % q;
        a.c0[j] = (a.c0[j] + q - cp.c0[j]) % q;
        a.c1[j] = (a.c1[j] + q - cp.c1[j]) % q;
      }
    }
    r.c0.resize(K * N); r.c1.resize(K * N);
    for (size_t m = 0; m < K; m++) {
      negacyclic_shift(a.c0.data() + m * N, N, (size_t)(-(int64_t)k), qs[m], r.c0.data() + m * N);
      negacyclic_shift(a.c1.data() + m * N, N, (size_t)(-(int64_t)k), qs[m], r.c1.data() + m * N);
    }
  }
  return std::vector<RlweCt>(cts.begin() + cap, cts.begin() + cap + useful);
}

std::vector<GSWCt> PirServer::complete_selectors(size_t id, const std::vector<RlweCt> &exp) const {
  // server.cpp:823-855
  const size_t l = params.l_ep(), fst = params.fst_dim_sz, ns = params.num_dims - 1;
  std::vector<GSWCt> out(ns);
  if (ns == 0) return out;
  if (exp.size() < fst + ns * l) throw std::invalid_argument("complete_selectors expects one expanded row group per selector");
  const GSWCt &key = gsw_keys.at(id);
  GSWEval kg(params, params.l_key(), params.base_log2_key);
  for (size_t i = 1; i <= ns; i++) {
    std::vector<RlweCt> v(exp.begin() + fst + (i - 1) * l, exp.begin() + fst + i * l);
    kg.query_to_gsw(v, key, out[i - 1]);
  }
  return out;
}

std::vector<RlweCt> PirServer::evaluate_first_dim(std::vector<RlweCt> &q) {
  // server.cpp:313-405, 529-597
  const size_t fst = params.fst_dim_sz, other = params.other_dim_sz(), N = params.N(), K = params.K();
  const auto &qs = params.rns_mods;
  std::vector<RlweCt> result;
  if (params.composite) {
    const u64 Q = qs[0];
    for (auto &ct : q) { ntt_fwd(ct.c0.data(), N, Q); ntt_fwd(ct.c1.data(), N, Q); ct.ntt_form = true; }
    std::vector<uint32_t> blo(N * fst * 2), bhi(N * fst * 2);
    for (size_t c = 0; c < N; c++)
      for (size_t i = 0; i < fst; i++)
        for (size_t p = 0; p < 2; p++) {
          const u64 v = q[i].data(p)[c];
          blo[c * fst * 2 + i * 2 + p] = (uint32_t)(v % params.q1);
          bhi[c * fst * 2 + i * 2 + p] = (uint32_t)(v % params.q2);
        }
    std::vector<u64> ilo(other * N * 2), ihi(other * N * 2);
    level_mat_mat_32(db_lo.data(), blo.data(), ilo.data(), other, fst, N, params.q1);
    level_mat_mat_32(db_hi.data(), bhi.data(), ihi.data(), other, fst, N, params.q2);
    for (size_t j = 0; j < other; j++) {
      RlweCt ct;
      ct.c0.assign(N, 0); ct.c1.assign(N, 0);
      for (size_t c = 0; c < N; c++)
        for (size_t p = 0; p < 2; p++) {
          const size_t b = c * other * 2 + j * 2 + p;
          const u64 lo = ilo[b], hi = ihi[b];
          const u64 diff = (hi + params.q2 - lo % params.q2) % params.q2;
          const u64 kk = (diff * params.q1_inv_mod_q2) % params.q2;
          ct.data(p)[c] = lo + params.q1 * kk;
        }
      ntt_inv(ct.c0.data(), N, Q);
      ntt_inv(ct.c1.data(), N, Q);
      ct.ntt_form = false;
      result.push_back(std::move(ct));
    }
    return result;
  }
  // Standard path: server.cpp:205-262,363-527
  for (auto &ct : q) {
    for (size_t m = 0; m < K; m++) {
      ntt_fwd(ct.c0.data() + m * N, N, qs[m]);
      ntt_fwd(ct.c1.data() + m * N, N, qs[m]);
    }
    ct.ntt_form = true;
  }
  const size_t L = N * K;
  std::vector<u64> B(L * fst * 2), lq(L);
  for (size_t lv = 0; lv < L; lv++) {
    lq[lv] = qs[lv / N];
    for (size_t i = 0; i < fst; i++) {
      B[lv * fst * 2 + i * 2] = q[i].c0[lv];
      B[lv * fst * 2 + i * 2 + 1] = q[i].c1[lv];
    }
  }
  std::vector<u128> inter(other * L * 2);
  level_mat_mat_u64(db_u64.data(), B.data(), inter.data(), other, fst, L, lq.data());
  for (size_t j = 0; j < other; j++) {
    RlweCt ct;
    ct.c0.assign(L, 0); ct.c1.assign(L, 0);
    for (size_t lv = 0; lv < L; lv++) {
      ct.c0[lv] = (u64)inter[lv * other * 2 + j * 2];
      ct.c1[lv] = (u64)inter[lv * other * 2 + j * 2 + 1];
    }
    for (size_t m = 0; m < K; m++) {
      ntt_inv(ct.c0.data() + m * N, N, qs[m]);
      ntt_inv(ct.c1.data() + m * N, N, qs[m]);
    }
    result.push_back(std::move(ct));
  }
  return result;
}

void PirServer::ext_prod_mux(RlweCt &x, RlweCt &y, const GSWCt &sel, RlweCt &res) const {
  // server.cpp:652-722: res = x + INTT(EP(sel, y - x)).
  const size_t N = params.N(), K = params.K();
  const auto &qs = params.rns_mods;
  for (size_t m = 0; m < K; m++)
    for (size_t c = 0; c < N; c++) {
      const size_t j = m * N + c;
      y.c0[j] = (y.c0[j] + qs[m] - x.c0[j]) % qs[m];
      y.c1[j] = (y.c1[j] + qs[m] - x.c1[j]) % qs[m];
    }
  GSWEval dg(params, params.l_ep(), params.base_log2);
  dg.external_product(sel, y, y);
  for (size_t m = 0; m < K; m++) {
    ntt_inv(y.c0.data() + m * N, N, qs[m]);
    ntt_inv(y.c1.data() + m * N, N, qs[m]);
  }
  y.ntt_form = false;
  RlweCt out;
  out.c0.resize(K * N); out.c1.resize(K * N);
  for (size_t m = 0; m < K; m++)
    for (size_t c = 0; c < N; c++) {
      const size_t j = m * N + c;
      out.c0[j] = (x.c0[j] + y.c0[j]) % qs[m];
      out.c1[j] = (x.c1[j] + y.c1[j]) % qs[m];
    }
  out.ntt_form = x.ntt_form;
  res = std::move(out);
}

RlweCt PirServer::evaluate_other_dim(std::vector<RlweCt> &mid, std::vector<GSWCt> &sel) const {
  // server.cpp:599-649
  if (params.num_dims == 1) return mid[0];
  const size_t h = params.num_dims - 1, other = params.other_dim_sz();
  const size_t perfect = size_t{1} << (h - 1);
  if (other < perfect)
    throw std::invalid_argument("evaluate_other_dim: layout has fewer candidates than its selector tree; query-shape views cannot answer make_query");
  const size_t last = 2 * other - (size_t{1} << h), offset = other - last;
  for (size_t i = 0; i < last; i += 2) {
    const size_t ci = i + offset;
    ext_prod_mux(mid[ci], mid[ci + 1], sel[0], mid[i / 2 + offset]);
  }
  for (size_t a = 1; a < sel.size(); a++) {
    const size_t half = (size_t{1} << (h - a)) >> 1;
    for (size_t i = 0; i < half; i++) ext_prod_mux(mid[i], mid[i + half], sel[a], mid[i]);
  }
  return mid[0];
}

u64 drop_q1(u64 r0, u64 r1, u64 q0, u64 q1, u64 q0_inv_mod_q1) {
  // server.cpp:1053-1061: Garner CRT of (r0 mod q0, r1 mod q1), then
  // round-half-up division by q1, wrapped into [0, q0)
  const u64 diff = (r1 + q1 - r0 % q1) % q1;
  const u64 s = mulmod(diff, q0_inv_mod_q1, q1);
  const u128 x = (u128)q0 * s + r0;
  const u64 o = (u64)((x + ((u128)q1 >> 1)) / q1);
  return o >= q0 ? o - q0 : o;
}

void PirServer::mod_switch_inplace(RlweCt &ct, u64 sq) const {
  // server.cpp:1025-1075
  const size_t N = params.N(), K = params.K();
  const auto &qs = params.rns_mods;
  if (K == 1) {
    for (size_t i = 0; i < N; i++) {
      ct.c0[i] = rescale(ct.c0[i], qs[0], sq);
      ct.c1[i] = rescale(ct.c1[i], qs[0], sq);
    }
    return;
  }
  const u64 q0 = qs[0], q1 = qs[1];
  for (size_t i = 0; i < N; i++) {
    ct.c0[i] = rescale(drop_q1(ct.c0[i], ct.c0[N + i], q0, q1, params.q0_inv_mod_q1), q0, sq);
    ct.c1[i] = rescale(drop_q1(ct.c1[i], ct.c1[N + i], q0, q1, params.q0_inv_mod_q1), q0, sq);
  }
  ct.c0.resize(N);
  ct.c1.resize(N);
}

RlweCt PirServer::make_query(size_t id, const RlweCt &query) {
  // server.cpp:895-947, timed under the reference stage names (logging.h:122-134)
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  auto t0 = clk::now();
  std::vector<RlweCt> qv = fast_expand_qry(id, query);
  auto t1 = clk::now();
  std::vector<GSWCt> gv = complete_selectors(id, qv);
  auto t2 = clk::now();
  qv.resize(params.fst_dim_sz);
  std::vector<RlweCt> mid = evaluate_first_dim(qv);
  auto t3 = clk::now();
  RlweCt res = evaluate_other_dim(mid, gv);
  auto t4 = clk::now();
  if (params.cfg.small_q_width < params.cfg.rns_mod_bits[0]) mod_switch_inplace(res, params.small_q);
  auto t5 = clk::now();
  last_times = {ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, t4), ms(t4, t5)};
  return res;
}

std::string PirServer::save_resp_to_stream(const RlweCt &resp) const {
  // server.cpp:950-1009: c0 then c1, small_q_width bits per coefficient, LSB-first.
  const size_t w = small_q_width(params.small_q), N = params.N();
  std::string out;
  uint8_t byte = 0;
  size_t filled = 0;
  for (size_t p = 0; p < 2; p++) {
    const u64 *d = resp.data(p);
    for (size_t i = 0; i < N; i++) {
      const u64 c = d[i] & ((1ULL << w) - 1);
      size_t wr = 0;
      while (wr < w) {
        const size_t take = std::min(8 - filled, w - wr);
        byte |= (uint8_t)(((c >> wr) & ((1ULL << take) - 1)) << filled);
        filled += take;
        wr += take;
        if (filled == 8) { out.push_back((char)byte); byte = 0; filled = 0; }
      }
    }
  }
  if (filled) out.push_back((char)byte);
  return out;
}

}  // namespace oracle
