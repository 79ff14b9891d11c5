// PirClient on the CPU: secret key, BV Galois keys, RGSW(s), packed query,
// response decoding and small-q decryption.
// Follows src/client.cpp:14-430, src/rlwe.cpp:174-276, src/gsw.cpp:359-422,
// src/bv_keyswitch.cpp:166-276, src/utils.cpp:407-443, for one logical
// modulus (K = 1: a prime, or the composite q1*q2 of CONFIG_N2048_K1_COMP)
// and for K = 2 RNS limbs (CONFIG_N4096_K2_MP). Polynomials are limb-major,
// K*N values, as in the reference.
#include <bit>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "codec.h"
#include "onionpir_host.h"
#include "prg.h"

namespace onionpir {

using u128 = unsigned __int128;

namespace {

u64 mmul(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }

u64 splitmix64(u64 x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// utils.cpp:421-430: rejection-sampled uniform values in [0, q)
void uniform_poly(u64 *out, size_t N, u64 q, std::mt19937_64 &rng) {
  const u64 limit = ~u64{0} - (~u64{0} % q);
  for (size_t i = 0; i < N; i++) {
    u64 r;
    do r = rng(); while (r >= limit);
    out[i] = r % q;
  }
}

// rlwe.cpp:196-215 / bv_keyswitch.cpp:166-185: one rounded Gaussian per
// coefficient, the same integer in every limb
void gaussian_shared(u64 *out, size_t N, const std::vector<u64> &qs, double sigma, std::mt19937_64 &rng) {
  std::normal_distribution<double> dist(0.0, sigma);
  std::vector<int64_t> e(N);
  for (size_t i = 0; i < N; i++) e[i] = std::llround(dist(rng));
  for (size_t k = 0; k < qs.size(); k++)
    for (size_t i = 0; i < N; i++) {
      const u64 mag = (u64)(e[i] < 0 ? -e[i] : e[i]) % qs[k];
      out[k * N + i] = (e[i] < 0 && mag) ? qs[k] - mag : mag;
    }
}

std::vector<std::shared_ptr<RingMod>> rings_for(const PirParams &p) {
  const size_t N = p.get_poly_degree();
  std::vector<std::shared_ptr<RingMod>> r;
  if (p.q1()) {  // composite Q = q1*q2: one ring over Q, run as two CRT limbs
    r.push_back(std::make_shared<RingMod>(N, p.get_rns_mods()[0], p.ntt_root(), p.q1(), p.q2()));
    return r;
  }
  for (u64 q : p.get_rns_mods()) r.push_back(std::make_shared<RingMod>(N, q, utils::minimal_primitive_root(2 * N, q)));
  return r;
}

}  // namespace

PirClient::PirClient(const PirParams &params, size_t client_id, u64 seed)
    : client_id_(client_id), params_(params), rng_(seed), seed_(seed), qs_(params.get_rns_mods()),
      rings_(rings_for(params)) {
  // rlwe.cpp:177-193,219-228: ternary secret {0, 1, -1} from uniform_int(0, 2),
  // the same integers in every limb, NTT form per limb
  const size_t N = params.get_poly_degree(), K = qs_.size();
  std::uniform_int_distribution<int> dist(0, 2);
  std::vector<int> tern(N);
  for (size_t i = 0; i < N; i++) tern[i] = dist(rng_);
  sk_.data.resize(K * N);
  for (size_t k = 0; k < K; k++) {
    for (size_t i = 0; i < N; i++) sk_.data[k * N + i] = tern[i] == 0 ? 0 : (tern[i] == 1 ? 1 : qs_[k] - 1);
    rings_[k]->ntt_fwd(sk_.data.data() + k * N);
  }
}

PirClient::PirClient(const PirParams &params, size_t client_id, const std::array<uint8_t, 32> &key)
    : client_id_(client_id), params_(params), rng_(0), seed_(0), qs_(params.get_rns_mods()),
      rings_(rings_for(params)), prg_(true), key_(key) {
  const size_t N = params.get_poly_degree(), K = qs_.size();
  cdt_ = prg::gaussian_cdt(params.get_noise_std_dev());
  std::vector<int8_t> tern(N);
  prg::ternary(key_, N, tern.data());
  sk_.data.resize(K * N);
  for (size_t k = 0; k < K; k++) {
    for (size_t i = 0; i < N; i++) sk_.data[k * N + i] = tern[i] == 0 ? 0 : (tern[i] == 1 ? 1 : qs_[k] - 1);
    rings_[k]->ntt_fwd(sk_.data.data() + k * N);
  }
}

void PirClient::shared_error(uint32_t domain, uint32_t i, std::vector<u64> &e) const {
  const size_t N = params_.get_poly_degree(), K = qs_.size();
  std::vector<int64_t> v(N);
  prg::gaussian(key_, domain, i, cdt_, N, v.data());
  e.assign(K * N, 0);
  for (size_t k = 0; k < K; k++)
    for (size_t j = 0; j < N; j++) {
      const u64 mag = (u64)(v[j] < 0 ? -v[j] : v[j]) % qs_[k];
      e[k * N + j] = (v[j] < 0 && mag) ? qs_[k] - mag : mag;
    }
}

void PirClient::encrypt_zero_seeded(RlweCt &ct, const std::array<uint8_t, 32> &qseed) {
  // c1 = uniform(qseed, kQueryA, 0, limb), c0 = -(c1*s + e), coefficient form
  const size_t N = params_.get_poly_degree(), K = qs_.size();
  ct.c0.assign(K * N, 0);
  ct.c1.assign(K * N, 0);
  for (size_t k = 0; k < K; k++) prg::uniform(qseed, prg::kQueryA, 0, (uint32_t)k, qs_[k], N, ct.c1.data() + k * N);
  std::vector<int64_t> ev(N);
  prg::gaussian(qseed, prg::kQueryE, 0, cdt_, N, ev.data());
  std::vector<u64> a(N);
  for (size_t k = 0; k < K; k++) {
    const u64 q = qs_[k];
    u64 *c0 = ct.c0.data() + k * N;
    std::memcpy(a.data(), ct.c1.data() + k * N, N * 8);
    rings_[k]->ntt_fwd(a.data());
    rings_[k]->mul(a.data(), sk_.data.data() + k * N, c0);
    rings_[k]->ntt_inv(c0);
    for (size_t i = 0; i < N; i++) {
      const u64 mag = (u64)(ev[i] < 0 ? -ev[i] : ev[i]) % q;
      const u64 e = (ev[i] < 0 && mag) ? q - mag : mag;
      const u64 s = (c0[i] + e) % q;
      c0[i] = s ? q - s : 0;
    }
  }
  ct.ntt_form = false;
}

void PirClient::encrypt_zero(RlweCt &ct) {
  // rlwe.cpp:230-276: c1 = a (uniform per limb), c0 = -(a*s + e), coefficient form
  const size_t N = params_.get_poly_degree(), K = qs_.size();
  ct.c0.assign(K * N, 0);
  ct.c1.assign(K * N, 0);
  for (size_t k = 0; k < K; k++) uniform_poly(ct.c1.data() + k * N, N, qs_[k], rng_);
  std::vector<u64> e(K * N), a(N);
  gaussian_shared(e.data(), N, qs_, params_.get_noise_std_dev(), rng_);
  for (size_t k = 0; k < K; k++) {
    const u64 q = qs_[k];
    u64 *c0 = ct.c0.data() + k * N;
    std::memcpy(a.data(), ct.c1.data() + k * N, N * 8);
    rings_[k]->ntt_fwd(a.data());
    rings_[k]->mul(a.data(), sk_.data.data() + k * N, c0);
    rings_[k]->ntt_inv(c0);
    for (size_t i = 0; i < N; i++) {
      const u64 s = (c0[i] + e[k * N + i]) % q;
      c0[i] = s ? q - s : 0;
    }
  }
  ct.ntt_form = false;
}

std::vector<size_t> PirClient::get_query_indices(const PirParams &qp, size_t pt_idx) const {
  // client.cpp:48-98: {first-dim column, deepest folded-pair bit, MSB-first bits}
  const size_t N0 = qp.get_fst_dim_sz(), col = pt_idx % N0, row = pt_idx / N0;
  const size_t other = qp.get_other_dim_sz(), d = qp.get_num_dims();
  std::vector<size_t> idx{col};
  if (d == 1) return idx;
  const size_t h = d - 1;
  const size_t deep = 2 * other - (size_t{1} << h);
  const size_t shallow = other - deep;
  size_t perfect = row;
  if (row < shallow) {
    idx.push_back(0);
  } else {
    idx.push_back((row - shallow) % 2);
    perfect = shallow + (row - shallow) / 2;
  }
  for (size_t k = h - 1; k-- > 0;) idx.push_back((perfect >> k) & 1);
  return idx;
}

RlweCt PirClient::fast_generate_query(size_t pt_idx) { return fast_generate_query(params_, pt_idx); }

RlweCt PirClient::fresh_zero_ct() {
  RlweCt ct;
  encrypt_zero(ct);
  return ct;
}

RlweCt PirClient::fast_generate_query(const PirParams &qp, size_t pt_idx) {
  // client.cpp:107-166
  if (!params_.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  if (pt_idx >= qp.get_num_pt()) throw std::out_of_range("PIR plaintext index exceeds the runtime layout");
  return generate_query_for_indices(qp, get_query_indices(qp, pt_idx));
}

RlweCt PirClient::generate_query_for_indices(const PirParams &qp, const std::vector<size_t> &idx) {
  // Delta * (2^h)^-1 (mod t rounding) at BitRev(column), then the selector rows
  if (!params_.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  const size_t N = qp.get_poly_degree(), K = qs_.size();
  const u64 t = qp.get_plain_mod();
  const size_t h = qp.get_expan_height();
  if (idx.empty() || idx[0] >= (size_t{1} << h)) throw std::invalid_argument("query column is outside the capacity");
  u64 inv = 0;
  utils::try_invert_uint_mod(size_t{1} << h, t, inv);
  const size_t slot = utils::bit_reverse(idx[0], h);
  RlweCt q;
  if (prg_) {
    last_qseed_ = prg::query_seed(key_, query_ctr_++);
    encrypt_zero_seeded(q, last_qseed_);
  } else {
    encrypt_zero(q);
  }
  if (K == 1) {
    const u64 Q = qs_[0];
    const u64 scaled = (u64)(((u128)Q * inv + (t >> 1)) / t) % Q;  // round(Q * inv / t)
    q.c0[slot] = (q.c0[slot] + scaled) % Q;
  } else {
    // Q = q0*q1 < 2^128: Delta*inv + round((Q - Delta*t) * inv / t), reduced per limb
    const u128 Q = (u128)qs_[0] * qs_[1];
    const u128 Delta = Q / t;
    const u64 r = (u64)(Q - Delta * t);
    const u64 rr = (u64)(((u128)r * inv + (t >> 1)) / t);
    const u128 sm = Delta * inv + rr;
    for (size_t k = 0; k < K; k++) {
      const size_t j = k * N + slot;
      q.c0[j] = (u64)(((u128)q.c0[j] + (u64)(sm % qs_[k])) % qs_[k]);
    }
  }
  add_gsw_to_query(qp, q, idx);
  return q;
}

void PirClient::add_gsw_to_query(const PirParams &qp, RlweCt &query, const std::vector<size_t> &idx) {
  // client.cpp:174-225: selector i with bit 1 adds B^(l-1-k) / 2^h at BitRev(N0 + (i-1)l + k)
  if (!params_.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  if (idx.size() == 1) return;
  const size_t N = qp.get_poly_degree(), K = qs_.size();
  const size_t h = qp.get_expan_height(), l = qp.get_l();
  std::vector<u64> inv(K);
  for (size_t m = 0; m < K; m++) utils::try_invert_uint_mod(size_t{1} << h, qs_[m], inv[m]);
  const auto gad = utils::gsw_gadget(l, qp.get_base_log2(), qs_);
  for (size_t i = 1; i < idx.size(); i++) {
    if (idx[i] != 1) continue;
    for (size_t k = 0; k < l; k++) {
      const size_t slot = utils::bit_reverse(qp.get_fst_dim_sz() + (i - 1) * l + k, h);
      for (size_t m = 0; m < K; m++) {
        u64 &v = query.c0[m * N + slot];
        v = (v + mmul(gad[m][k], inv[m], qs_[m])) % qs_[m];
      }
    }
  }
}

bvks::BvGaloisKeys PirClient::create_bv_galois_keys() {
  // bv_keyswitch.cpp:191-276; key RNG seeded from the client seed (the
  // reference draws that seed from random_device).
  const size_t N = params_.get_poly_degree(), L = params_.get_l_ks(), b = params_.get_bv_base_log2();
  const size_t K = qs_.size();
  std::mt19937_64 rng(splitmix64(seed_));
  bvks::BvGaloisKeys keys;
  std::vector<u64> s_coef(sk_.data), sigma_s(K * N), e(K * N);
  for (size_t k = 0; k < K; k++) rings_[k]->ntt_inv(s_coef.data() + k * N);
  for (size_t lvl = 0; lvl < params_.get_expan_height(); lvl++) {
    const uint32_t g = (uint32_t)((N >> lvl) + 1);
    // sigma_g(s) in coefficient form per limb, then NTT (utils.cpp:58-82)
    std::fill(sigma_s.begin(), sigma_s.end(), 0);
    for (size_t k = 0; k < K; k++) {
      const u64 q = qs_[k];
      const u64 *sc = s_coef.data() + k * N;
      u64 *ss = sigma_s.data() + k * N;
      for (size_t i = 0; i < N; i++) {
        const size_t dest = (i * g) % (2 * N);
        if (dest < N) ss[dest] = sc[i];
        else ss[dest - N] = sc[i] ? q - sc[i] : 0;
      }
      rings_[k]->ntt_fwd(ss);
    }
    bvks::BvKeySwitchKey key;
    key.galois_k = g;
    key.cts.resize(L);
    for (size_t r = 0; r < L; r++) {
      auto &row = key.cts[r];
      row.a.assign(K * N, 0);
      row.b.assign(K * N, 0);
      const uint32_t sid = (uint32_t)(lvl * L + r);
      if (prg_) {  // a drawn directly in NTT form; the error in coefficient form
        for (size_t k = 0; k < K; k++) prg::uniform(key_, prg::kBvA, sid, (uint32_t)k, qs_[k], N, row.a.data() + k * N);
        shared_error(prg::kBvE, sid, e);
      } else {
        for (size_t k = 0; k < K; k++) uniform_poly(row.a.data() + k * N, N, qs_[k], rng);
        gaussian_shared(e.data(), N, qs_, params_.get_noise_std_dev(), rng);
      }
      for (size_t k = 0; k < K; k++) {
        const u64 q = qs_[k];
        u64 Br = 1 % q;
        for (size_t z = 0; z < r * b; z++) Br = (u64)(((u128)Br << 1) % q);
        u64 *a = row.a.data() + k * N, *bb = row.b.data() + k * N, *ek = e.data() + k * N;
        if (!prg_) rings_[k]->ntt_fwd(a);
        rings_[k]->ntt_fwd(ek);
        for (size_t j = 0; j < N; j++) {
          // b = sigma(s) * B^r - a*s + e, NTT form
          const u64 msg = mmul(sigma_s[k * N + j], Br, q), as = mmul(a[j], sk_.data[k * N + j], q);
          bb[j] = ((msg + q - as) % q + ek[j]) % q;
        }
      }
    }
    keys.keys.push_back(std::move(key));
  }
  return keys;
}

GSWCt PirClient::generate_gsw_from_key() {
  // client.cpp:25-38 + gsw.cpp:359-422: RGSW(s) with the key gadget; the
  // secret's coefficients (0, 1, -1) are carried into every limb.
  const size_t N = params_.get_poly_degree(), l = params_.get_l_key(), K = qs_.size();
  std::vector<u64> s_coef(sk_.data.begin(), sk_.data.begin() + N);
  rings_[0]->ntt_inv(s_coef.data());
  const u64 q0 = qs_[0];
  const auto gad = utils::gsw_gadget(l, params_.get_base_log2_key(), qs_);
  GSWCt out(2 * l, std::vector<u64>(2 * K * N));
  if (prg_) {
    // row = half * l + r: c1 = uniform(kGswA, row) in NTT form, c0 = -(c1*s + e)
    // + g_r * (half ? s^2 : s): the same phase as the unseeded rows, whose c1
    // is a + s*g_r (bottom half) with a uniform
    std::vector<u64> e;
    for (size_t half = 0; half < 2; half++)
      for (size_t r = 0; r < l; r++) {
        const uint32_t row_id = (uint32_t)(half * l + r);
        auto &row = out[row_id];
        shared_error(prg::kGswE, row_id, e);
        for (size_t k = 0; k < K; k++) {
          const u64 q = qs_[k];
          u64 *c0 = row.data() + k * N, *c1 = row.data() + K * N + k * N, *ek = e.data() + k * N;
          const u64 *s = sk_.data.data() + k * N;
          prg::uniform(key_, prg::kGswA, row_id, (uint32_t)k, q, N, c1);
          rings_[k]->ntt_fwd(ek);
          for (size_t j = 0; j < N; j++) {
            const u64 as = mmul(c1[j], s[j], q);
            const u64 m = mmul(gad[k][r], half ? mmul(s[j], s[j], q) : s[j], q);
            const u64 neg = (as + ek[j]) % q;
            c0[j] = ((neg ? q - neg : 0) + m) % q;
          }
        }
      }
    return out;
  }
  RlweCt ct;
  for (size_t half = 0; half < 2; half++)
    for (size_t r = 0; r < l; r++) {
      encrypt_zero(ct);
      for (size_t k = 0; k < K; k++) {
        const u64 qk = qs_[k];
        u64 *tgt = ct.data(half) + k * N;
        for (size_t j = 0; j < N; j++) {
          const u64 s = s_coef[j];
          const u64 canon = (K == 1 || s <= q0 / 2) ? s % qk : (q0 - s >= qk ? (qk - (q0 - s) % qk) % qk : qk - (q0 - s));
          tgt[j] = (tgt[j] + mmul(canon, gad[k][r], qk)) % qk;
        }
      }
      for (size_t k = 0; k < K; k++) {
        rings_[k]->ntt_fwd(ct.c0.data() + k * N);
        rings_[k]->ntt_fwd(ct.c1.data() + k * N);
      }
      auto &row = out[half * l + r];
      std::memcpy(row.data(), ct.c0.data(), K * N * 8);
      std::memcpy(row.data() + K * N, ct.c1.data(), K * N * 8);
    }
  return out;
}

SharedPirSessionKeys PirClient::create_session_keys() {
  auto keys = std::make_shared<PirSessionKeys>();
  keys->bv_galois_keys = create_bv_galois_keys();
  keys->gsw_key = generate_gsw_from_key();
  return keys;
}

RlweCt PirClient::load_resp_from_stream(std::stringstream &stream) const {
  // client.cpp:315-361: c0 then c1, ceil(log2 q') bits per value, LSB-first
  const size_t N = params_.get_poly_degree();
  const size_t w = std::bit_width(params_.get_small_q() - 1);
  const size_t need = codec::packed_bytes(2 * N, w);
  std::string buf(need, '\0');
  stream.read(buf.data(), (std::streamsize)need);
  if ((size_t)stream.gcount() != need) throw std::runtime_error("unexpected end of response stream");
  std::vector<u64> v(2 * N);
  size_t pos = 0;
  codec::unpack_bits(buf, pos, 2 * N, w, v.data());
  RlweCt r;
  r.c0.assign(v.begin(), v.begin() + N);
  r.c1.assign(v.begin() + N, v.end());
  return r;
}

RlwePt PirClient::decrypt_mod_q(const RlweCt &ct, int *budget) const {
  // client.cpp:364-430: phase = c0 + c1*s mod q', m = round(phase * t / q')
  const size_t N = params_.get_poly_degree();
  const u64 q = params_.get_small_q(), t = params_.get_plain_mod();
  RingMod small(N, q, utils::minimal_primitive_root(2 * N, q));
  std::vector<u64> s(sk_.data.begin(), sk_.data.begin() + N), c1(N), ph(N);
  rings_[0]->ntt_inv(s.data());
  for (size_t i = 0; i < N; i++) s[i] = s[i] > 1 ? q - 1 : s[i];
  small.ntt_fwd(s.data());
  for (size_t i = 0; i < N; i++) c1[i] = ct.c1[i] % q;
  small.ntt_fwd(c1.data());
  small.mul(c1.data(), s.data(), ph.data());
  small.ntt_inv(ph.data());
  RlwePt pt;
  pt.data.resize(N);
  u64 max_noise = 0;
  for (size_t i = 0; i < N; i++) {
    ph[i] = (ph[i] + ct.c0[i] % q) % q;
    const u64 m = (u64)(((u128)ph[i] * t + (q >> 1)) / q) % t;
    pt.data[i] = m;
    const u64 approx = (u64)(((u128)q * m + (t >> 1)) / t) % q;
    const u64 d = ph[i] >= approx ? ph[i] - approx : q - approx + ph[i];
    const u64 mag = d > q / 2 ? q - d : d;
    if (mag > max_noise) max_noise = mag;
  }
  if (budget)
    *budget = max_noise ? (int)std::log2((double)q / (2.0 * t * max_noise)) : (int)std::log2((double)q / (2.0 * t));
  return pt;
}

}  // namespace onionpir
