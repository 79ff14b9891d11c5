// PirClient on the CPU: secret key, BV Galois keys, RGSW(s), packed query,
// response decoding and small-q decryption.
// Follows src/client.cpp:14-430, src/rlwe.cpp:174-276, src/gsw.cpp:359-422,
// src/bv_keyswitch.cpp:166-276, src/utils.cpp:407-443. Single logical
// modulus schemes (K = 1, composite or prime).
#include <bit>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "onionpir_host.h"

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

// rlwe.cpp:196-215 / bv_keyswitch.cpp:166-185: rounded Gaussian mod q
void gaussian_poly(u64 *out, size_t N, u64 q, double sigma, std::mt19937_64 &rng) {
  std::normal_distribution<double> dist(0.0, sigma);
  for (size_t i = 0; i < N; i++) {
    const int64_t e = std::llround(dist(rng));
    const u64 mag = (u64)(e < 0 ? -e : e) % q;
    out[i] = (e < 0 && mag) ? q - mag : mag;
  }
}

std::shared_ptr<RingMod> ring_for(const PirParams &p) {
  if (p.K() != 1) throw std::invalid_argument("host PirClient supports single-modulus schemes");
  const u64 Q = p.get_rns_mods()[0];
  if (p.q1()) return std::make_shared<RingMod>(p.get_poly_degree(), Q, p.ntt_root(), p.q1(), p.q2());
  return std::make_shared<RingMod>(p.get_poly_degree(), Q, utils::minimal_primitive_root(2 * p.get_poly_degree(), Q));
}

}  // namespace

PirClient::PirClient(const PirParams &params, size_t client_id, u64 seed)
    : client_id_(client_id), params_(params), rng_(seed), seed_(seed), ring_(ring_for(params)) {
  // rlwe.cpp:177-193,219-228: ternary secret {0, 1, -1} from uniform_int(0, 2), NTT form
  const size_t N = params.get_poly_degree();
  const u64 Q = params.get_rns_mods()[0];
  std::uniform_int_distribution<int> dist(0, 2);
  sk_.data.resize(N);
  for (size_t i = 0; i < N; i++) {
    const int v = dist(rng_);
    sk_.data[i] = v == 0 ? 0 : (v == 1 ? 1 : Q - 1);
  }
  ring_->ntt_fwd(sk_.data.data());
}

void PirClient::encrypt_zero(RlweCt &ct) {
  // rlwe.cpp:230-276 (K = 1): c1 = a, c0 = -(a*s + e), coefficient form
  const size_t N = params_.get_poly_degree();
  const u64 Q = params_.get_rns_mods()[0];
  ct.c0.assign(N, 0);
  ct.c1.assign(N, 0);
  uniform_poly(ct.c1.data(), N, Q, rng_);
  std::vector<u64> e(N), a(ct.c1);
  gaussian_poly(e.data(), N, Q, params_.get_noise_std_dev(), rng_);
  ring_->ntt_fwd(a.data());
  ring_->mul(a.data(), sk_.data.data(), ct.c0.data());
  ring_->ntt_inv(ct.c0.data());
  for (size_t i = 0; i < N; i++) {
    const u64 s = (ct.c0[i] + e[i]) % Q;
    ct.c0[i] = s ? Q - s : 0;
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

RlweCt PirClient::fast_generate_query(const PirParams &qp, size_t pt_idx) {
  // client.cpp:107-166 (K = 1)
  if (!params_.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  if (pt_idx >= qp.get_num_pt()) throw std::out_of_range("PIR plaintext index exceeds the runtime layout");
  const u64 Q = qp.get_rns_mods()[0], t = qp.get_plain_mod();
  const auto idx = get_query_indices(qp, pt_idx);
  const size_t h = qp.get_expan_height();
  u64 inv = 0;
  utils::try_invert_uint_mod(size_t{1} << h, t, inv);
  const size_t slot = utils::bit_reverse(idx[0], h);
  RlweCt q;
  encrypt_zero(q);
  const u64 scaled = (u64)(((u128)Q * inv + (t >> 1)) / t) % Q;  // round(Q * inv / t)
  q.c0[slot] = (q.c0[slot] + scaled) % Q;
  add_gsw_to_query(qp, q, idx);
  return q;
}

void PirClient::add_gsw_to_query(const PirParams &qp, RlweCt &query, const std::vector<size_t> &idx) {
  // client.cpp:174-225: selector i with bit 1 adds B^(l-1-k) / 2^h at BitRev(N0 + (i-1)l + k)
  if (!params_.scheme_compatible(qp))
    throw std::invalid_argument("PirClient query layout is not scheme-compatible with its session");
  if (idx.size() == 1) return;
  const u64 Q = qp.get_rns_mods()[0];
  const size_t h = qp.get_expan_height(), l = qp.get_l();
  u64 inv = 0;
  utils::try_invert_uint_mod(size_t{1} << h, Q, inv);
  const auto gad = utils::gsw_gadget(l, qp.get_base_log2(), qp.get_rns_mods());
  for (size_t i = 1; i < idx.size(); i++) {
    if (idx[i] != 1) continue;
    for (size_t k = 0; k < l; k++) {
      const size_t slot = utils::bit_reverse(qp.get_fst_dim_sz() + (i - 1) * l + k, h);
      query.c0[slot] = (query.c0[slot] + mmul(gad[0][k], inv, Q)) % Q;
    }
  }
}

bvks::BvGaloisKeys PirClient::create_bv_galois_keys() {
  // bv_keyswitch.cpp:191-276; key RNG seeded from the client seed (the
  // reference draws that seed from random_device).
  const size_t N = params_.get_poly_degree(), L = params_.get_l_ks(), b = params_.get_bv_base_log2();
  const u64 Q = params_.get_rns_mods()[0];
  std::mt19937_64 rng(splitmix64(seed_));
  bvks::BvGaloisKeys keys;
  std::vector<u64> s_coef(sk_.data), sigma_s(N), e(N);
  ring_->ntt_inv(s_coef.data());
  for (size_t lvl = 0; lvl < params_.get_expan_height(); lvl++) {
    const uint32_t g = (uint32_t)((N >> lvl) + 1);
    // sigma_g(s) in coefficient form, then NTT (utils.cpp:58-82)
    std::fill(sigma_s.begin(), sigma_s.end(), 0);
    for (size_t i = 0; i < N; i++) {
      const size_t dest = (i * g) % (2 * N);
      if (dest < N) sigma_s[dest] = s_coef[i];
      else sigma_s[dest - N] = s_coef[i] ? Q - s_coef[i] : 0;
    }
    ring_->ntt_fwd(sigma_s.data());
    bvks::BvKeySwitchKey key;
    key.galois_k = g;
    key.cts.resize(L);
    for (size_t r = 0; r < L; r++) {
      u64 Br = 1;
      for (size_t k = 0; k < r * b; k++) Br = (u64)(((u128)Br << 1) % Q);
      auto &row = key.cts[r];
      row.a.assign(N, 0);
      row.b.assign(N, 0);
      uniform_poly(row.a.data(), N, Q, rng);
      gaussian_poly(e.data(), N, Q, params_.get_noise_std_dev(), rng);
      ring_->ntt_fwd(row.a.data());
      ring_->ntt_fwd(e.data());
      for (size_t j = 0; j < N; j++) {
        // b = sigma(s) * B^r - a*s + e, NTT form
        const u64 msg = mmul(sigma_s[j], Br, Q), as = mmul(row.a[j], sk_.data[j], Q);
        row.b[j] = ((msg + Q - as) % Q + e[j]) % Q;
      }
    }
    keys.keys.push_back(std::move(key));
  }
  return keys;
}

GSWCt PirClient::generate_gsw_from_key() {
  // client.cpp:25-38 + gsw.cpp:359-422 (K = 1): RGSW(s) with the key gadget
  const size_t N = params_.get_poly_degree(), l = params_.get_l_key();
  const u64 Q = params_.get_rns_mods()[0];
  std::vector<u64> s_coef(sk_.data);
  ring_->ntt_inv(s_coef.data());
  const auto gad = utils::gsw_gadget(l, params_.get_base_log2_key(), params_.get_rns_mods());
  GSWCt out(2 * l, std::vector<u64>(2 * N));
  RlweCt ct;
  for (size_t half = 0; half < 2; half++)
    for (size_t r = 0; r < l; r++) {
      encrypt_zero(ct);
      u64 *tgt = ct.data(half);
      for (size_t j = 0; j < N; j++) tgt[j] = (tgt[j] + mmul(s_coef[j] % Q, gad[0][r], Q)) % Q;
      ring_->ntt_fwd(ct.c0.data());
      ring_->ntt_fwd(ct.c1.data());
      auto &row = out[half * l + r];
      std::memcpy(row.data(), ct.c0.data(), N * 8);
      std::memcpy(row.data() + N, ct.c1.data(), N * 8);
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
  RlweCt r;
  r.c0.assign(N, 0);
  r.c1.assign(N, 0);
  u64 acc = 0;
  size_t have = 0;
  for (size_t p = 0; p < 2; p++)
    for (size_t i = 0; i < N; i++) {
      while (have < w) {
        const int ch = stream.get();
        if (ch == EOF) throw std::runtime_error("unexpected end of response stream");
        acc |= (u64)(uint8_t)ch << have;
        have += 8;
      }
      r.data(p)[i] = acc & ((u64{1} << w) - 1);
      acc >>= w;
      have -= w;
    }
  return r;
}

RlwePt PirClient::decrypt_mod_q(const RlweCt &ct, int *budget) const {
  // client.cpp:364-430: phase = c0 + c1*s mod q', m = round(phase * t / q')
  const size_t N = params_.get_poly_degree();
  const u64 q = params_.get_small_q(), t = params_.get_plain_mod();
  RingMod small(N, q, utils::minimal_primitive_root(2 * N, q));
  std::vector<u64> s(sk_.data), c1(N), ph(N);
  ring_->ntt_inv(s.data());
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
