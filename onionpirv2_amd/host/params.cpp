// Scheme parameters and number theory for the host API.
// Follows src/pir.cpp:14-194 and src/utils.cpp:96-401.
#include <bit>
#include <stdexcept>

#include "onionpir_host.h"

namespace onionpir {

using u128 = unsigned __int128;

static u64 mmul(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
static u64 mpow(u64 b, u64 e, u64 m) {
  u64 r = 1 % m;
  for (b %= m; e; e >>= 1, b = mmul(b, b, m))
    if (e & 1) r = mmul(r, b, m);
  return r;
}

SchemeConfig SchemeConfig::by_name(const std::string &name) {
  SchemeConfig c;  // defaults are CONFIG_N2048_K1_COMP (database_constants.h:77-94)
  if (name == "CONFIG_N2048_K1_COMP") return c;
  c.name = name;
  c.composite_first_dim = false;
  c.first_dim_rns_bits.clear();
  if (name == "CONFIG_N2048_K1") {  // database_constants.h:61-75
    c.l_ep = 5; c.l_key = 8; c.plain_mod_bits = 14; c.rns_mod_bits = {60};
  } else if (name == "CONFIG_N2048_K2_MP") {  // database_constants.h:96-109
    c.l_ep = 5; c.l_key = 8; c.plain_mod_bits = 10; c.rns_mod_bits = {29, 29};
  } else if (name == "CONFIG_N4096_K2_MP") {  // database_constants.h:111-124
    c.poly_degree = 4096; c.l_ep = 5; c.l_key = 8; c.plain_mod_bits = 40; c.small_q_width = 50;
    c.rns_mod_bits = {60, 60};
  } else {
    throw std::invalid_argument("unknown scheme config " + name);
  }
  return c;
}

namespace utils {

bool is_prime(u64 n) {
  // deterministic Miller-Rabin over the first twelve primes (utils.cpp:126-153)
  static const u64 bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2) return false;
  for (u64 p : bases)
    if (n % p == 0) return n == p;
  int r = 0;
  u64 d = n - 1;
  for (; !(d & 1); d >>= 1) r++;
  for (u64 a : bases) {
    u64 x = mpow(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool witness = true;
    for (int i = 1; i < r && witness; i++) {
      x = mmul(x, x, n);
      if (x == n - 1) witness = false;
    }
    if (witness) return false;
  }
  return true;
}

u64 generate_prime(size_t bw) {  // utils.cpp:240-251
  if (bw < 2) throw std::invalid_argument("Bit width must be at least 2.");
  u64 c = u64{1} << (bw - 1);
  do c = (c + 1) | 1; while (!is_prime(c));
  return c;
}

std::vector<u64> generate_ntt_friendly_primes(const std::vector<size_t> &bws, size_t N) {  // utils.cpp:253-287
  const u64 step = 2 * (u64)N;
  std::map<size_t, u64> resume;
  std::vector<u64> out;
  for (size_t bw : bws) {
    if (bw < 2 || bw > 63) throw std::invalid_argument("bit width out of range [2, 63]");
    u64 c;
    if (resume.count(bw)) {
      c = resume[bw];
    } else {
      c = (u64{1} << bw) - 1;
      c -= (c - 1) % step;
    }
    while (!is_prime(c)) {
      if (c < step + 2) throw std::runtime_error("No NTT-friendly prime found");
      c -= step;
    }
    out.push_back(c);
    resume[bw] = c >= step ? c - step : 0;
  }
  return out;
}

u64 minimal_primitive_root(u64 degree, u64 q) {
  // smallest x with x^(degree/2) = -1 mod prime q (the definition HEXL's
  // MinimalPrimitiveRoot computes, used at pir.cpp:32-33)
  u64 g = 0;
  for (u64 c = 2; c < q && !g; c++) {
    const u64 x = mpow(c, (q - 1) / degree, q);
    if (mpow(x, degree / 2, q) == q - 1) g = x;
  }
  if (!g) throw std::runtime_error("no primitive root of unity");
  u64 best = g, x = g;
  const u64 g2 = mmul(g, g, q);
  for (u64 k = 1; k < degree / 2; k++) {
    x = mmul(x, g2, q);
    if (x < best) best = x;
  }
  return best;
}

bool try_invert_uint_mod(u64 value, u64 modulus, u64 &result) {
  if (value == 0) return false;
  __int128 r0 = modulus, r1 = value % modulus, s0 = 0, s1 = 1;
  while (r1) {
    const __int128 q = r0 / r1, r2 = r0 - q * r1, s2 = s0 - q * s1;
    r0 = r1; r1 = r2; s0 = s1; s1 = s2;
  }
  if (r0 != 1) return false;
  result = (u64)((s0 % (__int128)modulus + modulus) % modulus);
  return true;
}

u64 crt_combine(u64 w1, u64 q1, u64 w2, u64 q2) {  // utils.cpp:229-238
  u64 inv;
  if (!try_invert_uint_mod(q1 % q2, q2, inv)) throw std::invalid_argument("crt_combine: q1, q2 must be coprime");
  const u64 k = mmul((w2 + q2 - w1 % q2) % q2, inv, q2);
  return (u64)((u128)q1 * k + w1);
}

std::pair<size_t, size_t> calculate_db_shape(size_t target, size_t l, size_t h, bool pow2) {  // utils.cpp:376-401
  if (target == 0) throw std::invalid_argument("calculate_db_shape target must be positive");
  if (l == 0) throw std::invalid_argument("calculate_db_shape gadget length must be positive");
  if (h >= 64) throw std::invalid_argument("calculate_db_shape height is too large");
  const size_t cap = size_t{1} << h;
  for (size_t dims = 1; l * (dims - 1) < cap; dims++) {
    const size_t slack = cap - l * (dims - 1);
    const size_t fst = pow2 ? std::bit_floor(slack) : slack;
    if (fst * (size_t{1} << (dims - 1)) >= target) return {fst, dims};
  }
  throw std::runtime_error("Failed to calculate database shape");
}

size_t bit_reverse(size_t x, size_t logn) {
  size_t y = 0;
  for (size_t i = 0; i < logn; i++, x >>= 1) y = (y << 1) | (x & 1);
  return y;
}

std::vector<std::vector<u64>> gsw_gadget(size_t l, u64 b, const std::vector<u64> &mods) {  // utils.cpp:206-222
  std::vector<std::vector<u64>> g(mods.size(), std::vector<u64>(l));
  for (size_t k = 0; k < mods.size(); k++) {
    u64 pw = 1;
    for (size_t r = l; r-- > 0;) {
      g[k][r] = pw;
      pw = (u64)(((u128)pw << b) % mods[k]);
    }
  }
  return g;
}

}  // namespace utils

PirParams::PirParams(const SchemeConfig &cfg) : cfg_(cfg) {
  const size_t N = cfg.poly_degree;
  if (cfg.composite_first_dim) {  // pir.cpp:14-47
    const auto pr = utils::generate_ntt_friendly_primes(cfg.first_dim_rns_bits, N);
    q1_ = pr[0];
    q2_ = pr[1];
    w_crt_ = utils::crt_combine(utils::minimal_primitive_root(2 * N, q1_), q1_,
                                utils::minimal_primitive_root(2 * N, q2_), q2_);
    rns_mods_ = {q1_ * q2_};
  } else {
    rns_mods_ = utils::generate_ntt_friendly_primes(cfg.rns_mod_bits, N);
  }
  plain_mod_ = utils::generate_prime(cfg.plain_mod_bits);
  small_q_ = utils::generate_ntt_friendly_primes({cfg.small_q_width}, N)[0];
  const size_t w = get_ct_mod_width();
  base_log2_ = (w + cfg.l_ep - 1) / cfg.l_ep;
  base_log2_key_ = (w + cfg.l_key - 1) / cfg.l_key;
  apply_layout({cfg.db_size_mb * 1024 * 1024 / get_pt_size(), cfg.tree_height, cfg.fst_dim_pow2});
}

size_t PirParams::get_ct_mod_width() const {
  size_t w = 0;
  for (size_t b : cfg_.rns_mod_bits) w += b;
  return w;
}

void PirParams::apply_layout(const PirLayoutConfig &lay) {  // pir.cpp:100-128
  if (lay.target_num_pt == 0) throw std::invalid_argument("PirParams: target_num_pt must be positive");
  if (lay.expansion_height > (size_t)std::bit_width(cfg_.poly_degree) - 1)
    throw std::invalid_argument("PirParams: expansion_height exceeds the ring automorphism capacity");
  const auto [fst, dims] = utils::calculate_db_shape(lay.target_num_pt, cfg_.l_ep, lay.expansion_height,
                                                     lay.fst_dim_pow2);
  target_num_pt_ = lay.target_num_pt;
  expansion_height_ = lay.expansion_height;
  fst_dim_pow2_ = lay.fst_dim_pow2;
  fst_dim_sz_ = fst;
  num_dims_ = dims;
  num_pt_ = fst * ((lay.target_num_pt + fst - 1) / fst);
}

PirParams PirParams::with_layout(const PirLayoutConfig &layout) const {
  PirParams r(*this);
  r.apply_layout(layout);
  return r;
}

PirParams PirParams::with_query_shape(size_t fst, size_t nsel, size_t h) const {  // pir.cpp:136-167
  if (h > (size_t)std::bit_width(cfg_.poly_degree) - 1)
    throw std::invalid_argument("PirParams: query shape exceeds the ring automorphism capacity");
  if (fst == 0) throw std::invalid_argument("PirParams: query shape needs a positive first dimension");
  const size_t cap = size_t{1} << h;
  if (fst > cap || nsel > (cap - fst) / cfg_.l_ep)
    throw std::invalid_argument("PirParams: query shape does not fit the expansion capacity");
  PirParams r(*this);
  r.fst_dim_sz_ = fst;
  r.num_dims_ = 1 + nsel;
  r.expansion_height_ = h;
  r.num_pt_ = r.target_num_pt_ = fst;
  return r;
}

bool PirParams::scheme_compatible(const PirParams &o) const {  // pir.cpp:169-186
  return rns_mods_ == o.rns_mods_ && q1_ == o.q1_ && q2_ == o.q2_ && w_crt_ == o.w_crt_ &&
         plain_mod_ == o.plain_mod_ && small_q_ == o.small_q_ && base_log2_ == o.base_log2_ &&
         base_log2_key_ == o.base_log2_key_ && cfg_.rns_mod_bits == o.cfg_.rns_mod_bits &&
         cfg_.composite_first_dim == o.cfg_.composite_first_dim && cfg_.poly_degree == o.cfg_.poly_degree;
}

const bvks::BvKeySwitchKey &bvks::BvGaloisKeys::get(uint32_t galois_k) const {  // bv_keyswitch.cpp:92-98
  for (auto &k : keys)
    if (k.galois_k == galois_k) return k;
  throw std::out_of_range("BvGaloisKeys::get: galois_k not found");
}

}  // namespace onionpir
