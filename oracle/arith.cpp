// OnionPIRv2 CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Number theory, Barrett, NTT, automorphisms and runtime parameters.
#include <bit>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "oracle.h"

namespace oracle {

SchemeConfig config_by_name(const std::string &name) {
  // src/includes/database_constants.h:47-124
  SchemeConfig c;
  c.name = name;
  if (name == "CONFIG_N2048_K1_COMP") {
    c.poly_degree = 2048; c.l_ep = 6; c.l_key = 10; c.l_ks = 8; c.tree_height = 10;
    c.plain_mod_bits = 13; c.small_q_width = 22; c.rns_mod_bits = {58};
    c.composite_first_dim = true; c.first_dim_rns_bits = {29, 29};
  } else if (name == "CONFIG_N2048_K1") {
    c.poly_degree = 2048; c.l_ep = 5; c.l_key = 8; c.l_ks = 8; c.tree_height = 10;
    c.plain_mod_bits = 14; c.small_q_width = 22; c.rns_mod_bits = {60};
    c.composite_first_dim = false;
  } else if (name == "CONFIG_N2048_K2_MP") {
    c.poly_degree = 2048; c.l_ep = 5; c.l_key = 8; c.l_ks = 8; c.tree_height = 10;
    c.plain_mod_bits = 10; c.small_q_width = 22; c.rns_mod_bits = {29, 29};
    c.composite_first_dim = false;
  } else if (name == "CONFIG_N4096_K2_MP") {
    c.poly_degree = 4096; c.l_ep = 5; c.l_key = 8; c.l_ks = 8; c.tree_height = 10;
    c.plain_mod_bits = 40; c.small_q_width = 50; c.rns_mod_bits = {60, 60};
    c.composite_first_dim = false;
  } else {
    throw std::invalid_argument("unknown config " + name);
  }
  return c;
}

u64 mulmod(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }

u64 powmod(u64 b, u64 e, u64 m) {
  u64 r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod(r, b, m);
    b = mulmod(b, b, m);
    e >>= 1;
  }
  return r;
}

// utils.cpp:126-153: deterministic Miller-Rabin with the first 12 primes.
static bool mr_witness(u64 n, u64 d, int r, u64 a) {
  u64 x = powmod(a, d, n);
  if (x == 1 || x == n - 1) return false;
  for (int i = 0; i < r - 1; ++i) {
    x = mulmod(x, x, n);
    if (x == n - 1) return false;
  }
  return true;
}

bool is_prime(u64 n) {
  if (n < 2) return false;
  static const u64 sp[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (u64 p : sp) {
    if (n == p) return true;
    if (n % p == 0) return false;
  }
  u64 d = n - 1;
  int r = 0;
  while ((d & 1) == 0) { d >>= 1; ++r; }
  for (u64 a : sp)
    if (mr_witness(n, d, r, a)) return false;
  return true;
}

// utils.cpp:96-106,155-163: extended Euclid on signed 64-bit values.
bool try_invert_uint_mod(u64 value, u64 modulus, u64 &out) {
  if (value == 0) return false;
  int64_t s0 = 1, s1 = 0, r0 = (int64_t)value, r1 = (int64_t)modulus;
  while (r1 != 0) {
    int64_t qq = r0 / r1;
    int64_t t = r1; r1 = r0 - qq * r1; r0 = t;
    t = s1; s1 = s0 - qq * s1; s0 = t;
  }
  if (r0 != 1) return false;
  out = (s0 < 0) ? (u64)(s0 + (int64_t)modulus) : (u64)s0;
  return true;
}

// utils.cpp:240-251: smallest odd prime above 2^(bw-1), stepping c = (c+1)|1.
u64 generate_prime(size_t bw) {
  if (bw < 2) throw std::invalid_argument("Bit width must be at least 2.");
  u64 c = 1ULL << (bw - 1);
  do {
    c++;
    c |= 1;
  } while (!is_prime(c));
  return c;
}

// utils.cpp:253-287: largest primes p < 2^bw with p = 1 mod 2N, distinct per width.
std::vector<u64> generate_ntt_friendly_primes(const std::vector<size_t> &bws, size_t N) {
  const u64 step = 2 * (u64)N;
  std::unordered_map<size_t, u64> next;
  std::vector<u64> out;
  for (size_t bw : bws) {
    if (bw < 2 || bw > 63) throw std::invalid_argument("bit width out of range [2, 63]");
    u64 c;
    auto it = next.find(bw);
    if (it == next.end()) {
      c = (1ULL << bw) - 1;
      c -= (c - 1) % step;
    } else {
      c = it->second;
    }
    while (true) {
      if (c < 2) throw std::runtime_error("No NTT-friendly prime found");
      if (is_prime(c)) break;
      c -= step;
    }
    out.push_back(c);
    next[bw] = (c >= step) ? c - step : 0;
  }
  return out;
}

// utils.cpp:229-238: Garner combine.
u64 crt_combine(u64 w1, u64 q1, u64 w2, u64 q2) {
  u64 inv;
  if (!try_invert_uint_mod(q1 % q2, q2, inv)) throw std::invalid_argument("crt_combine: not coprime");
  const u64 diff = (w2 + q2 - (w1 % q2)) % q2;
  const u64 k = mulmod(diff, inv, q2);
  return (u64)((u128)q1 * k + w1);
}

// HEXL MinimalPrimitiveRoot(degree, q) (pir.cpp:32-33): the smallest primitive
// degree-th root of unity mod prime q, i.e. min over the odd powers of any
// primitive root. Computed from the definition: find x with x^(degree/2) = -1.
u64 minimal_primitive_root(u64 degree, u64 q) {
  if ((q - 1) % degree != 0) throw std::invalid_argument("q != 1 mod degree");
  u64 x = 0;
  for (u64 c = 2; c < q; ++c) {
    u64 cand = powmod(c, (q - 1) / degree, q);
    if (powmod(cand, degree / 2, q) == q - 1) { x = cand; break; }
  }
  if (!x) throw std::runtime_error("no primitive root");
  const u64 x2 = mulmod(x, x, q);
  u64 cur = x, best = x;
  for (u64 i = 0; i < degree / 2; ++i) {
    if (cur < best) best = cur;
    cur = mulmod(cur, x2, q);
  }
  return best;
}

// utils.cpp:376-401
std::pair<size_t, size_t> calculate_db_shape(size_t target, size_t l, size_t h, bool pow2) {
  if (target == 0) throw std::invalid_argument("calculate_db_shape target must be positive");
  if (l == 0) throw std::invalid_argument("calculate_db_shape gadget length must be positive");
  if (h >= 64) throw std::invalid_argument("calculate_db_shape height is too large");
  const size_t cap = size_t{1} << h;
  const size_t max_dims = 1 + (cap - 1) / l;
  for (size_t d = 1; d <= max_dims + 1; d++) {
    const size_t reserved = l * (d - 1);
    if (reserved >= cap) break;
    const size_t slack = cap - reserved;
    const size_t fst = pow2 ? (size_t{1} << (std::bit_width(slack) - 1)) : slack;
    if (fst * (size_t{1} << (d - 1)) >= target) return {fst, d};
  }
  throw std::runtime_error("Failed to calculate database shape");
}

size_t bit_reverse(size_t x, size_t logn) {
  size_t y = 0;
  for (size_t i = 0; i < logn; i++) { y = (y << 1) | (x & 1); x >>= 1; }
  return y;
}

// utils.cpp:445-462: centered round(a * out / inp) reduced into [0, out).
u64 rescale(u64 a, u64 inp_mod, u64 out_mod) {
  const int64_t inp_i = (int64_t)inp_mod;
  const i128 out_i = (i128)out_mod;
  int64_t v = (int64_t)(a % inp_mod);
  if (v >= inp_i / 2) v -= inp_i;
  const int64_t sign = (v >= 0) ? 1 : -1;
  const i128 val = (i128)v * (i128)out_mod;
  i128 r = (val + (i128)(sign * (inp_i / 2))) / (i128)inp_mod;
  r = (r + (i128)((inp_mod / out_mod) * out_mod) + 2 * out_i) % out_i;
  return (u64)((r + out_i) % out_i);
}

// utils.h:131-203
BarrettU64 barrett_u64_setup(u64 q) {
  return {q, q == 1 ? 0 : (u64)(((u128)1 << 64) / q)};
}
u64 barrett_reduce_u64(u64 x, const BarrettU64 &b) {
  if (b.q == 1) return 0;
  const u64 qe = (u64)(((u128)x * b.mu) >> 64);
  u64 r = x - qe * b.q;
  if (r >= b.q) r -= b.q;
  return r;
}
BarrettU128 barrett_u128_setup(u64 q) {
  const u128 mu = (~(u128)0) / q;
  return {q, (u64)mu, (u64)(mu >> 64)};
}
u64 barrett_reduce_u128(u128 x, const BarrettU128 &b) {
  const u64 xl = (u64)x, xh = (u64)(x >> 64);
  const u64 c1 = (u64)(((u128)xl * b.mu_lo) >> 64);
  const u128 t1 = (u128)xl * b.mu_hi + c1;
  const u64 t1l = (u64)t1, t1h = (u64)(t1 >> 64);
  const u128 t2 = (u128)xh * b.mu_lo + t1l;
  const u64 c2 = (u64)(t2 >> 64);
  const u64 qe = xh * b.mu_hi + t1h + c2;
  u64 r = xl - qe * b.q;
  if (r >= b.q) r -= b.q;
  return r;
}

// ---------------------------------------------------------------------------
// NTT (utils.cpp:13-56,84-90 via HEXL).
// ---------------------------------------------------------------------------
namespace {
struct RootKey {
  size_t N; u64 q;
  bool operator<(const RootKey &o) const { return N != o.N ? N < o.N : q < o.q; }
};
std::map<RootKey, u64> &root_registry() { static std::map<RootKey, u64> m; return m; }
std::map<RootKey, std::unique_ptr<NttTable>> &table_cache() {
  static std::map<RootKey, std::unique_ptr<NttTable>> m; return m;
}
}  // namespace

void register_ntt_root(size_t N, u64 q, u64 root) {
  auto &m = root_registry();
  auto it = m.find({N, q});
  if (it != m.end() && it->second != root)
    throw std::invalid_argument("register_ntt_root: a different root is already registered");
  m[{N, q}] = root;
}

const NttTable &ntt_table(size_t N, u64 q) {
  auto &cache = table_cache();
  auto it = cache.find({N, q});
  if (it != cache.end()) return *it->second;
  auto t = std::make_unique<NttTable>();
  t->N = N; t->q = q;
  auto rit = root_registry().find({N, q});
  t->root = (rit != root_registry().end()) ? rit->second : minimal_primitive_root(2 * N, q);
  const size_t logn = std::bit_width(N) - 1;
  u64 root_inv;
  if (!try_invert_uint_mod(t->root % q, q, root_inv)) throw std::runtime_error("root not invertible");
  t->psi_brv.assign(N, 0);
  t->psi_inv_brv.assign(N, 0);
  u64 p = 1, pi = 1;
  for (size_t i = 0; i < N; i++) {
    t->psi_brv[bit_reverse(i, logn)] = p;
    t->psi_inv_brv[bit_reverse(i, logn)] = pi;
    p = mulmod(p, t->root, q);
    pi = mulmod(pi, root_inv, q);
  }
  if (!try_invert_uint_mod(N % q, q, t->inv_n)) throw std::runtime_error("N not invertible");
  auto &ref = *t;
  cache[{N, q}] = std::move(t);
  return ref;
}

// Forward: Cooley-Tukey, natural order in, bit-reversed evaluation order out:
// out[j] = sum_i a_i * root^((2*bitrev(j)+1)*i) mod q.
void ntt_fwd(u64 *a, size_t N, u64 q) {
  const NttTable &T = ntt_table(N, q);
  for (size_t m = 1, t = N / 2; m < N; m <<= 1, t >>= 1) {
    for (size_t i = 0; i < m; i++) {
      const u64 W = T.psi_brv[m + i];
      for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
        const u64 X = a[j];
        const u64 Y = mulmod(a[j + t], W, q);
        a[j] = (X + Y) % q;
        a[j + t] = (X + q - Y) % q;
      }
    }
  }
}

// Inverse: Gentleman-Sande, exact inverse of ntt_fwd, scaled by N^-1.
void ntt_inv(u64 *a, size_t N, u64 q) {
  const NttTable &T = ntt_table(N, q);
  for (size_t m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1) {
    for (size_t i = 0; i < m; i++) {
      const u64 W = T.psi_inv_brv[m + i];
      for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
        const u64 X = a[j], Y = a[j + t];
        a[j] = (X + Y) % q;
        a[j + t] = mulmod((X + q - Y) % q, W, q);
      }
    }
  }
  for (size_t j = 0; j < N; j++) a[j] = mulmod(a[j], T.inv_n, q);
}

// utils.cpp:58-72
void automorphism_coeff(const u64 *in, size_t N, uint32_t k, u64 q, u64 *out) {
  const size_t two_n = 2 * N;
  for (size_t i = 0; i < N; i++) {
    const size_t dest = ((u64)i * k) % two_n;
    if (dest < N) out[dest] = in[i];
    else out[dest - N] = (in[i] == 0) ? 0 : q - in[i];
  }
}

// utils.cpp:74-82
void automorphism_ntt(const u64 *in, size_t N, uint32_t k, u64 q, u64 *out) {
  std::vector<u64> coeff(in, in + N);
  ntt_inv(coeff.data(), N, q);
  automorphism_coeff(coeff.data(), N, k, q, out);
  ntt_fwd(out, N, q);
}

// utils.cpp:166-186: multiply by X^shift (shift taken mod 2N via size_t wrap).
void negacyclic_shift(const u64 *poly, size_t N, size_t shift, u64 q, u64 *out) {
  if (shift == 0) { std::memcpy(out, poly, N * sizeof(u64)); return; }
  u64 idx_raw = shift;
  const u64 mask = (u64)N - 1;
  for (size_t i = 0; i < N; i++, idx_raw++) {
    const u64 idx = idx_raw & mask;
    if (!(idx_raw & (u64)N) || !poly[i]) out[idx] = poly[i];
    else out[idx] = q - poly[i];
  }
}

// utils.cpp:206-222: MSB-first gadget, row p = B^(l-1-p) mod q.
std::vector<std::vector<u64>> gsw_gadget(size_t l, u64 base_log2, const std::vector<u64> &mods) {
  std::vector<std::vector<u64>> g(mods.size(), std::vector<u64>(l));
  for (size_t i = 0; i < mods.size(); i++) {
    u64 pw = 1;
    for (int j = (int)l - 1; j >= 0; j--) {
      g[i][j] = pw;
      pw = (u64)(((u128)pw << base_log2) % mods[i]);
    }
  }
  return g;
}

size_t small_q_width(u64 small_q) { return std::bit_width(small_q - 1); }

// ---------------------------------------------------------------------------
// PirParams. src/pir.cpp:14-194
// ---------------------------------------------------------------------------
PirParams::PirParams(const SchemeConfig &c) : cfg(c), rns_mod_bits(c.rns_mod_bits) {
  const size_t N = c.poly_degree;
  if (c.composite_first_dim) {
    // pir.cpp:14-47
    auto pr = generate_ntt_friendly_primes(c.first_dim_rns_bits, N);
    q1 = pr[0]; q2 = pr[1];
    const u64 Q = q1 * q2;
    w1 = minimal_primitive_root(2 * N, q1);
    w2 = minimal_primitive_root(2 * N, q2);
    w_crt = crt_combine(w1, q1, w2, q2);
    register_ntt_root(N, Q, w_crt);
    rns_mods = {Q};
    composite = true;
    if (!try_invert_uint_mod(q1 % q2, q2, q1_inv_mod_q2)) throw std::runtime_error("q1,q2 not coprime");
  } else {
    rns_mods = generate_ntt_friendly_primes(rns_mod_bits, N);
  }
  plain_mod = generate_prime(c.plain_mod_bits);                           // pir.cpp:59
  small_q = generate_ntt_friendly_primes({c.small_q_width}, N)[0];        // pir.cpp:62-63
  r64_mod_q.resize(rns_mods.size());
  for (size_t i = 0; i < rns_mods.size(); i++) r64_mod_q[i] = (u64)(((u128)1 << 64) % rns_mods[i]);
  if (rns_mods.size() == 2 && !try_invert_uint_mod(rns_mods[0] % rns_mods[1], rns_mods[1], q0_inv_mod_q1))
    throw std::runtime_error("PirParams: coeff moduli not coprime");
  const size_t w = ct_mod_width();
  base_log2 = (w + c.l_ep - 1) / c.l_ep;                                  // pir.cpp:82
  base_log2_key = (w + c.l_key - 1) / c.l_key;                            // pir.cpp:83
  const size_t target = c.db_size_mb * 1024 * 1024 / pt_size();           // pir.cpp:92-95
  apply_layout(target, c.tree_height, c.fst_dim_pow2);
}

size_t PirParams::ct_mod_width() const {
  size_t w = 0;
  for (size_t b : rns_mod_bits) w += b;
  return w;
}

void PirParams::apply_layout(size_t target, size_t h, bool pow2) {  // pir.cpp:100-128
  if (target == 0) throw std::invalid_argument("PirParams: target_num_pt must be positive");
  const size_t max_h = std::bit_width(N()) - 1;
  if (h > max_h) throw std::invalid_argument("PirParams: expansion_height exceeds the ring automorphism capacity");
  auto [fst, dims] = calculate_db_shape(target, cfg.l_ep, h, pow2);
  const size_t other = (target + fst - 1) / fst;
  target_num_pt = target;
  expansion_height = h;
  fst_dim_pow2 = pow2;
  fst_dim_sz = fst;
  num_dims = dims;
  num_pt = fst * other;
}

PirParams PirParams::with_layout(size_t target, size_t h, bool pow2) const {  // pir.cpp:130-134
  PirParams r(*this);
  r.apply_layout(target, h, pow2);
  return r;
}

PirParams PirParams::with_query_shape(size_t fst, size_t nsel, size_t h) const {  // pir.cpp:136-167
  const size_t max_h = std::bit_width(N()) - 1;
  if (h > max_h) throw std::invalid_argument("PirParams: query shape exceeds the ring automorphism capacity");
  if (fst == 0) throw std::invalid_argument("PirParams: query shape needs a positive first dimension");
  const size_t cap = size_t{1} << h;
  if (fst > cap || nsel > (cap - fst) / cfg.l_ep)
    throw std::invalid_argument("PirParams: query shape does not fit the expansion capacity");
  PirParams r(*this);
  r.fst_dim_sz = fst; r.num_dims = 1 + nsel; r.expansion_height = h;
  r.num_pt = fst; r.target_num_pt = fst;
  return r;
}

bool PirParams::scheme_compatible(const PirParams &o) const {  // pir.cpp:169-186
  return composite == o.composite && q1 == o.q1 && q2 == o.q2 && w1 == o.w1 && w2 == o.w2 &&
         w_crt == o.w_crt && q1_inv_mod_q2 == o.q1_inv_mod_q2 && small_q == o.small_q &&
         base_log2 == o.base_log2 && base_log2_key == o.base_log2_key && plain_mod == o.plain_mod &&
         rns_mod_bits == o.rns_mod_bits && rns_mods == o.rns_mods && q0_inv_mod_q1 == o.q0_inv_mod_q1 &&
         r64_mod_q == o.r64_mod_q;
}

}  // namespace oracle
