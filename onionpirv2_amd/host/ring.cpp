// CPU negacyclic NTT for the client side of the host API.
// Same transform as utils::ntt_fwd/ntt_inv (src/utils.cpp:84-90): root powers
// in bit-reversed order, Cooley-Tukey forward into the bit-reversed
// evaluation order, Gentleman-Sande inverse scaled by N^-1. A composite
// modulus Q = q1*q2 is transformed limb by limb (root w mod q_i) and
// recombined by CRT, which is exact because Z_Q = Z_q1 x Z_q2.
#include <bit>
#include <stdexcept>

#include "onionpir_host.h"

namespace onionpir {

using u128 = unsigned __int128;

namespace {
u64 mmul(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
inline u32 shoup(u32 a, u32 w, u32 ws, u32 q) {  // a*w mod q in [0, 2q)
  const u32 hi = (u32)(((u64)a * ws) >> 32);
  return a * w - hi * q;
}
inline u32 red2(u32 x, u32 q) { return x >= q ? x - q : x; }
}  // namespace

RingMod::RingMod(size_t N, u64 q, u64 root, u64 q1, u64 q2) : N_(N), q_(q) {
  const int logn = std::bit_width(N) - 1;
  std::vector<u64> mods;
  split_ = q1 && q2;
  if (split_) mods = {q1, q2};
  else if (q < (u64{1} << 30)) mods = {q};
  if (!mods.empty()) {
    for (u64 m : mods) {
      Limb L;
      L.q = (u32)m;
      const u64 w = root % m;
      u64 wi;
      if (!utils::try_invert_uint_mod(w, m, wi)) throw std::invalid_argument("NTT root not invertible");
      L.w.resize(N); L.ws.resize(N); L.iw.resize(N); L.iws.resize(N);
      u64 p = 1, pi = 1;
      for (size_t i = 0; i < N; i++) {
        const size_t r = utils::bit_reverse(i, logn);
        L.w[r] = (u32)p; L.ws[r] = (u32)(((u64)p << 32) / m);
        L.iw[r] = (u32)pi; L.iws[r] = (u32)(((u64)pi << 32) / m);
        p = mmul(p, w, m);
        pi = mmul(pi, wi, m);
      }
      u64 ni;
      utils::try_invert_uint_mod(N % m, m, ni);
      L.ninv = (u32)ni;
      L.ninvs = (u32)(((u64)ni << 32) / m);
      limbs_.push_back(std::move(L));
    }
    if (split_) utils::try_invert_uint_mod(q1 % q2, q2, q1inv_mod_q2_);
    return;
  }
  // direct 64-bit path
  u64 wi;
  if (!utils::try_invert_uint_mod(root % q, q, wi)) throw std::invalid_argument("NTT root not invertible");
  w64_.resize(N);
  iw64_.resize(N);
  u64 p = 1, pi = 1;
  for (size_t i = 0; i < N; i++) {
    const size_t r = utils::bit_reverse(i, logn);
    w64_[r] = p;
    iw64_[r] = pi;
    p = mmul(p, root, q);
    pi = mmul(pi, wi, q);
  }
  utils::try_invert_uint_mod(N % q, q, ninv64_);
}

void RingMod::limb_fwd(const Limb &L, u32 *a) const {
  const u32 q = L.q, q2 = 2 * q;
  for (size_t m = 1, t = N_ / 2; m < N_; m <<= 1, t >>= 1)
    for (size_t i = 0; i < m; i++) {
      const u32 w = L.w[m + i], ws = L.ws[m + i];
      for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
        const u32 x = a[j] >= q2 ? a[j] - q2 : a[j];
        const u32 y = shoup(a[j + t], w, ws, q);
        a[j] = x + y;
        a[j + t] = x - y + q2;
      }
    }
  for (size_t j = 0; j < N_; j++) a[j] = red2(red2(a[j], q2), q);
}

void RingMod::limb_inv(const Limb &L, u32 *a) const {
  const u32 q = L.q, q2 = 2 * q;
  for (size_t m = N_ / 2, t = 1; m >= 1; m >>= 1, t <<= 1)
    for (size_t i = 0; i < m; i++) {
      const u32 w = L.iw[m + i], ws = L.iws[m + i];
      for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
        const u32 x = a[j], y = a[j + t];
        a[j] = red2(x + y, q2);
        a[j + t] = shoup(x - y + q2, w, ws, q);
      }
    }
  for (size_t j = 0; j < N_; j++) a[j] = red2(shoup(a[j], L.ninv, L.ninvs, q), q);
}

void RingMod::ntt_fwd(u64 *a) const {
  if (!limbs_.empty()) {
    std::vector<u32> x(N_), y;
    for (size_t j = 0; j < N_; j++) x[j] = (u32)(a[j] % limbs_[0].q);
    limb_fwd(limbs_[0], x.data());
    if (!split_) {
      for (size_t j = 0; j < N_; j++) a[j] = x[j];
      return;
    }
    y.resize(N_);
    for (size_t j = 0; j < N_; j++) y[j] = (u32)(a[j] % limbs_[1].q);
    limb_fwd(limbs_[1], y.data());
    const u64 q1 = limbs_[0].q, q2 = limbs_[1].q;
    for (size_t j = 0; j < N_; j++) {
      const u64 k = mmul((y[j] + q2 - x[j] % q2) % q2, q1inv_mod_q2_, q2);
      a[j] = x[j] + q1 * k;
    }
    return;
  }
  for (size_t m = 1, t = N_ / 2; m < N_; m <<= 1, t >>= 1)
    for (size_t i = 0; i < m; i++)
      for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
        const u64 x = a[j], y = mmul(a[j + t], w64_[m + i], q_);
        a[j] = (x + y) % q_;
        a[j + t] = (x + q_ - y) % q_;
      }
}

void RingMod::ntt_inv(u64 *a) const {
  if (!limbs_.empty()) {
    std::vector<u32> x(N_), y;
    for (size_t j = 0; j < N_; j++) x[j] = (u32)(a[j] % limbs_[0].q);
    limb_inv(limbs_[0], x.data());
    if (!split_) {
      for (size_t j = 0; j < N_; j++) a[j] = x[j];
      return;
    }
    y.resize(N_);
    for (size_t j = 0; j < N_; j++) y[j] = (u32)(a[j] % limbs_[1].q);
    limb_inv(limbs_[1], y.data());
    const u64 q1 = limbs_[0].q, q2 = limbs_[1].q;
    for (size_t j = 0; j < N_; j++) {
      const u64 k = mmul((y[j] + q2 - x[j] % q2) % q2, q1inv_mod_q2_, q2);
      a[j] = x[j] + q1 * k;
    }
    return;
  }
  for (size_t m = N_ / 2, t = 1; m >= 1; m >>= 1, t <<= 1)
    for (size_t i = 0; i < m; i++)
      for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
        const u64 x = a[j], y = a[j + t];
        a[j] = (x + y) % q_;
        a[j + t] = mmul((x + q_ - y) % q_, iw64_[m + i], q_);
      }
  for (size_t j = 0; j < N_; j++) a[j] = mmul(a[j], ninv64_, q_);
}

void RingMod::mul(const u64 *a, const u64 *b, u64 *out) const {
  for (size_t j = 0; j < N_; j++) out[j] = mmul(a[j], b[j], q_);
}

}  // namespace onionpir
