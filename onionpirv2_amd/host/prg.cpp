// Counter-based sampler (prg.h).
#include "prg.h"

#include <cmath>
#include <cstring>

namespace onionpir {
namespace prg {

namespace {
inline uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
inline void quarter(uint32_t *s, int a, int b, int c, int d) {
  s[a] += s[b]; s[d] ^= s[a]; s[d] = rotl(s[d], 16);
  s[c] += s[d]; s[b] ^= s[c]; s[b] = rotl(s[b], 12);
  s[a] += s[b]; s[d] ^= s[a]; s[d] = rotl(s[d], 8);
  s[c] += s[d]; s[b] ^= s[c]; s[b] = rotl(s[b], 7);
}
inline uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
}  // namespace

void chacha20_block(const Seed &key, uint32_t counter, const uint32_t nonce[3], uint32_t out[16]) {
  uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
  for (int i = 0; i < 8; i++) st[4 + i] = le32(key.data() + 4 * i);
  st[12] = counter;
  st[13] = nonce[0];
  st[14] = nonce[1];
  st[15] = nonce[2];
  uint32_t w[16];
  std::memcpy(w, st, sizeof w);
  for (int r = 0; r < 10; r++) {  // column round, then diagonal round
    quarter(w, 0, 4, 8, 12); quarter(w, 1, 5, 9, 13); quarter(w, 2, 6, 10, 14); quarter(w, 3, 7, 11, 15);
    quarter(w, 0, 5, 10, 15); quarter(w, 1, 6, 11, 12); quarter(w, 2, 7, 8, 13); quarter(w, 3, 4, 9, 14);
  }
  for (int i = 0; i < 16; i++) out[i] = w[i] + st[i];
}

void uniform(const Seed &key, uint32_t domain, uint32_t i, uint32_t j, uint64_t q, size_t n, uint64_t *out) {
  const uint32_t nonce[3] = {domain, i, j};
  uint32_t w[16];
  for (size_t k = 0; k < n; k++) {
    if (k % 4 == 0) chacha20_block(key, (uint32_t)(k / 4), nonce, w);
    const uint32_t *d = w + 4 * (k % 4);
    const unsigned __int128 x = ((unsigned __int128)((uint64_t)d[3] << 32 | d[2]) << 64) | ((uint64_t)d[1] << 32 | d[0]);
    out[k] = (uint64_t)(x % q);
  }
}

void ternary(const Seed &key, size_t n, int8_t *out) {
  const uint32_t nonce[3] = {kSecret, 0, 0};
  uint32_t w[16];
  for (size_t k = 0; k < n; k++) {
    if (k % 8 == 0) chacha20_block(key, (uint32_t)(k / 8), nonce, w);
    const uint64_t u = (uint64_t)w[2 * (k % 8) + 1] << 32 | w[2 * (k % 8)];
    const int r = (int)(u % 3);
    out[k] = (int8_t)(r == 2 ? -1 : r);
  }
}

std::vector<uint64_t> gaussian_cdt(double sigma) {
  // P(sample <= v) = Phi((v + 1/2) / sigma) = erfc(-(v + 1/2) / (sigma sqrt 2)) / 2
  std::vector<uint64_t> t(2 * kCdtMax + 1);
  for (int v = -kCdtMax; v <= kCdtMax; v++) {
    const double p = 0.5 * std::erfc(-(v + 0.5) / (sigma * std::sqrt(2.0)));
    const double scaled = std::ldexp(p, 64);
    t[v + kCdtMax] = scaled >= 18446744073709551615.0 ? ~uint64_t{0} : (uint64_t)scaled;
  }
  t[2 * kCdtMax] = ~uint64_t{0};
  return t;
}

int64_t cdt_sample(const std::vector<uint64_t> &cdt, uint64_t u) {
  int v = 0;
  while (u > cdt[v]) v++;
  return v - kCdtMax;
}

void gaussian(const Seed &key, uint32_t domain, uint32_t i, const std::vector<uint64_t> &cdt, size_t n, int64_t *out) {
  const uint32_t nonce[3] = {domain, i, 0};
  uint32_t w[16];
  for (size_t k = 0; k < n; k++) {
    if (k % 8 == 0) chacha20_block(key, (uint32_t)(k / 8), nonce, w);
    out[k] = cdt_sample(cdt, (uint64_t)w[2 * (k % 8) + 1] << 32 | w[2 * (k % 8)]);
  }
}

Seed query_seed(const Seed &key, uint64_t counter) {
  const uint32_t nonce[3] = {kQuerySeed, (uint32_t)counter, (uint32_t)(counter >> 32)};
  uint32_t w[16];
  chacha20_block(key, 0, nonce, w);
  Seed s;
  std::memcpy(s.data(), w, 32);  // little-endian words
  return s;
}

}  // namespace prg
}  // namespace onionpir
