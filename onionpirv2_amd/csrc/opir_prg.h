// Device copy of the counter-based sampler of the host client
// (onionpirv2_amd/host/prg.h states the rules; both must draw the same
// values): ChaCha20 blocks (RFC 8439 section 2.3) keyed by a 32-byte seed,
// nonce = (domain, i, j), block counter = block index; 128-bit draws for
// uniform residues, 64-bit draws for the secret and the errors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace opir {
namespace prg {

enum : uint32_t { kSecret = 1, kQueryA = 2, kQueryE = 3, kBvA = 4, kBvE = 5, kGswA = 6, kGswE = 7, kQuerySeed = 8 };
constexpr int kCdtMax = 24;

struct Key {
  uint32_t w[8];  // the seed as eight little-endian words
};

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

__device__ __forceinline__ void chacha20_block(const Key &k, uint32_t counter, uint32_t n0, uint32_t n1, uint32_t n2,
                                               uint32_t out[16]) {
  uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, k.w[0], k.w[1], k.w[2], k.w[3],
                     k.w[4],      k.w[5],      k.w[6],      k.w[7],      counter, n0,     n1,     n2};
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 16; i++) x[i] = st[i];
#define OPIR_QR(a, b, c, d)                                        \
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16);                    \
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12);                    \
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);                     \
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
#pragma unroll 1
  for (int r = 0; r < 10; r++) {
    OPIR_QR(0, 4, 8, 12) OPIR_QR(1, 5, 9, 13) OPIR_QR(2, 6, 10, 14) OPIR_QR(3, 7, 11, 15)
    OPIR_QR(0, 5, 10, 15) OPIR_QR(1, 6, 11, 12) OPIR_QR(2, 7, 8, 13) OPIR_QR(3, 4, 9, 14)
  }
#undef OPIR_QR
#pragma unroll
  for (int i = 0; i < 16; i++) out[i] = x[i] + st[i];
}

// 64-bit draw m (0..7) of a block.
__device__ __forceinline__ uint64_t draw64(const uint32_t w[16], int m) {
  return (uint64_t)w[2 * m + 1] << 32 | w[2 * m];
}
// 128-bit draw m (0..3) of a block: (hi, lo).
__device__ __forceinline__ void draw128(const uint32_t w[16], int m, uint64_t &hi, uint64_t &lo) {
  lo = (uint64_t)w[4 * m + 1] << 32 | w[4 * m];
  hi = (uint64_t)w[4 * m + 3] << 32 | w[4 * m + 2];
}

}  // namespace prg
}  // namespace opir
