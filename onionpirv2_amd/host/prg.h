// Counter-based sampler shared by the host client, the wire codec and the
// device kernels (csrc/opir_prg.h holds the device copy of the same rules).
//
// Every random polynomial of a seeded client is a pure function of a 32-byte
// seed and a stream id (domain, i, j): ChaCha20 blocks (RFC 8439 section 2.3:
// key = the seed as eight little-endian words, block counter = block index,
// nonce = (domain, i, j)). From block b's sixteen output words w:
//   128-bit draw k = 4b + m:  lo = w[4m] | w[4m+1] << 32, hi = w[4m+2] | w[4m+3] << 32
//    64-bit draw k = 8b + m:  w[2m] | w[2m+1] << 32
// and
//   uniform mod q: coefficient k = (hi * 2^64 + lo) mod q of 128-bit draw k
//                  (statistical distance q / 2^128 from uniform);
//   ternary secret: coefficient k = {0, 1, -1}[64-bit draw k mod 3];
//   error: the rounded Gaussian of standard deviation sigma by inversion of
//          a cumulative table over 64-bit draws (gaussian_cdt).
// The reference draws all of these from std::mt19937_64 / std::
// distributions seeded by random_device (src/utils.cpp:407-443); a seeded
// client here is a new, reproducible source with the same distributions, so
// that a server can regenerate the uniform halves of keys and queries from
// their seeds (the seed-compressed sizes of src/includes/pir.h:267-291).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace onionpir {
namespace prg {

using Seed = std::array<uint8_t, 32>;

// Stream domains (nonce word 0).
enum Domain : uint32_t {
  kSecret = 1,     // ternary secret key: (1, 0, 0)
  kQueryA = 2,     // query c1, keyed by the query seed: (2, 0, limb)
  kQueryE = 3,     // query error, keyed by the query seed: (3, 0, 0)
  kBvA = 4,        // BV key row a: (4, level * l_ks + row, limb)
  kBvE = 5,        // BV key row error: (5, level * l_ks + row, 0)
  kGswA = 6,       // RGSW(s) row c1: (6, row, limb)
  kGswE = 7,       // RGSW(s) row error: (7, row, 0)
  kQuerySeed = 8,  // per-query seed: first 32 bytes of block 0 of (8, counter_lo, counter_hi)
};

// Largest |e| the error table covers (the mass beyond is below 2^-64 for
// the reference's sigma = 2.55 and is folded into the outermost entries).
constexpr int kCdtMax = 24;

void chacha20_block(const Seed &key, uint32_t counter, const uint32_t nonce[3], uint32_t out[16]);
// n coefficients uniform mod q of stream (domain, i, j).
void uniform(const Seed &key, uint32_t domain, uint32_t i, uint32_t j, uint64_t q, size_t n, uint64_t *out);
// n ternary coefficients in {-1, 0, 1} of stream (kSecret, 0, 0).
void ternary(const Seed &key, size_t n, int8_t *out);
// Cumulative table of the rounded Gaussian: entry v (0 <= v < 2 kCdtMax + 1)
// is the number of 64-bit draws u with sample <= v - kCdtMax, minus one,
// i.e. the sample of u is the first v with u <= cdt[v], minus kCdtMax.
std::vector<uint64_t> gaussian_cdt(double sigma);
int64_t cdt_sample(const std::vector<uint64_t> &cdt, uint64_t u);
// n errors of stream (domain, i, 0).
void gaussian(const Seed &key, uint32_t domain, uint32_t i, const std::vector<uint64_t> &cdt, size_t n, int64_t *out);
// The 32-byte seed of query number `counter` of a client keyed by `key`.
Seed query_seed(const Seed &key, uint64_t counter);

}  // namespace prg
}  // namespace onionpir
