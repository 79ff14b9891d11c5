// Wire codecs of the host API.
//
// Bit packing is the reference's response format (src/server.cpp:950-1009,
// src/client.cpp:315-361): values of `width` bits, least significant bit
// first, concatenated in order; the final partial byte is zero-padded.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace onionpir {
namespace codec {

// Appends n values of `width` bits (1..64; each value's bits above width
// are dropped) to out. Returns the bytes appended.
size_t pack_bits(const uint64_t *vals, size_t n, size_t width, std::string &out);
// Reads n values of `width` bits from data[pos...]; advances pos past the
// bytes consumed (the last one may be partial). Throws std::runtime_error if
// the buffer ends early.
void unpack_bits(const std::string &data, size_t &pos, size_t n, size_t width, uint64_t *vals);
inline size_t packed_bytes(size_t n, size_t width) { return (n * width + 7) / 8; }

}  // namespace codec
}  // namespace onionpir

// Seed-compressed wire formats of queries and client keys. The reference
// models their sizes only (src/includes/pir.h:267-291: 32 seed bytes plus
// the non-uniform half at ct_mod_width bits per coefficient) and has no
// encoder; these formats realise that model for seeded clients
// (PirClient(params, id, key)). Each limb k's values are packed at
// width_k = bit length of (modulus_k - 1) bits (Q for the composite scheme,
// the RNS primes otherwise), limbs in order, with pack_bits.
//   query: [32-byte query seed][c0: K limbs x N values]
//          c1 = uniform(query seed, kQueryA, 0, limb) is regenerated.
//   keys:  [32-byte client key][h: 1 byte]
//          [BV rows b: h levels x l_ks rows x K limbs x N, NTT form]
//          [RGSW(s) rows c0: 2 l_key rows x K limbs x N, NTT form]
//          a / c1 = uniform(key, kBvA, level * l_ks + row, limb) and
//          uniform(key, kGswA, row, limb), NTT form, are regenerated.
#include "onionpir_host.h"
#include "prg.h"

namespace onionpir {
namespace codec {

struct SeededKeys {
  prg::Seed key{};
  size_t h = 0;
  std::vector<uint64_t> bv_b;    // [h][l_ks][K*N]
  std::vector<uint64_t> gsw_c0;  // [2 l_key][K*N]
};

size_t query_wire_bytes(const PirParams &p);
size_t keys_wire_bytes(const PirParams &p, size_t h);
std::string encode_query(const PirParams &p, const prg::Seed &qseed, const RlweCt &query);
// Throws std::invalid_argument on a wrong length or an unreduced value.
RlweCt decode_query(const PirParams &p, const std::string &wire);
std::string encode_keys(const PirParams &p, const prg::Seed &key, const bvks::BvGaloisKeys &bv, const GSWCt &gsw);
SeededKeys decode_keys(const PirParams &p, const std::string &wire);
// The full keys (uniform halves regenerated on the host).
void expand_keys(const PirParams &p, const SeededKeys &k, bvks::BvGaloisKeys &bv, GSWCt &gsw);

}  // namespace codec
}  // namespace onionpir
