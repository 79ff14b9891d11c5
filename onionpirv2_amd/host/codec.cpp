// Wire codecs of the host API (codec.h).
#include "codec.h"

#include <bit>
#include <cstring>
#include <stdexcept>

namespace onionpir {
namespace codec {

size_t pack_bits(const uint64_t *vals, size_t n, size_t width, std::string &out) {
  if (width == 0 || width > 64) throw std::invalid_argument("bit width must be 1..64");
  const uint64_t mask = width == 64 ? ~uint64_t{0} : (uint64_t{1} << width) - 1;
  // at most 7 pending bits plus one value of up to 64 bits
  unsigned __int128 acc = 0;
  size_t have = 0, bytes = 0;
  for (size_t i = 0; i < n; i++) {
    acc |= (unsigned __int128)(vals[i] & mask) << have;
    have += width;
    while (have >= 8) {
      out.push_back((char)(uint8_t)acc);
      acc >>= 8;
      have -= 8;
      bytes++;
    }
  }
  if (have) {
    out.push_back((char)(uint8_t)acc);
    bytes++;
  }
  return bytes;
}

void unpack_bits(const std::string &data, size_t &pos, size_t n, size_t width, uint64_t *vals) {
  if (width == 0 || width > 64) throw std::invalid_argument("bit width must be 1..64");
  const uint64_t mask = width == 64 ? ~uint64_t{0} : (uint64_t{1} << width) - 1;
  unsigned __int128 acc = 0;
  size_t have = 0;
  for (size_t i = 0; i < n; i++) {
    while (have < width) {
      if (pos >= data.size()) throw std::runtime_error("unexpected end of stream");
      acc |= (unsigned __int128)(uint8_t)data[pos++] << have;
      have += 8;
    }
    vals[i] = (uint64_t)acc & mask;
    acc >>= width;
    have -= width;
  }
}

}  // namespace codec
}  // namespace onionpir

namespace onionpir {
namespace codec {

namespace {
size_t width(uint64_t m) { return (size_t)std::bit_width(m - 1); }


// One K-limb polynomial (limb-major) packed limb by limb.
void put_poly(const PirParams &p, const uint64_t *v, std::string &out) {
  const size_t N = p.get_poly_degree();
  const auto &mods = p.get_rns_mods();
  for (size_t k = 0; k < mods.size(); k++) pack_bits(v + k * N, N, width(mods[k]), out);
}

void get_poly(const PirParams &p, const std::string &in, size_t &pos, uint64_t *v) {
  const size_t N = p.get_poly_degree();
  const auto &mods = p.get_rns_mods();
  for (size_t k = 0; k < mods.size(); k++) {
    unpack_bits(in, pos, N, width(mods[k]), v + k * N);
    for (size_t j = 0; j < N; j++)
      if (v[k * N + j] >= mods[k]) throw std::invalid_argument("wire value not reduced mod its modulus");
  }
}

prg::Seed get_seed(const std::string &in, size_t &pos) {
  if (in.size() < pos + 32) throw std::invalid_argument("wire too short for its seed");
  prg::Seed s;
  std::memcpy(s.data(), in.data() + pos, 32);
  pos += 32;
  return s;
}
}  // namespace

size_t query_wire_bytes(const PirParams &p) {
  // each limb's values are packed on their own, so every limb rounds up to whole bytes
  size_t b = 32;
  for (uint64_t m : p.get_rns_mods()) b += packed_bytes(p.get_poly_degree(), width(m));
  return b;
}

size_t keys_wire_bytes(const PirParams &p, size_t h) {
  size_t poly = 0;
  for (uint64_t m : p.get_rns_mods()) poly += packed_bytes(p.get_poly_degree(), width(m));
  return 33 + (h * p.get_l_ks() + 2 * p.get_l_key()) * poly;
}

std::string encode_query(const PirParams &p, const prg::Seed &qseed, const RlweCt &query) {
  const size_t KN = p.get_poly_degree() * p.K();
  if (query.c0.size() != KN) throw std::invalid_argument("query c0 must hold K*N values");
  std::string out(reinterpret_cast<const char *>(qseed.data()), 32);
  put_poly(p, query.c0.data(), out);
  return out;
}

RlweCt decode_query(const PirParams &p, const std::string &wire) {
  if (wire.size() != query_wire_bytes(p)) throw std::invalid_argument("query wire has the wrong length");
  const size_t N = p.get_poly_degree(), K = p.K();
  size_t pos = 0;
  const prg::Seed qs = get_seed(wire, pos);
  RlweCt q;
  q.c0.resize(K * N);
  q.c1.resize(K * N);
  get_poly(p, wire, pos, q.c0.data());
  for (size_t k = 0; k < K; k++) prg::uniform(qs, prg::kQueryA, 0, (uint32_t)k, p.get_rns_mods()[k], N, q.c1.data() + k * N);
  q.ntt_form = false;
  return q;
}

std::string encode_keys(const PirParams &p, const prg::Seed &key, const bvks::BvGaloisKeys &bv, const GSWCt &gsw) {
  const size_t KN = p.get_poly_degree() * p.K(), h = bv.keys.size();
  if (h > 255) throw std::invalid_argument("too many BV keys");
  if (gsw.size() != 2 * p.get_l_key()) throw std::invalid_argument("RGSW(s) must hold 2 l_key rows");
  std::string out(reinterpret_cast<const char *>(key.data()), 32);
  out.push_back((char)(uint8_t)h);
  for (const auto &k : bv.keys) {
    if (k.cts.size() != p.get_l_ks()) throw std::invalid_argument("BV key must hold l_ks rows");
    for (const auto &row : k.cts) {
      if (row.b.size() != KN) throw std::invalid_argument("BV row must hold K*N values");
      put_poly(p, row.b.data(), out);
    }
  }
  for (const auto &row : gsw) {
    if (row.size() != 2 * KN) throw std::invalid_argument("RGSW(s) row must hold 2*K*N values");
    put_poly(p, row.data(), out);
  }
  return out;
}

SeededKeys decode_keys(const PirParams &p, const std::string &wire) {
  const size_t KN = p.get_poly_degree() * p.K(), L = p.get_l_ks(), lk = p.get_l_key();
  SeededKeys k;
  size_t pos = 0;
  k.key = get_seed(wire, pos);
  if (wire.size() < 33) throw std::invalid_argument("key wire too short");
  k.h = (uint8_t)wire[pos++];
  if (wire.size() != keys_wire_bytes(p, k.h)) throw std::invalid_argument("key wire has the wrong length");
  k.bv_b.resize(k.h * L * KN);
  for (size_t r = 0; r < k.h * L; r++) get_poly(p, wire, pos, k.bv_b.data() + r * KN);
  k.gsw_c0.resize(2 * lk * KN);
  for (size_t r = 0; r < 2 * lk; r++) get_poly(p, wire, pos, k.gsw_c0.data() + r * KN);
  return k;
}

void expand_keys(const PirParams &p, const SeededKeys &k, bvks::BvGaloisKeys &bv, GSWCt &gsw) {
  const size_t N = p.get_poly_degree(), K = p.K(), KN = K * N, L = p.get_l_ks(), lk = p.get_l_key();
  const auto &mods = p.get_rns_mods();
  bv.keys.clear();
  for (size_t lvl = 0; lvl < k.h; lvl++) {
    bvks::BvKeySwitchKey key;
    key.galois_k = (uint32_t)((N >> lvl) + 1);
    key.cts.resize(L);
    for (size_t r = 0; r < L; r++) {
      auto &row = key.cts[r];
      row.b.assign(k.bv_b.begin() + (lvl * L + r) * KN, k.bv_b.begin() + (lvl * L + r + 1) * KN);
      row.a.resize(KN);
      for (size_t m = 0; m < K; m++)
        prg::uniform(k.key, prg::kBvA, (uint32_t)(lvl * L + r), (uint32_t)m, mods[m], N, row.a.data() + m * N);
    }
    bv.keys.push_back(std::move(key));
  }
  gsw.assign(2 * lk, std::vector<uint64_t>(2 * KN));
  for (size_t r = 0; r < 2 * lk; r++) {
    std::copy(k.gsw_c0.begin() + r * KN, k.gsw_c0.begin() + (r + 1) * KN, gsw[r].begin());
    for (size_t m = 0; m < K; m++)
      prg::uniform(k.key, prg::kGswA, (uint32_t)r, (uint32_t)m, mods[m], N, gsw[r].data() + KN + m * N);
  }
}

}  // namespace codec
}  // namespace onionpir
