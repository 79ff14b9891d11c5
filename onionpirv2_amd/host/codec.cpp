// Wire codecs of the host API (codec.h).
#include "codec.h"

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
