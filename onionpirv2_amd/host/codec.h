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
