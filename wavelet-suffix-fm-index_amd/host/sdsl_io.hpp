// sdsl_io.hpp -- the byte images the drop-in classes serialize to and load
// from, in sdsl-lite's stream format:
//   member       its bytes as laid out in memory (write_member / read_member);
//   int_vector   the size in BITS as uint64, then -- only for a variable-width
//                int_vector<0> -- the element width as uint8, then
//                ceil(bits/64) little-endian 64-bit words; element i of width w
//                occupies bit positions [i*w, i*w + w) (bit j of word k is
//                position 64k + j), the bits past the end are zero;
//   bit_vector   int_vector<1>: the size in bits, then the words.
// The same images whether or not sdsl-lite itself is on the include path.
#pragma once
#include <cstdint>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace wtg {
namespace sdslio {

template <class T>
inline uint64_t write_member(const T& x, std::ostream& out) {
  out.write(reinterpret_cast<const char*>(&x), sizeof(T));
  return sizeof(T);
}
template <class T>
inline void read_member(T& x, std::istream& in) {
  in.read(reinterpret_cast<char*>(&x), sizeof(T));
  if (!in) throw std::runtime_error("sdsl image: unexpected end of stream");
}

inline uint64_t write_words(std::ostream& out, const uint64_t* w, uint64_t nwords) {
  if (nwords) out.write(reinterpret_cast<const char*>(w), (std::streamsize)(nwords * 8));
  return nwords * 8;
}
inline void read_words(std::istream& in, uint64_t* w, uint64_t nwords) {
  if (nwords) in.read(reinterpret_cast<char*>(w), (std::streamsize)(nwords * 8));
  if (!in) throw std::runtime_error("sdsl image: unexpected end of stream");
}

// bit_vector image of nbits bits held in words (bits past nbits must be zero)
inline uint64_t write_bits(std::ostream& out, const uint64_t* words, uint64_t nbits) {
  return write_member(nbits, out) + write_words(out, words, (nbits + 63) / 64);
}
// reads a bit_vector image: words gets ceil(nbits/64) words plus one zero padding word
inline void read_bits(std::istream& in, std::vector<uint64_t>& words, uint64_t& nbits) {
  read_member(nbits, in);
  words.assign((nbits + 63) / 64 + 1, 0);
  read_words(in, words.data(), (nbits + 63) / 64);
}

// int_vector<64> image of `count` 64-bit values
inline uint64_t write_u64s(std::ostream& out, const uint64_t* v, uint64_t count) {
  return write_member((uint64_t)(count * 64), out) + write_words(out, v, count);
}
inline void read_u64s(std::istream& in, std::vector<uint64_t>& v) {
  uint64_t bits = 0;
  read_member(bits, in);
  if (bits % 64) throw std::runtime_error("sdsl image: int_vector<64> size is not a multiple of 64 bits");
  v.resize(bits / 64);
  read_words(in, v.data(), v.size());
}

// int_vector<0> image of `count` values get(i), each < 2^width (1 <= width <= 64)
template <class Get>
inline uint64_t write_packed(std::ostream& out, uint64_t count, uint8_t width, Get get) {
  if (width == 0 || width > 64) throw std::runtime_error("sdsl image: int_vector width outside 1..64");
  const uint64_t bits = count * width, nw = (bits + 63) / 64;
  std::vector<uint64_t> w(nw, 0);
  const uint64_t mask = width == 64 ? ~0ull : ((1ull << width) - 1);
  for (uint64_t i = 0; i < count; i++) {
    const uint64_t v = get(i) & mask, p = i * width, k = p >> 6, o = p & 63;
    w[k] |= v << o;
    if (o + width > 64) w[k + 1] |= v >> (64 - o);
  }
  uint64_t n = write_member(bits, out) + write_member(width, out);
  return n + write_words(out, w.data(), nw);
}
inline void read_packed(std::istream& in, std::vector<uint64_t>& vals, uint8_t& width) {
  uint64_t bits = 0;
  read_member(bits, in);
  read_member(width, in);
  if (width == 0 || width > 64 || bits % width)
    throw std::runtime_error("sdsl image: bad int_vector<0> header");
  const uint64_t count = bits / width, nw = (bits + 63) / 64;
  std::vector<uint64_t> w(nw + 1, 0);
  read_words(in, w.data(), nw);
  const uint64_t mask = width == 64 ? ~0ull : ((1ull << width) - 1);
  vals.resize(count);
  for (uint64_t i = 0; i < count; i++) {
    const uint64_t p = i * width, k = p >> 6, o = p & 63;
    uint64_t v = w[k] >> o;
    if (o + width > 64) v |= w[k + 1] << (64 - o);
    vals[i] = v & mask;
  }
}

}  // namespace sdslio
}  // namespace wtg
