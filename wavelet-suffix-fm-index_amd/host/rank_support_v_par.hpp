// rank_support_v_par.hpp -- drop-in for sdsl::rank_support_v<b,1> of the
// reference (rank/rank_support_v_par.hpp:21-183): same include guard, class
// name, constructor, rank()/operator()/size()/serialize()/load()/set_vector()/
// swap() and the interleaved directory m_basic_block (superblock count + seven
// 9-bit block counts per 512 bits).  The directory is built on the GPU
// (wtg_rank9_build) and kept on the host, so rank() is a const, thread-safe
// host call exactly as in the reference; rank_batch() answers many queries on
// the device.
//
// Bit patterns: `1` (rank_1_type), `0` (bit_vector::rank_0_type, counting
// zeros over whole words, so the directory is 64*words - the ones' directory)
// and the two-bit patterns `10` and `01` of the reference (:57-58, template
// arguments <10,2> and <01,2>): an occurrence sits at position i when bits
// i-1, i read the pattern (sdsl's rank_support_trait: the previous word's top
// bit carries into the next word, and position 0 is never one).  Their
// directory is the ones' directory of the argument words (wtg_pattern_words),
// counted over whole words as the reference's block pass (:84-105) does.
#ifndef INCLUDED_SDSL_RANK_SUPPORT_V
#define INCLUDED_SDSL_RANK_SUPPORT_V
#include <cstdint>
#include <istream>
#include <ostream>
#include <string>
#include <vector>
#include "sdsl_io.hpp"
#include "sdsl_standin.hpp"
#include "wtg_runtime.hpp"

namespace sdsl {

template <uint8_t t_b = 1, uint8_t t_pat_len = 1>
class rank_support_v {
  static_assert((t_pat_len == 1 && (t_b == 1 || t_b == 0)) || (t_pat_len == 2 && (t_b == 10 || t_b == 1)),
                "rank_support_v: bit pattern must be `0`, `1`, `10` or `01`");
  static constexpr bool kPair = t_pat_len == 2;

 public:
  typedef uint64_t size_type;
  typedef bit_vector bit_vector_type;
  enum { bit_pat = t_b };
  enum { bit_pat_len = t_pat_len };

 private:
  const bit_vector* m_v = nullptr;
  std::vector<uint64_t> m_basic_block;

  // the arguments of word k: pattern `0`, the zeros of whole words (bits past
  // size() included, as the reference's trait counts ~word); two-bit
  // patterns, bit i = (bit i-1, bit i) == the pattern, the carry being the
  // previous word's top bit (before word 0: 0 for `10`, 1 for `01`)
  uint64_t argword(uint64_t k) const {
    const uint64_t* d = m_v->data();
    if (!kPair) return t_b ? d[k] : ~d[k];
    const uint64_t x = d[k], prev = (x << 1) | (k ? d[k - 1] >> 63 : (t_b == 10 ? 0ull : 1ull));
    return t_b == 10 ? prev & ~x : ~prev & x;
  }

  // the ones' directory (GPU) -> the zeros' directory: every count c over j
  // whole words becomes 64*j - c
  void to_pattern(uint64_t nbits) {
    if (t_b || nbits == 0) return;
    const uint64_t W = (nbits + 63) / 64, S = m_basic_block.size() / 2 - 1;
    for (uint64_t b = 0; b <= S; b++) {
      m_basic_block[2 * b] = 512 * b - m_basic_block[2 * b];
      const uint64_t m = W > 8 * b ? (W - 8 * b < 8 ? W - 8 * b : 8) : 0;
      uint64_t F = m_basic_block[2 * b + 1], G = 0;
      for (uint64_t j = 1; j < 8 && j <= m; j++) {
        const uint64_t c = (F >> (63 - 9 * j)) & 0x1FF;
        G |= (64 * j - c) << (63 - 9 * j);
      }
      m_basic_block[2 * b + 1] = G;
    }
  }

 public:
  explicit rank_support_v(const bit_vector* v = nullptr) {  // :69-123
    set_vector(v);
    if (v == nullptr) return;
    if (v->empty()) {
      m_basic_block.assign(2, 0);
      return;
    }
    const uint64_t nbits = v->size(), nw = (nbits + 63) / 64;
    wtg::dbuf dw(nw * 8);
    dw.upload(v->data(), nw * 8);
    build_from_device(dw.as<uint64_t>(), nbits);
  }
  rank_support_v(const rank_support_v&) = default;
  rank_support_v(rank_support_v&&) = default;
  rank_support_v& operator=(const rank_support_v&) = default;
  rank_support_v& operator=(rank_support_v&&) = default;

  // directory of words already resident on the device
  void build_from_device(const uint64_t* d_words, uint64_t nbits) {
    const uint64_t nd = wtg_rank9_words(nbits);
    wtg::dbuf dd(nd * 8), dp;
    if (kPair) {  // the argument words, whole words counted
      dp.alloc((nbits + 63) / 64 * 8);
      wtg::check(wtg_pattern_words(wtg::default_context(), d_words, nbits, t_b, 0, dp.as<uint64_t>()));
      d_words = dp.as<uint64_t>();
    }
    wtg::check(wtg_rank9_build(wtg::default_context(), d_words, nbits, dd.as<uint64_t>()));
    m_basic_block.resize(nd);
    dd.download(m_basic_block.data(), nd * 8);
    to_pattern(nbits);
  }
  // the ones' directory of nbits bits, already built on the device (of the
  // argument words, for a two-bit pattern)
  void assign_directory(const uint64_t* d_dir, uint64_t nbits) {
    const uint64_t nd = wtg_rank9_words(nbits);
    m_basic_block.resize(nd);
    wtg::check(wtg_memcpy_d2h(wtg::default_context(), m_basic_block.data(), d_dir, nd * 8));
    to_pattern(nbits);
  }

  size_type rank(size_type idx) const {  // :131-142
    const uint64_t* p = m_basic_block.data() + ((idx >> 8) & 0xFFFFFFFFFFFFFFFEULL);
    size_type r = *p + ((*(p + 1) >> (63 - 9 * ((idx & 0x1FF) >> 6))) & 0x1FF);
    if (idx & 0x3F) r += __builtin_popcountll(argword(idx >> 6) & ((1ULL << (idx & 63)) - 1));
    return r;
  }
  size_type operator()(size_type idx) const { return rank(idx); }
  size_type size() const { return m_v->size(); }

  // m_basic_block as an sdsl int_vector<64> image (rank_support_v_par.hpp:154-164)
  size_type serialize(std::ostream& out, structure_tree_node* = nullptr, std::string = "") const {
    return wtg::sdslio::write_u64s(out, m_basic_block.data(), m_basic_block.size());
  }
  void load(std::istream& in, const bit_vector* v = nullptr) {  // :166-170
    set_vector(v);
    wtg::sdslio::read_u64s(in, m_basic_block);
  }
  void set_vector(const bit_vector* v = nullptr) { m_v = v; }
  void swap(rank_support_v& rs) {
    if (this != &rs) m_basic_block.swap(rs.m_basic_block);
  }
  const std::vector<uint64_t>& basic_block() const { return m_basic_block; }

  // rank(idx) for a batch of device indices, answered on the device from the
  // ones' directory
  void rank_batch(const uint64_t* d_words, const uint64_t* d_idx, uint64_t nq, uint64_t* d_out) const {
    std::vector<uint64_t> ones_dir(m_basic_block);
    const uint64_t nbits = m_v ? m_v->size() : 0;
    if (!t_b) {  // back to the ones' directory
      const uint64_t W = (nbits + 63) / 64, S = ones_dir.size() / 2 - 1;
      for (uint64_t b = 0; b <= S && nbits; b++) {
        ones_dir[2 * b] = 512 * b - ones_dir[2 * b];
        const uint64_t m = W > 8 * b ? (W - 8 * b < 8 ? W - 8 * b : 8) : 0;
        uint64_t F = ones_dir[2 * b + 1], G = 0;
        for (uint64_t j = 1; j < 8 && j <= m; j++) G |= (64 * j - ((F >> (63 - 9 * j)) & 0x1FF)) << (63 - 9 * j);
        ones_dir[2 * b + 1] = G;
      }
    }
    wtg::dbuf dd(ones_dir.size() * 8), dp;
    dd.upload(ones_dir.data(), ones_dir.size() * 8);
    if (kPair && nbits) {  // the argument words' ones
      dp.alloc((nbits + 63) / 64 * 8);
      wtg::check(wtg_pattern_words(wtg::default_context(), d_words, nbits, t_b, 0, dp.as<uint64_t>()));
      d_words = dp.as<uint64_t>();
    }
    wtg::check(wtg_rank9_batch_bit(wtg::default_context(), d_words, dd.as<uint64_t>(), nbits, kPair ? 1 : t_b, d_idx,
                                   nq, d_out));
  }
};

}  // namespace sdsl
#endif
