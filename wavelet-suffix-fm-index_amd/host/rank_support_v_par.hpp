// rank_support_v_par.hpp -- drop-in for sdsl::rank_support_v<1,1> of the
// reference (rank/rank_support_v_par.hpp:21-183): same include guard, class
// name, constructor, rank()/operator()/size()/set_vector()/swap() and the
// interleaved directory m_basic_block (superblock count + seven 9-bit block
// counts per 512 bits).  The directory is built on the GPU (wtg_rank9_build)
// and kept on the host, so rank() is a const, thread-safe host call exactly as
// in the reference; rank_batch() answers many queries on the device.
#ifndef INCLUDED_SDSL_RANK_SUPPORT_V
#define INCLUDED_SDSL_RANK_SUPPORT_V
#include <cstdint>
#include <istream>
#include <ostream>
#include <string>
#include <vector>
#include "sdsl_standin.hpp"
#include "wtg_runtime.hpp"

namespace sdsl {

template <uint8_t t_b = 1, uint8_t t_pat_len = 1>
class rank_support_v {
  static_assert(t_b == 1 && t_pat_len == 1,
                "the GPU rank directory is built for the bit pattern `1` (rank_1_type)");

 public:
  typedef uint64_t size_type;
  typedef bit_vector bit_vector_type;
  enum { bit_pat = t_b };
  enum { bit_pat_len = t_pat_len };

 private:
  const bit_vector* m_v = nullptr;
  std::vector<uint64_t> m_basic_block;

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
    wtg::dbuf dd(nd * 8);
    wtg::check(wtg_rank9_build(wtg::default_context(), d_words, nbits, dd.as<uint64_t>()));
    m_basic_block.resize(nd);
    dd.download(m_basic_block.data(), nd * 8);
  }
  void assign_directory(const uint64_t* d_dir, uint64_t nbits) {
    const uint64_t nd = wtg_rank9_words(nbits);
    m_basic_block.resize(nd);
    wtg::check(wtg_memcpy_d2h(wtg::default_context(), m_basic_block.data(), d_dir, nd * 8));
  }

  size_type rank(size_type idx) const {  // :131-142
    const uint64_t* p = m_basic_block.data() + ((idx >> 8) & 0xFFFFFFFFFFFFFFFEULL);
    size_type r = *p + ((*(p + 1) >> (63 - 9 * ((idx & 0x1FF) >> 6))) & 0x1FF);
    if (idx & 0x3F) r += __builtin_popcountll(m_v->data()[idx >> 6] & ((1ULL << (idx & 63)) - 1));
    return r;
  }
  size_type operator()(size_type idx) const { return rank(idx); }
  size_type size() const { return m_v->size(); }

  // sdsl int_vector<64> layout of m_basic_block: bit size (uint64), then the words
  size_type serialize(std::ostream& out, structure_tree_node* = nullptr, std::string = "") const {
    const uint64_t bits = m_basic_block.size() * 64;
    out.write((const char*)&bits, 8);
    out.write((const char*)m_basic_block.data(), m_basic_block.size() * 8);
    return 8 + m_basic_block.size() * 8;
  }
  void load(std::istream& in, const bit_vector* v = nullptr) {
    set_vector(v);
    uint64_t bits = 0;
    in.read((char*)&bits, 8);
    m_basic_block.resize(bits / 64);
    in.read((char*)m_basic_block.data(), m_basic_block.size() * 8);
  }
  void set_vector(const bit_vector* v = nullptr) { m_v = v; }
  void swap(rank_support_v& rs) {
    if (this != &rs) m_basic_block.swap(rs.m_basic_block);
  }
  const std::vector<uint64_t>& basic_block() const { return m_basic_block; }

  // rank(idx) for a batch of device indices, answered on the device
  void rank_batch(const uint64_t* d_words, const uint64_t* d_idx, uint64_t nq, uint64_t* d_out) const {
    wtg::dbuf dd(m_basic_block.size() * 8);
    dd.upload(m_basic_block.data(), m_basic_block.size() * 8);
    wtg::check(wtg_rank9_batch(wtg::default_context(), d_words, dd.as<uint64_t>(),
                               m_v ? m_v->size() : 0, d_idx, nq, d_out));
  }
};

}  // namespace sdsl
#endif
