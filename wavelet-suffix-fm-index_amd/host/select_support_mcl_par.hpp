// select_support_mcl_par.hpp -- drop-in for sdsl::select_support_mcl<b,1> of
// the reference (select/select_support_mcl_par.hpp:21-719): same include guard,
// class name, constructor, init_slow, select()/operator(), set_vector, swap,
// copy and move.  Construction runs on the GPU (wtg_select_build over a rank
// directory); the content (superblock starts, 64 mini offsets or 4096 long
// positions per 4096 arguments, logn^4 long/mini rule) is kept on the host so
// select() is the reference's host algorithm (:564-611); select_batch() answers
// many queries on the device.
#ifndef INCLUDED_SDSL_SELECT_SUPPORT_MCL
#define INCLUDED_SDSL_SELECT_SUPPORT_MCL
#include <cassert>
#include <cstdint>
#include <vector>
#include "sdsl_standin.hpp"
#include "wtg_runtime.hpp"

namespace sdsl {

template <uint8_t t_b = 1, uint8_t t_pat_len = 1>
class select_support_mcl {
  static_assert((t_b == 1 || t_b == 0) && t_pat_len == 1,
                "the GPU select samples are built for the bit patterns `1` and `0`");

 public:
  typedef uint64_t size_type;
  typedef bit_vector bit_vector_type;
  enum { bit_pat = t_b };
  enum { bit_pat_len = t_pat_len };

 private:
  const bit_vector* m_v = nullptr;
  uint32_t m_logn = 0, m_logn2 = 0, m_logn4 = 0;
  std::vector<uint64_t> m_superblock;   // sb
  std::vector<uint32_t> m_mini;         // sb*64
  std::vector<int64_t> m_long_index;    // sb, -1 = mini superblock
  std::vector<uint64_t> m_long_pool;    // n_long*4096
  size_type m_arg_cnt = 0;

  static uint64_t argword(uint64_t w) { return t_b ? w : ~w; }
  static uint32_t sel_in_word(uint64_t x, uint64_t i) {  // 1-based i-th set bit
    while (--i) x &= x - 1;
    return (uint32_t)__builtin_ctzll(x);
  }

 public:
  explicit select_support_mcl(const bit_vector* v = nullptr) {  // :120-129
    if (v != nullptr) init_slow(v);
  }
  select_support_mcl(const select_support_mcl&) = default;
  select_support_mcl(select_support_mcl&&) = default;
  select_support_mcl& operator=(const select_support_mcl&) = default;
  select_support_mcl& operator=(select_support_mcl&&) = default;

  // init_slow and init_fast (:227-276, :483-561) produce the same content; both
  // build it on the device here
  void init_slow(const bit_vector* v = nullptr) {
    set_vector(v);
    m_arg_cnt = 0;
    m_superblock.clear(); m_mini.clear(); m_long_index.clear(); m_long_pool.clear();
    if (v == nullptr) return;
    const uint64_t nbits = v->size(), nw = (nbits + 63) / 64;
    wtg::dbuf dw(nw * 8);
    dw.upload(v->data(), nw * 8);
    build_from_device(dw.as<uint64_t>(), nbits);
  }
  void build_from_device(const uint64_t* d_words, uint64_t nbits) {
    wtg::dbuf dd(wtg_rank9_words(nbits) * 8);
    wtg_select s{};
    wtg::check(wtg_rank9_select_build(wtg::default_context(), d_words, nbits, dd.as<uint64_t>(),
                                      t_b ? &s : nullptr, t_b ? nullptr : &s));
    assign(s);
    wtg_select_free(wtg::default_context(), &s);
  }
  // copy a device-resident structure (e.g. from wtg_wtint_select1/0)
  void assign(const wtg_select& s) {
    wtg_ctx* c = wtg::default_context();
    m_arg_cnt = s.arg_cnt;
    m_logn = s.logn;
    m_logn2 = m_logn * m_logn;
    m_logn4 = s.logn4;
    m_superblock.resize(s.sb);
    m_mini.resize(s.sb * 64);
    m_long_index.resize(s.sb);
    m_long_pool.resize(s.n_long * 4096);
    if (s.sb) {
      wtg::check(wtg_memcpy_d2h(c, m_superblock.data(), s.superblock, s.sb * 8));
      wtg::check(wtg_memcpy_d2h(c, m_mini.data(), s.mini, s.sb * 64 * 4));
      wtg::check(wtg_memcpy_d2h(c, m_long_index.data(), s.long_index, s.sb * 8));
    }
    if (s.n_long) wtg::check(wtg_memcpy_d2h(c, m_long_pool.data(), s.long_pool, s.n_long * 4096 * 8));
  }

  size_type select(size_type i) const {  // :564-611
    assert(i > 0 && i <= m_arg_cnt);
    i = i - 1;
    const size_type sb_idx = i >> 12, offset = i & 0xFFF;
    if (m_long_index[sb_idx] >= 0) return m_long_pool[(uint64_t)m_long_index[sb_idx] * 4096 + offset];
    if ((offset & 0x3F) == 0) return m_superblock[sb_idx] + m_mini[sb_idx * 64 + (offset >> 6)];
    i = i - (sb_idx << 12) - ((offset >> 6) << 6);
    const size_type pos = m_superblock[sb_idx] + m_mini[sb_idx * 64 + (offset >> 6)] + 1;
    size_type word_pos = pos >> 6;
    const size_type word_off = pos & 0x3F;
    const uint64_t* data = m_v->data();
    uint64_t x = argword(data[word_pos]) >> word_off;
    size_type args = __builtin_popcountll(x);
    if (args >= i) return (word_pos << 6) + word_off + sel_in_word(x, i);
    size_type sum = args;
    for (;;) {
      x = argword(data[++word_pos]);
      args = __builtin_popcountll(x);
      if (sum + args >= i) return (word_pos << 6) + sel_in_word(x, i - sum);
      sum += args;
    }
  }
  size_type operator()(size_type i) const { return select(i); }
  void set_vector(const bit_vector* v = nullptr) { m_v = v; }
  void swap(select_support_mcl& ss) {
    std::swap(m_logn, ss.m_logn); std::swap(m_logn2, ss.m_logn2); std::swap(m_logn4, ss.m_logn4);
    m_superblock.swap(ss.m_superblock); m_mini.swap(ss.m_mini);
    m_long_index.swap(ss.m_long_index); m_long_pool.swap(ss.m_long_pool);
    std::swap(m_arg_cnt, ss.m_arg_cnt);
  }
  size_type arg_cnt() const { return m_arg_cnt; }
  uint32_t logn() const { return m_logn; }
  uint32_t logn4() const { return m_logn4; }
  const std::vector<uint64_t>& superblock() const { return m_superblock; }
  const std::vector<uint32_t>& mini() const { return m_mini; }
  const std::vector<int64_t>& long_index() const { return m_long_index; }
  const std::vector<uint64_t>& long_pool() const { return m_long_pool; }
};

}  // namespace sdsl
#endif
