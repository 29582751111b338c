// select_support_mcl_par.hpp -- drop-in for sdsl::select_support_mcl<b,1> of
// the reference (select/select_support_mcl_par.hpp:21-719): same include guard,
// class name, constructor, init_slow / init_fast, select()/operator(),
// serialize / load, set_vector, swap, copy and move.  Construction runs on the
// GPU (wtg_rank9_select_build); the content (superblock starts, 64 mini offsets
// or 4096 long positions per 4096 arguments, logn^4 long/mini rule, and each
// block's bit width) is kept on the host so select() is the reference's host
// algorithm (:564-611).
//
// Bit patterns `1`, `0` and the two-bit patterns `10` and `01` of the
// reference (:76-77, template arguments <10,2> and <01,2>): an occurrence sits
// at position i < size() when bits i-1, i read the pattern (position 0 never;
// init_slow's found_arg scan, :240-245).  Their content is the ones' content
// of the argument words (wtg_pattern_words, clipped at size()).
#ifndef INCLUDED_SDSL_SELECT_SUPPORT_MCL
#define INCLUDED_SDSL_SELECT_SUPPORT_MCL
#include <cassert>
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
class select_support_mcl {
  static_assert((t_pat_len == 1 && (t_b == 1 || t_b == 0)) || (t_pat_len == 2 && (t_b == 10 || t_b == 1)),
                "select_support_mcl: bit pattern must be `0`, `1`, `10` or `01`");
  static constexpr bool kPair = t_pat_len == 2;

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
  std::vector<uint8_t> m_mini_width;    // sb: hi(end-start)+1, 0 for a long superblock
  std::vector<int64_t> m_long_index;    // sb, -1 = mini superblock
  std::vector<uint8_t> m_long_width;    // sb: hi(end)+1, 0 for a mini superblock
  std::vector<uint64_t> m_long_pool;    // n_long*4096
  size_type m_arg_cnt = 0;

  // the arguments of word k (for a two-bit pattern: bit i = (bit i-1, bit i)
  // == the pattern, the previous word's top bit carried in)
  uint64_t argword(uint64_t k) const {
    const uint64_t* d = m_v->data();
    if (!kPair) return t_b ? d[k] : ~d[k];
    const uint64_t x = d[k], prev = (x << 1) | (k ? d[k - 1] >> 63 : (t_b == 10 ? 0ull : 1ull));
    return t_b == 10 ? prev & ~x : ~prev & x;
  }
  static uint32_t sel_in_word(uint64_t x, uint64_t i) {  // 1-based i-th set bit
    while (--i) x &= x - 1;
    return (uint32_t)__builtin_ctzll(x);
  }
  void clear() {
    m_arg_cnt = 0;
    m_logn = m_logn2 = m_logn4 = 0;
    m_superblock.clear(); m_mini.clear(); m_mini_width.clear();
    m_long_index.clear(); m_long_width.clear(); m_long_pool.clear();
  }

 public:
  explicit select_support_mcl(const bit_vector* v = nullptr) {  // :120-129
    if (v != nullptr) init_slow(v);
  }
  select_support_mcl(const select_support_mcl&) = default;
  select_support_mcl(select_support_mcl&&) = default;
  select_support_mcl& operator=(const select_support_mcl&) = default;
  select_support_mcl& operator=(select_support_mcl&&) = default;

  // init_slow (:227-276) and init_fast (:483-561) give the same content; the
  // reference picks one by size (:120-129), here both build it on the device
  void init_slow(const bit_vector* v = nullptr) {
    set_vector(v);
    clear();
    if (v == nullptr) return;
    const uint64_t nbits = v->size(), nw = (nbits + 63) / 64;
    wtg::dbuf dw(nw * 8);
    dw.upload(v->data(), nw * 8);
    build_from_device(dw.as<uint64_t>(), nbits);
  }
  void init_fast(const bit_vector* v = nullptr) { init_slow(v); }

  void build_from_device(const uint64_t* d_words, uint64_t nbits) {
    wtg::dbuf dd(wtg_rank9_words(nbits) * 8), dp;
    if (kPair && nbits) {  // the argument words, positions below nbits
      dp.alloc((nbits + 63) / 64 * 8);
      wtg::check(wtg_pattern_words(wtg::default_context(), d_words, nbits, t_b, 1, dp.as<uint64_t>()));
      d_words = dp.as<uint64_t>();
    }
    wtg_select s{};
    const bool ones = kPair || t_b;
    wtg::check(wtg_rank9_select_build(wtg::default_context(), d_words, nbits, dd.as<uint64_t>(),
                                      ones ? &s : nullptr, ones ? nullptr : &s));
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
    m_mini_width.resize(s.sb);
    m_long_index.resize(s.sb);
    m_long_width.resize(s.sb);
    m_long_pool.resize(s.n_long * 4096);
    if (s.sb) {
      wtg::check(wtg_memcpy_d2h(c, m_superblock.data(), s.superblock, s.sb * 8));
      wtg::check(wtg_memcpy_d2h(c, m_mini.data(), s.mini, s.sb * 64 * 4));
      wtg::check(wtg_memcpy_d2h(c, m_mini_width.data(), s.mini_width, s.sb));
      wtg::check(wtg_memcpy_d2h(c, m_long_index.data(), s.long_index, s.sb * 8));
      wtg::check(wtg_memcpy_d2h(c, m_long_width.data(), s.long_width, s.sb));
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
    uint64_t x = argword(word_pos) >> word_off;
    size_type args = __builtin_popcountll(x);
    if (args >= i) return (word_pos << 6) + word_off + sel_in_word(x, i);
    size_type sum = args;
    for (;;) {
      x = argword(++word_pos);
      args = __builtin_popcountll(x);
      if (sum + args >= i) return (word_pos << 6) + sel_in_word(x, i - sum);
      sum += args;
    }
  }
  size_type operator()(size_type i) const { return select(i); }

  // The reference's stream layout (:644-682): the argument count (uint64);
  // then, if it is nonzero, the superblock starts as an int_vector<0> of width
  // logn, a bit_vector with bit k = 1 for a mini superblock k (empty when no
  // superblock is long), and per superblock either its 4096 positions (long,
  // width hi(end)+1) or its 64 offsets (mini, width hi(end-start)+1), each an
  // int_vector<0>.
  size_type serialize(std::ostream& out, structure_tree_node* = nullptr, std::string = "") const {
    using namespace wtg::sdslio;
    size_type n = write_member((uint64_t)m_arg_cnt, out);
    if (!m_arg_cnt) return n;
    const uint64_t sb = m_superblock.size();
    n += write_packed(out, sb, (uint8_t)m_logn, [&](uint64_t k) { return m_superblock[k]; });
    bool any_long = false;
    for (uint64_t k = 0; k < sb; k++) any_long |= m_long_index[k] >= 0;
    std::vector<uint64_t> ml(any_long ? (sb + 63) / 64 : 0, 0);
    if (any_long)
      for (uint64_t k = 0; k < sb; k++)
        if (m_long_index[k] < 0) ml[k >> 6] |= 1ull << (k & 63);
    n += write_bits(out, ml.data(), any_long ? sb : 0);
    for (uint64_t k = 0; k < sb; k++) {
      if (m_long_index[k] >= 0) {
        const uint64_t* p = m_long_pool.data() + (uint64_t)m_long_index[k] * 4096;
        n += write_packed(out, 4096, m_long_width[k], [&](uint64_t t) { return p[t]; });
      } else {
        const uint32_t* p = m_mini.data() + k * 64;
        n += write_packed(out, 64, m_mini_width[k], [&](uint64_t t) { return (uint64_t)p[t]; });
      }
    }
    return n;
  }
  void load(std::istream& in, const bit_vector* v = nullptr) {  // :685-717
    using namespace wtg::sdslio;
    set_vector(v);
    clear();
    uint64_t m = 0;
    read_member(m, in);
    m_arg_cnt = m;
    if (!m) return;
    const uint64_t sb = (m + 4095) >> 12;
    std::vector<uint64_t> vals;
    uint8_t w = 0;
    read_packed(in, vals, w);
    if (vals.size() != sb) throw std::runtime_error("select_support_mcl::load: superblock count mismatch");
    m_superblock = vals;
    m_logn = w;
    m_logn2 = m_logn * m_logn;
    m_logn4 = m_logn2 * m_logn2;
    std::vector<uint64_t> ml;
    uint64_t mlbits = 0;
    read_bits(in, ml, mlbits);
    m_mini.assign(sb * 64, 0);
    m_mini_width.assign(sb, 0);
    m_long_index.assign(sb, -1);
    m_long_width.assign(sb, 0);
    for (uint64_t k = 0; k < sb; k++) {
      const bool is_long = mlbits && !((ml[k >> 6] >> (k & 63)) & 1);
      read_packed(in, vals, w);
      if (is_long) {
        if (vals.size() != 4096) throw std::runtime_error("select_support_mcl::load: long block size");
        m_long_index[k] = (int64_t)(m_long_pool.size() / 4096);
        m_long_width[k] = w;
        m_long_pool.insert(m_long_pool.end(), vals.begin(), vals.end());
      } else {
        if (vals.size() != 64) throw std::runtime_error("select_support_mcl::load: mini block size");
        m_mini_width[k] = w;
        for (int t = 0; t < 64; t++) m_mini[k * 64 + t] = (uint32_t)vals[t];
      }
    }
  }

  void set_vector(const bit_vector* v = nullptr) { m_v = v; }
  void swap(select_support_mcl& ss) {
    std::swap(m_logn, ss.m_logn); std::swap(m_logn2, ss.m_logn2); std::swap(m_logn4, ss.m_logn4);
    m_superblock.swap(ss.m_superblock); m_mini.swap(ss.m_mini); m_mini_width.swap(ss.m_mini_width);
    m_long_index.swap(ss.m_long_index); m_long_width.swap(ss.m_long_width); m_long_pool.swap(ss.m_long_pool);
    std::swap(m_arg_cnt, ss.m_arg_cnt);
  }
  size_type arg_cnt() const { return m_arg_cnt; }
  uint32_t logn() const { return m_logn; }
  uint32_t logn4() const { return m_logn4; }
  const std::vector<uint64_t>& superblock() const { return m_superblock; }
  const std::vector<uint32_t>& mini() const { return m_mini; }
  const std::vector<uint8_t>& mini_width() const { return m_mini_width; }
  const std::vector<int64_t>& long_index() const { return m_long_index; }
  const std::vector<uint8_t>& long_width() const { return m_long_width; }
  const std::vector<uint64_t>& long_pool() const { return m_long_pool; }
};

}  // namespace sdsl
#endif
