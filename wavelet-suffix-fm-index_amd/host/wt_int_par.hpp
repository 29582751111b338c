// wt_int_par.hpp -- drop-in for sdsl::wt_int<> of the reference
// (wt_int/wt_int_par.hpp:57-1086): same include guard, constructor
// wt_int(int_vector<w>& buf, size, max_level=0) and the query surface size /
// empty / sigma / tree / max_level / operator[] / rank / inverse_select /
// select / interval_symbols / lex_count / lex_smaller_count / range_search_2d and
// the node interface (root / is_leaf / sym / empty / expand / path).  The tree bits, rank_support_v and select_support_mcl (for 1 and 0)
// are built on the GPU (wtg_wtint_build) and copied into host supports, so the
// scalar queries run the reference's host algorithms; *_batch() run batches on
// the device against the device copy.  select() throws std::logic_error when
// c does not occur i times, as the reference (:644-646).
#ifndef INCLUDED_SDSL_INT_WAVELET_TREE_PAR
#define INCLUDED_SDSL_INT_WAVELET_TREE_PAR
#include <stdexcept>
#include <algorithm>
#include <string>
#include <tuple>
#include <utility>
#include <vector>
#include "rank_support_v_par.hpp"
#include "select_support_mcl_par.hpp"
#include "sdsl_standin.hpp"
#include "wtg_runtime.hpp"

namespace sdsl {

template <class t_bitvector = bit_vector, class t_rank = rank_support_v<1, 1>,
          class t_select = select_support_mcl<1, 1>, class t_select_zero = select_support_mcl<0, 1>>
class wt_int {
 public:
  typedef uint64_t size_type;
  typedef uint64_t value_type;
  typedef t_bitvector bit_vector_type;
  typedef t_rank rank_1_type;
  typedef t_select select_1_type;
  typedef t_select_zero select_0_type;
  enum { lex_ordered = 1 };

  bit_vector_type m_tree;

 protected:
  size_type m_size = 0;
  size_type m_sigma = 0;
  rank_1_type m_tree_rank;
  select_1_type m_tree_select1;
  select_0_type m_tree_select0;
  uint32_t m_max_level = 0;
  wtg_wtint* m_dev = nullptr;  // device copy for the batched queries

 public:
  const size_type& sigma = m_sigma;
  const bit_vector_type& tree = m_tree;
  const uint32_t& max_level = m_max_level;

  wt_int() = default;

  template <uint8_t int_width>
  wt_int(int_vector<int_width>& buf, size_type size, uint32_t max_level = 0) : m_size(size) {  // :368-406
    if (0 == m_size) return;
    if (buf.size() < m_size)
      throw std::logic_error("n=" + std::to_string(buf.size()) + " < " + std::to_string(m_size) + "=m_size");
    // the reference requires whole-word symbols (:371); copy them packed
    const int w = buf.width() <= 8 ? 1 : buf.width() <= 16 ? 2 : buf.width() <= 32 ? 4 : 8;
    std::vector<unsigned char> packed(m_size * w);
    for (size_type i = 0; i < m_size; i++) {
      const uint64_t v = buf[i];
      for (int b = 0; b < w; b++) packed[i * w + b] = (unsigned char)(v >> (8 * b));
    }
    wtg::dbuf ds(m_size * w);
    ds.upload(packed.data(), m_size * w);
    wtg::check(wtg_wtint_build(wtg::default_context(), ds.get(), w, m_size, max_level, &m_dev));
    fetch();
  }
  wt_int(const wt_int& o) { copy(o); }
  wt_int& operator=(const wt_int& o) {
    if (this != &o) copy(o);
    return *this;
  }
  ~wt_int() {
    if (m_dev) wtg_wtint_destroy(wtg::default_context(), m_dev);
  }

  size_type size() const { return m_size; }
  bool empty() const { return m_size == 0; }

  value_type operator[](size_type i) const {  // :519-541
    size_type offset = 0, node_size = m_size;
    value_type res = 0;
    for (uint32_t k = 0; k < m_max_level; ++k) {
      res <<= 1;
      const size_type o = m_tree_rank(offset), oi = m_tree_rank(offset + i) - o,
                      oe = m_tree_rank(offset + node_size) - o;
      if (m_tree[offset + i]) { offset += node_size - oe; node_size = oe; i = oi; res |= 1; }
      else { node_size = node_size - oe; i = i - oi; }
      offset += m_size;
    }
    return res;
  }

  size_type rank(size_type i, value_type c) const {  // :553-577
    if (((1ULL) << m_max_level) <= c) return 0;
    size_type offset = 0, node_size = m_size;
    uint64_t mask = 1ULL << (m_max_level - 1);
    for (uint32_t k = 0; k < m_max_level && i; ++k) {
      const size_type o = m_tree_rank(offset), oi = m_tree_rank(offset + i) - o,
                      oe = m_tree_rank(offset + node_size) - o;
      if (c & mask) { offset += node_size - oe; node_size = oe; i = oi; }
      else { node_size = node_size - oe; i = i - oi; }
      offset += m_size;
      mask >>= 1;
    }
    return i;
  }

  std::pair<size_type, value_type> inverse_select(size_type i) const {  // :588-611
    value_type c = 0;
    size_type node_size = m_size, offset = 0;
    for (uint32_t k = 0; k < m_max_level; ++k) {
      const size_type o = m_tree_rank(offset), oi = m_tree_rank(offset + i) - o,
                      oe = m_tree_rank(offset + node_size) - o;
      c <<= 1;
      if (m_tree[offset + i]) { offset += node_size - oe; node_size = oe; i = oi; c |= 1; }
      else { node_size = node_size - oe; i = i - oi; }
      offset += m_size;
    }
    return std::make_pair(i, c);
  }

  // the k distinct values of [i, j) in ascending order, with rank(i, c) and
  // rank(j, c) for each (cs / rank_c_i / rank_c_j hold at least sigma entries)
  void interval_symbols(size_type i, size_type j, size_type& k, std::vector<value_type>& cs,
                        std::vector<size_type>& rank_c_i, std::vector<size_type>& rank_c_j) const {  // :684-702
    k = 0;
    if (i == j) return;
    if (i + 1 == j) {
      const auto r = inverse_select(i);
      cs[0] = r.second;
      rank_c_i[0] = r.first;
      rank_c_j[0] = r.first + 1;
      k = 1;
      return;
    }
    interval_walk(i, j, k, cs, rank_c_i, rank_c_j, 0, 0, m_size, 0);
  }

  // (rank(i,c), #values < c in [i,j), #values > c in [i,j))
  std::tuple<size_type, size_type, size_type> lex_count(size_type i, size_type j, value_type c) const {  // :712-741
    if (((1ULL) << m_max_level) <= c) return std::make_tuple(size_type(0), j - i, size_type(0));
    size_type offset = 0, node_size = m_size, smaller = 0, greater = 0;
    uint64_t mask = 1ULL << (m_max_level - 1);
    for (uint32_t k = 0; k < m_max_level; ++k) {
      const size_type o = m_tree_rank(offset), oi = m_tree_rank(offset + i) - o, oj = m_tree_rank(offset + j) - o,
                      oe = m_tree_rank(offset + node_size) - o;
      if (c & mask) { offset += node_size - oe; node_size = oe; smaller += (j - i) - (oj - oi); i = oi; j = oj; }
      else { node_size -= oe; greater += oj - oi; i -= oi; j -= oj; }
      offset += m_size;
      mask >>= 1;
    }
    return std::make_tuple(i, smaller, greater);
  }

  // (rank(i,c), #values < c in [0,i))
  std::tuple<size_type, size_type> lex_smaller_count(size_type i, value_type c) const {  // :753-780
    if (((1ULL) << m_max_level) <= c) return std::make_tuple(size_type(0), i);
    size_type offset = 0, node_size = m_size, below = 0;
    uint64_t mask = 1ULL << (m_max_level - 1);
    for (uint32_t k = 0; k < m_max_level && i; ++k) {
      const size_type o = m_tree_rank(offset), oi = m_tree_rank(offset + i) - o,
                      oe = m_tree_rank(offset + node_size) - o;
      if (c & mask) { offset += node_size - oe; node_size = oe; below += i - oi; i = oi; }
      else { node_size -= oe; i -= oi; }
      offset += m_size;
      mask >>= 1;
    }
    return std::make_tuple(i, below);
  }

  // points (index, value) with index in [lb, rb] and value in [vlb, vrb]: (count,
  // the points when report) -- :801-870
  std::pair<size_type, std::vector<std::pair<value_type, size_type>>> range_search_2d(
      size_type lb, size_type rb, value_type vlb, value_type vrb, bool report = true) const {
    std::pair<size_type, std::vector<std::pair<value_type, size_type>>> res;
    res.first = 0;
    if (vrb > (1ULL << m_max_level)) vrb = 1ULL << m_max_level;
    if (vlb > vrb) return res;
    std::vector<size_type> offs(m_max_level + 1), obs(m_max_level + 1);
    offs[0] = 0;
    range2d(lb, rb, vlb, vrb, 0, 0, m_size, offs, obs, 0, report, res);
    return res;
  }

  // ---- iterators (:877-886): over operator[]
  typedef random_access_const_iterator<wt_int> const_iterator;
  typedef const_iterator iterator;
  const_iterator begin() const { return const_iterator(this, 0); }
  const_iterator end() const { return const_iterator(this, size()); }

  // ---- node interface (:915-1085): a node is (offset of its bits, size, level,
  // value prefix)
  struct node_type {
    size_type offset = 0, size = 0, level = 0;
    value_type sym = 0;
    node_type(size_type o = 0, size_type sz = 0, size_type l = 0, value_type sy = 0)
        : offset(o), size(sz), level(l), sym(sy) {}
    bool operator==(const node_type& v) const { return offset == v.offset; }
    bool operator<(const node_type& v) const { return offset < v.offset; }
    bool operator>(const node_type& v) const { return offset > v.offset; }
  };
  bool is_leaf(const node_type& v) const { return v.level == m_max_level; }
  value_type sym(const node_type& v) const { return v.sym; }
  bool empty(const node_type& v) const { return v.size == 0; }
  node_type root() const { return node_type(0, m_size, 0, 0); }
  // the two children of an inner node
  std::pair<node_type, node_type> expand(const node_type& v) const {
    const size_type ones = m_tree_rank(v.offset + v.size) - m_tree_rank(v.offset);
    const node_type left(v.offset + m_size, v.size - ones, v.level + 1, v.sym << 1);
    const node_type right(v.offset + m_size + left.size, ones, v.level + 1, (v.sym << 1) | 1);
    return std::make_pair(left, right);
  }
  // a closed range [s, e] of v mapped into both children (an empty result is
  // [x, x-1], in size_type arithmetic as the reference)
  std::pair<range_type, range_type> expand(const node_type& v, const range_type& r) const {
    const size_type base = m_tree_rank(v.offset);
    const size_type sr = m_tree_rank(v.offset + r[0]);
    const size_type right_size = m_tree_rank(v.offset + r[1] + 1) - sr;
    const size_type left_size = (r[1] - r[0] + 1) - right_size;
    const size_type right_sp = sr - base, left_sp = r[0] - right_sp;
    return std::make_pair(range_type{{left_sp, left_sp + left_size - 1}},
                          range_type{{right_sp, right_sp + right_size - 1}});
  }
  std::pair<range_vec_type, range_vec_type> expand(const node_type& v, const range_vec_type& ranges) const {
    std::pair<range_vec_type, range_vec_type> res;
    res.first.reserve(ranges.size());
    res.second.reserve(ranges.size());
    for (const auto& r : ranges) {
      const auto lr = expand(v, r);
      res.first.push_back(lr.first);
      res.second.push_back(lr.second);
    }
    return res;
  }
  std::pair<uint64_t, uint64_t> path(value_type c) const { return {m_max_level, c}; }

  size_type select(size_type i, value_type c) const {  // :622-660
    std::vector<size_type> path_off(m_max_level + 1), path_rank(m_max_level + 1);
    size_type offset = 0, node_size = m_size;
    uint64_t mask = 1ULL << (m_max_level - 1);
    path_off[0] = path_rank[0] = 0;
    for (uint32_t k = 0; k < m_max_level && node_size; ++k) {
      const size_type o = m_tree_rank(offset);
      path_rank[k] = o;
      const size_type oe = m_tree_rank(offset + node_size) - o;
      if (c & mask) { offset += node_size - oe; node_size = oe; }
      else node_size = node_size - oe;
      offset += m_size;
      path_off[k + 1] = offset;
      mask >>= 1;
    }
    if (0ULL == node_size || node_size < i)
      throw std::logic_error("select(" + std::to_string(i) + "," + std::to_string(c) +
                             "): c does not occur i times in the WT");
    mask = 1ULL;
    for (uint32_t k = m_max_level; k > 0; --k) {
      offset = path_off[k - 1];
      const size_type o = path_rank[k - 1];
      if (c & mask) i = m_tree_select1(o + i) - offset + 1;
      else i = m_tree_select0(offset - o + i) - offset + 1;
      mask <<= 1;
    }
    return i - 1;
  }

  // ---- batched device queries (device buffers of nq uint64)
  void rank_batch(const uint64_t* d_i, const uint64_t* d_c, uint64_t nq, uint64_t* d_out) const {
    wtg::check(wtg_wtint_rank_batch(wtg::default_context(), m_dev, d_i, d_c, nq, d_out));
  }
  void select_batch(const uint64_t* d_i, const uint64_t* d_c, uint64_t nq, uint64_t* d_out) const {
    wtg::check(wtg_wtint_select_batch(wtg::default_context(), m_dev, d_i, d_c, nq, d_out));
  }
  void access_batch(const uint64_t* d_i, uint64_t nq, uint64_t* d_out) const {
    wtg::check(wtg_wtint_access_batch(wtg::default_context(), m_dev, d_i, nq, d_out));
  }
  void lex_count_batch(const uint64_t* d_i, const uint64_t* d_j, const uint64_t* d_c, uint64_t nq,
                       uint64_t* d_rank, uint64_t* d_smaller, uint64_t* d_greater) const {
    wtg::check(wtg_wtint_lex_count_batch(wtg::default_context(), m_dev, d_i, d_j, d_c, nq, d_rank, d_smaller,
                                         d_greater));
  }
  void lex_smaller_count_batch(const uint64_t* d_i, const uint64_t* d_c, uint64_t nq, uint64_t* d_rank,
                               uint64_t* d_smaller) const {
    wtg::check(wtg_wtint_lex_smaller_count_batch(wtg::default_context(), m_dev, d_i, d_c, nq, d_rank, d_smaller));
  }
  void range_count_batch(const uint64_t* d_lb, const uint64_t* d_rb, const uint64_t* d_vlb, const uint64_t* d_vrb,
                         uint64_t nq, uint64_t* d_out) const {
    wtg::check(wtg_wtint_range_count_batch(wtg::default_context(), m_dev, d_lb, d_rb, d_vlb, d_vrb, nq, d_out));
  }
  const wtg_wtint* device() const { return m_dev; }

  // The reference's stream layout (:886-899): size and sigma (uint64), the tree
  // bit_vector, rank_support_v, select_support_mcl<1>, <0>, max_level (uint32).
  size_type serialize(std::ostream& out, structure_tree_node* = nullptr, std::string = "") const {
    using namespace wtg::sdslio;
    size_type n = write_member((uint64_t)m_size, out) + write_member((uint64_t)m_sigma, out);
    n += write_bits(out, m_tree.data(), m_tree.size());
    n += m_tree_rank.serialize(out);
    n += m_tree_select1.serialize(out);
    n += m_tree_select0.serialize(out);
    n += write_member(m_max_level, out);
    return n;
  }
  // :901-911; the device copy for the batched queries is rebuilt from the tree
  void load(std::istream& in) {
    using namespace wtg::sdslio;
    if (m_dev) wtg_wtint_destroy(wtg::default_context(), m_dev);
    m_dev = nullptr;
    uint64_t size = 0, sigma = 0, nbits = 0;
    read_member(size, in);
    read_member(sigma, in);
    std::vector<uint64_t> words;
    read_bits(in, words, nbits);
    m_size = size;
    m_sigma = sigma;
    m_tree = bit_vector_type(nbits);
    if (nbits) std::copy(words.begin(), words.begin() + (nbits + 63) / 64, m_tree.data());
    m_tree_rank.load(in, &m_tree);
    m_tree_select1.load(in, &m_tree);
    m_tree_select0.load(in, &m_tree);
    read_member(m_max_level, in);
    if (m_size && m_max_level) {
      if (nbits != m_size * m_max_level) throw std::runtime_error("wt_int::load: tree size != size * max_level");
      wtg::dbuf dw(words.size() * 8);
      dw.upload(words.data(), words.size() * 8);
      wtg::check(wtg_wtint_from_tree(wtg::default_context(), dw.as<uint64_t>(), m_size, m_max_level, m_sigma,
                                     &m_dev));
    }
  }

 private:
  // one node of interval_symbols (:119-158): [i, j) of the node at (level, path),
  // left child before right child so the values come out ascending
  void interval_walk(size_type i, size_type j, size_type& k, std::vector<value_type>& cs,
                     std::vector<size_type>& rank_c_i, std::vector<size_type>& rank_c_j, uint32_t level,
                     value_type path, size_type node_size, size_type offset) const {
    if (level >= m_max_level) {
      rank_c_i[k] = i;
      rank_c_j[k] = j;
      cs[k++] = path;
      return;
    }
    const size_type o = m_tree_rank(offset), oi = m_tree_rank(offset + i) - o, oj = m_tree_rank(offset + j) - o,
                    oe = m_tree_rank(offset + node_size) - o;
    if ((j - i) - (oj - oi) > 0)
      interval_walk(i - oi, j - oj, k, cs, rank_c_i, rank_c_j, level + 1, path << 1, node_size - oe, offset + m_size);
    if (oj - oi > 0)
      interval_walk(oi, oj, k, cs, rank_c_i, rank_c_j, level + 1, (path << 1) | 1, oe,
                    offset + (node_size - oe) + m_size);
  }

  // one node of range_search_2d: [lb, rb] inside the node at (level, value range
  // starting at ilb); offs / obs: level offsets and ones before them on the path
  void range2d(size_type lb, size_type rb, value_type vlb, value_type vrb, uint32_t level, size_type ilb,
               size_type node_size, std::vector<size_type>& offs, std::vector<size_type>& obs, size_type path,
               bool report, std::pair<size_type, std::vector<std::pair<value_type, size_type>>>& res) const {
    if (lb > rb) return;
    if (level == m_max_level) {
      if (report) {
        for (size_type p = lb + 1; p <= rb + 1; ++p) {  // back to the text position by select up the path
          size_type i = p;
          value_type c = path;
          for (uint32_t k = m_max_level; k > 0; --k) {
            const size_type off = offs[k - 1], ob = obs[k - 1];
            i = (c & 1) ? m_tree_select1(ob + i) - off + 1 : m_tree_select0(off - ob + i) - off + 1;
            c >>= 1;
          }
          res.second.emplace_back(i - 1, path);
        }
      }
      res.first += rb - lb + 1;
      return;
    }
    const size_type mid = (2 * ilb + (1ULL << (m_max_level - level))) >> 1;
    const size_type off = offs[level];
    const size_type o = m_tree_rank(off), olb = m_tree_rank(off + lb), orb = m_tree_rank(off + rb + 1),
                    oe = m_tree_rank(off + node_size);
    obs[level] = o;
    const size_type zo = off - o, zlb = off + lb - olb, zrb = off + rb + 1 - orb, ze = off + node_size - oe;
    if (vlb < mid && mid) {
      offs[level + 1] = off + m_size;
      if (zrb - zo)
        range2d(zlb - zo, zrb - zo - 1, vlb, std::min<value_type>(vrb, mid - 1), level + 1, ilb, ze - zo, offs, obs,
                path << 1, report, res);
    }
    if (vrb >= mid) {
      offs[level + 1] = off + m_size + (ze - zo);
      if (orb - o)
        range2d(olb - o, orb - o - 1, std::max<value_type>(mid, vlb), vrb, level + 1, mid, oe - o, offs, obs,
                (path << 1) | 1, report, res);
    }
  }

  void fetch() {
    uint64_t n = 0, sg = 0;
    uint32_t L = 0;
    const uint64_t *dw = nullptr, *dd = nullptr;
    wtg::check(wtg_wtint_info(m_dev, &n, &L, &sg, &dw, &dd));
    m_max_level = L;
    m_sigma = sg;
    const uint64_t nbits = n * L;
    m_tree = bit_vector_type(nbits);
    wtg::check(wtg_memcpy_d2h(wtg::default_context(), m_tree.data(), dw, (nbits + 63) / 64 * 8));
    m_tree_rank = rank_1_type();
    m_tree_rank.assign_directory(dd, nbits);
    m_tree_rank.set_vector(&m_tree);
    m_tree_select1.assign(*wtg_wtint_select1(m_dev));
    m_tree_select1.set_vector(&m_tree);
    m_tree_select0.assign(*wtg_wtint_select0(m_dev));
    m_tree_select0.set_vector(&m_tree);
  }
  void copy(const wt_int& o) {
    m_size = o.m_size;
    m_sigma = o.m_sigma;
    m_tree = o.m_tree;
    m_max_level = o.m_max_level;
    m_tree_rank = o.m_tree_rank;
    m_tree_rank.set_vector(&m_tree);
    m_tree_select1 = o.m_tree_select1;
    m_tree_select1.set_vector(&m_tree);
    m_tree_select0 = o.m_tree_select0;
    m_tree_select0.set_vector(&m_tree);
    m_dev = nullptr;  // the copy answers host queries; batches need the original
  }
};

}  // namespace sdsl
#endif
