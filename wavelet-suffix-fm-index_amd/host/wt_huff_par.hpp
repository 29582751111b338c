// wt_huff_par.hpp -- drop-in for sdsl::wt_huff<> of the reference
// (wt_huff/wt_huff_par.hpp:68-78 over wt_huff/wt_pc_par.hpp): same include
// guard, the constructor wt_huff(int_vector<8>& text, size) (wt_pc_par.hpp:321-361)
// and the query surface size / empty / sigma / operator[] / rank /
// inverse_select / select / interval_symbols, begin / end, the node interface
// (root / is_leaf / sym / empty / expand / path) and symbol_gte / symbol_lte.  Counts, Huffman shape, tree bits, rank_support_v
// and select_support_mcl (1 and 0) are built on the GPU (wtg_huff_build) and
// copied into host supports, so scalar queries run on the host with the
// reference's results; rank_batch / select_batch / access_batch / count_batch
// answer device-resident batches.  select() throws std::logic_error when c does
// not occur i times (the reference asserts).
#ifndef INCLUDED_SDSL_WT_HUFF
#define INCLUDED_SDSL_WT_HUFF
#include <algorithm>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include "rank_support_v_par.hpp"
#include "select_support_mcl_par.hpp"
#include "sdsl_standin.hpp"
#include "wtg_runtime.hpp"

namespace sdsl {

struct byte_tree_standin {};

template <class t_bitvector = bit_vector, class t_rank = rank_support_v<1, 1>,
          class t_select = select_support_mcl<1, 1>, class t_select_zero = select_support_mcl<0, 1>,
          class t_tree_strat = byte_tree_standin>
class wt_huff {
 public:
  typedef uint64_t size_type;
  typedef uint8_t value_type;
  typedef t_bitvector bit_vector_type;
  typedef t_rank rank_1_type;
  typedef t_select select_1_type;
  typedef t_select_zero select_0_type;
  enum { lex_ordered = 0 };

 protected:
  size_type m_size = 0;
  size_type m_sigma = 0;
  bit_vector_type m_bv;
  rank_1_type m_bv_rank;
  select_1_type m_bv_select1;
  select_0_type m_bv_select0;
  wtg_huff_shape m_shape{};
  wtg_huff* m_dev = nullptr;

 public:
  const size_type& sigma = m_sigma;

  wt_huff() = default;

  template <uint8_t int_width>
  wt_huff(int_vector<int_width>& text, size_type size) : m_size(size) {  // wt_pc_par.hpp:321-361
    if (0 == m_size) return;
    if (text.size() < m_size) throw std::logic_error("Stream size is smaller than size!");
    std::vector<unsigned char> bytes(m_size);
    for (size_type i = 0; i < m_size; i++) bytes[i] = (unsigned char)text[i];
    build(bytes.data());
  }
  // host byte array (the FMIndex() drop-in uses this)
  wt_huff(const unsigned char* text, size_type size) : m_size(size) {
    if (m_size) build(text);
  }
  wt_huff(const wt_huff& o) { copy(o); }
  wt_huff(wt_huff&& o) noexcept { take(o); }
  wt_huff& operator=(const wt_huff& o) {
    if (this != &o) { release(); copy(o); }
    return *this;
  }
  wt_huff& operator=(wt_huff&& o) noexcept {
    if (this != &o) { release(); take(o); }
    return *this;
  }
  ~wt_huff() { release(); }

  size_type size() const { return m_size; }
  bool empty() const { return m_size == 0; }
  const wtg_huff_shape& shape() const { return m_shape; }
  const bit_vector_type& bits() const { return m_bv; }

  value_type operator[](size_type i) const { return inverse_select(i).second; }  // :430-448

  size_type rank(size_type i, value_type c) const {  // :462-485
    if (m_shape.c_to_leaf[c] < 0) return 0;
    if (m_sigma == 1) return i;
    uint64_t p = m_shape.path[c];
    int v = 0;
    for (uint32_t l = 0; l < m_shape.path_len[c] && i; ++l, p >>= 1) {
      const size_type r = m_bv_rank(m_shape.bv_pos[v] + i) - m_shape.bv_pos_rank[v];
      i = (p & 1) ? r : i - r;
      v = m_shape.child[v][p & 1];
    }
    return i;
  }

  std::pair<size_type, value_type> inverse_select(size_type i) const {  // :498-516
    if (m_sigma == 1) return std::make_pair(i, (value_type)m_shape.sym[0]);
    int v = 0;
    while (m_shape.child[v][0] >= 0) {
      const size_type r = m_bv_rank(m_shape.bv_pos[v] + i) - m_shape.bv_pos_rank[v];
      const int b = m_bv[m_shape.bv_pos[v] + i];
      i = b ? r : i - r;
      v = m_shape.child[v][b];
    }
    return std::make_pair(i, (value_type)m_shape.sym[v]);
  }

  // the k distinct symbols of [i, j) with rank(i, c) and rank(j, c): in tree order
  // (left child first) from three positions on, in text order for two, as the
  // reference (cs / rank_c_i / rank_c_j hold at least sigma entries)
  void interval_symbols(size_type i, size_type j, size_type& k, std::vector<value_type>& cs,
                        std::vector<size_type>& rank_c_i, std::vector<size_type>& rank_c_j) const {  // :579-617
    k = 0;
    if (i == j) return;
    if (m_sigma == 1) {
      k = 1;
      cs[0] = (value_type)m_shape.sym[0];
      rank_c_i[0] = std::min(i, m_size);
      rank_c_j[0] = std::min(j, m_size);
    } else if (j - i == 1) {
      k = 1;
      const auto r = inverse_select(i);
      rank_c_i[0] = r.first;
      cs[0] = r.second;
      rank_c_j[0] = r.first + 1;
    } else if (j - i == 2) {
      const auto r0 = inverse_select(i), r1 = inverse_select(i + 1);
      rank_c_i[0] = r0.first; cs[0] = r0.second;
      rank_c_i[1] = r1.first; cs[1] = r1.second;
      if (cs[0] == cs[1]) {
        k = 1;
        rank_c_j[0] = rank_c_i[0] + 2;
      } else {
        k = 2;
        rank_c_j[0] = rank_c_i[0] + 1;
        rank_c_j[1] = rank_c_i[1] + 1;
      }
    } else {
      interval_walk(i, j, k, cs, rank_c_i, rank_c_j, 0);
    }
  }

  size_type select(size_type i, value_type c) const {  // :527-554
    const int leaf = m_shape.c_to_leaf[c];
    if (leaf < 0 || i == 0 || i > m_shape.C[c + 1] - m_shape.C[c])
      throw std::logic_error("select(" + std::to_string(i) + "," + std::to_string((int)c) +
                             "): c does not occur i times");
    if (m_sigma == 1) return i - 1;
    size_type res = i - 1;
    int v = leaf;
    for (uint32_t l = 0; l < m_shape.path_len[c]; ++l) {
      const int p = m_shape.parent[v];
      const size_type bp = m_shape.bv_pos[p], br = m_shape.bv_pos_rank[p];
      if (m_shape.child[p][1] == v) res = m_bv_select1(br + res + 1) - bp;
      else res = m_bv_select0(bp - br + res + 1) - bp;
      v = p;
    }
    return res;
  }

  // one inner node of interval_symbols (:149-183): [i, j) of node v, j > i
  void interval_walk(size_type i, size_type j, size_type& k, std::vector<value_type>& cs,
                     std::vector<size_type>& rank_c_i, std::vector<size_type>& rank_c_j, int v) const {
    const size_type bp = m_shape.bv_pos[v], br = m_shape.bv_pos_rank[v];
    const size_type i1 = m_bv_rank(bp + i) - br, j1 = m_bv_rank(bp + j) - br;
    const size_type i0 = i - i1, j0 = j - j1;
    const size_type lo[2] = {i0, i1}, hi[2] = {j0, j1};
    for (int b = 0; b < 2; b++) {
      if (lo[b] == hi[b]) continue;
      const int c = m_shape.child[v][b];
      if (m_shape.child[c][0] >= 0) {
        interval_walk(lo[b], hi[b], k, cs, rank_c_i, rank_c_j, c);
      } else {
        rank_c_i[k] = lo[b];
        rank_c_j[k] = hi[b];
        cs[k++] = (value_type)m_shape.sym[c];
      }
    }
  }

  // ---- iterators (wt_pc_par.hpp:740-746): over operator[]
  typedef random_access_const_iterator<wt_huff> const_iterator;
  typedef const_iterator iterator;
  const_iterator begin() const { return const_iterator(this, 0); }
  const_iterator end() const { return const_iterator(this, size()); }

  // ---- node interface (wt_pc_par.hpp:778-909).  A node is its index in the
  // tree's BFS order (root 0); a leaf holds one symbol.
  typedef uint64_t node_type;
  bool is_leaf(const node_type& v) const { return m_shape.child[v][0] < 0; }
  value_type sym(const node_type& v) const { return (value_type)m_shape.sym[v]; }
  // the reference answers true for every node (:786-788)
  bool empty(const node_type&) const { return true; }
  node_type root() const { return 0; }
  std::pair<node_type, node_type> expand(const node_type& v) const {  // pre: !is_leaf(v)
    return std::make_pair((node_type)m_shape.child[v][0], (node_type)m_shape.child[v][1]);
  }
  // a closed range [s, e] of v mapped into both children (an empty result is
  // [x, x-1] in size_type arithmetic, as the reference)
  std::pair<range_type, range_type> expand(const node_type& v, const range_type& r) const {
    const size_type bp = m_shape.bv_pos[v], v_sp_rank = m_shape.bv_pos_rank[v];
    const size_type sp_rank = m_bv_rank(bp + r[0]);
    const size_type right_size = m_bv_rank(bp + r[1] + 1) - sp_rank;
    const size_type left_size = (r[1] - r[0] + 1) - right_size;
    const size_type right_sp = sp_rank - v_sp_rank, left_sp = r[0] - right_sp;
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
  // {code length, code} of symbol c, the root's branch as the code's highest
  // bit (:864-871)
  std::pair<uint64_t, uint64_t> path(value_type c) const {
    const uint32_t len = m_shape.c_to_leaf[c] >= 0 ? m_shape.path_len[c] : 0u;
    uint64_t code = 0;
    for (uint32_t l = 0; l < len; l++) code = (code << 1) | ((m_shape.path[c] >> l) & 1u);
    return {len, code};
  }
  // the smallest symbol >= c / the largest symbol <= c that occurs; first =
  // false when there is none (:880-899)
  std::pair<bool, value_type> symbol_gte(value_type c) const {
    for (uint32_t x = c; x < 256; x++)
      if (m_size && m_shape.c_to_leaf[x] >= 0) return {true, (value_type)x};
    return {false, 0};
  }
  std::pair<bool, value_type> symbol_lte(value_type c) const {
    for (int x = c; x >= 0; x--)
      if (m_size && m_shape.c_to_leaf[x] >= 0) return {true, (value_type)x};
    return {false, 0};
  }

  // ---- batched device queries (device buffers of nq entries)
  void rank_batch(const uint64_t* d_i, const uint32_t* d_c, uint64_t nq, uint64_t* d_out) const {
    wtg::check(wtg_huff_rank_batch(wtg::default_context(), m_dev, d_i, d_c, nq, d_out));
  }
  void select_batch(const uint64_t* d_i, const uint32_t* d_c, uint64_t nq, uint64_t* d_out) const {
    wtg::check(wtg_huff_select_batch(wtg::default_context(), m_dev, d_i, d_c, nq, d_out));
  }
  void access_batch(const uint64_t* d_i, uint64_t nq, uint64_t* d_sym) const {
    wtg::check(wtg_huff_access_batch(wtg::default_context(), m_dev, d_i, nq, d_sym, nullptr));
  }
  // FM-index backward-search count when this tree was built over a BWT
  void count_batch(const uint8_t* d_pats, uint32_t m, uint64_t npat, uint64_t* d_out) const {
    wtg::check(wtg_fm_count_batch(wtg::default_context(), m_dev, d_pats, m, npat, d_out));
  }
  const wtg_huff* device() const { return m_dev; }

  // The reference's stream layout (wt_pc_par.hpp:750-764): size and sigma
  // (uint64), the tree bit_vector, rank_support_v, select_support_mcl<1>, <0>,
  // then the tree: the node count (uint64); per node in BFS order bv_pos and
  // bv_pos_rank (uint64), parent, child[0], child[1] (uint16, 0xFFFF = none);
  // the leaf of each byte (256 x uint16, 0xFFFF = absent); each byte's code
  // (256 x uint64: branch bits from the root in bits 0..55, length in 56..63).
  size_type serialize(std::ostream& out, structure_tree_node* = nullptr, std::string = "") const {
    using namespace wtg::sdslio;
    size_type n = write_member((uint64_t)m_size, out) + write_member((uint64_t)m_sigma, out);
    n += write_bits(out, m_bv.data(), m_bv.size());
    n += m_bv_rank.serialize(out);
    n += m_bv_select1.serialize(out);
    n += m_bv_select0.serialize(out);
    const uint64_t nn = m_size ? m_shape.n_nodes : 0;
    n += write_member(nn, out);
    auto u16 = [](int32_t x) { return (uint16_t)(x < 0 ? 0xFFFF : x); };
    for (uint64_t v = 0; v < nn; v++) {
      const uint64_t pos = m_shape.bv_pos[v];
      const uint64_t prk = m_bv.size() ? m_bv_rank(pos) : 0;
      n += write_member(pos, out) + write_member(prk, out);
      n += write_member(u16(m_shape.parent[v]), out) + write_member(u16(m_shape.child[v][0]), out) +
           write_member(u16(m_shape.child[v][1]), out);
    }
    for (int c = 0; c < 256; c++) n += write_member(u16(m_size ? m_shape.c_to_leaf[c] : -1), out);
    for (int c = 0; c < 256; c++) {
      const uint64_t len = m_size && m_shape.c_to_leaf[c] >= 0 ? m_shape.path_len[c] : 0;
      n += write_member((uint64_t)((len ? m_shape.path[c] : 0) | (len << 56)), out);
    }
    return n;
  }
  // wt_pc_par.hpp:766-775.  The device copy for the batched queries is rebuilt
  // from the byte counts (the frequencies of the leaves) and the bits; a tree
  // that is not the Huffman tree of its counts is rejected.
  void load(std::istream& in) {
    using namespace wtg::sdslio;
    release();
    uint64_t size = 0, sigma = 0, nbits = 0, nn = 0;
    read_member(size, in);
    read_member(sigma, in);
    std::vector<uint64_t> words;
    read_bits(in, words, nbits);
    m_size = size;
    m_sigma = sigma;
    m_bv = bit_vector_type(nbits);
    if (nbits) std::copy(words.begin(), words.begin() + (nbits + 63) / 64, m_bv.data());
    m_bv_rank.load(in, &m_bv);
    m_bv_select1.load(in, &m_bv);
    m_bv_select0.load(in, &m_bv);
    read_member(nn, in);
    if (nn > WTG_HUFF_MAX_NODES) throw std::runtime_error("wt_huff::load: more than 511 tree nodes");
    m_shape = wtg_huff_shape{};
    m_shape.n = m_size;
    m_shape.bv_size = nbits;
    m_shape.n_nodes = (uint32_t)nn;
    auto i32 = [](uint16_t x) { return x == 0xFFFF ? -1 : (int32_t)x; };
    for (uint64_t v = 0; v < nn; v++) {
      uint64_t pos = 0, prk = 0;
      uint16_t p = 0, c0 = 0, c1 = 0;
      read_member(pos, in); read_member(prk, in);
      read_member(p, in); read_member(c0, in); read_member(c1, in);
      m_shape.bv_pos[v] = pos;
      m_shape.bv_pos_rank[v] = prk;
      m_shape.parent[v] = i32(p);
      m_shape.child[v][0] = i32(c0);
      m_shape.child[v][1] = i32(c1);
      m_shape.sym[v] = -1;
      if (c0 != 0xFFFF) m_shape.n_inner++;
    }
    for (int c = 0; c < 256; c++) {
      uint16_t x = 0;
      read_member(x, in);
      m_shape.c_to_leaf[c] = i32(x);
      if (x != 0xFFFF) {
        if (x >= nn) throw std::runtime_error("wt_huff::load: leaf index out of range");
        m_shape.sym[x] = c;
      }
    }
    for (int c = 0; c < 256; c++) {
      uint64_t p = 0;
      read_member(p, in);
      m_shape.path_len[c] = (uint32_t)(p >> 56);
      m_shape.path[c] = p & ((1ull << 56) - 1);
      m_shape.max_depth = std::max(m_shape.max_depth, m_shape.path_len[c]);
    }
    m_shape.sigma = (uint32_t)m_sigma;
    if (!m_size) return;
    // leaf frequencies: the root holds every position; an inner node's ones go
    // to child[1], its zeros to child[0] (BFS order: parents come first)
    std::vector<uint64_t> freq(nn, 0);
    if (nn == 0) throw std::runtime_error("wt_huff::load: a nonempty text without tree nodes");
    freq[0] = m_size;
    for (uint64_t v = 0; v < nn; v++) {
      const int32_t c0 = m_shape.child[v][0], c1 = m_shape.child[v][1];
      if (c0 < 0) continue;
      if (c0 >= (int32_t)nn || c1 < 0 || c1 >= (int32_t)nn || m_shape.bv_pos[v] + freq[v] > nbits)
        throw std::runtime_error("wt_huff::load: inconsistent tree");
      const uint64_t a = m_shape.bv_pos[v], ones = m_bv_rank(a + freq[v]) - m_bv_rank(a);
      freq[m_shape.child[v][1]] = ones;
      freq[m_shape.child[v][0]] = freq[v] - ones;
    }
    uint64_t counts[256] = {0};
    for (int c = 0; c < 256; c++)
      if (m_shape.c_to_leaf[c] >= 0) counts[c] = freq[m_shape.c_to_leaf[c]];
    m_shape.C[0] = 0;
    for (int c = 0; c < 256; c++) m_shape.C[c + 1] = m_shape.C[c] + counts[c];
    wtg::dbuf dw(words.size() * 8);
    dw.upload(words.data(), words.size() * 8);
    wtg::check(wtg_huff_from_bits(wtg::default_context(), counts, dw.as<uint64_t>(), nbits, &m_dev));
    wtg_huff_shape ds{};
    wtg::check(wtg_huff_get_shape(m_dev, &ds));
    bool same = ds.n_nodes == m_shape.n_nodes && ds.bv_size == m_shape.bv_size;
    for (uint32_t v = 0; same && v < ds.n_nodes; v++)
      same = ds.child[v][0] == m_shape.child[v][0] && ds.child[v][1] == m_shape.child[v][1] &&
             ds.sym[v] == m_shape.sym[v] && (ds.child[v][0] < 0 || ds.bv_pos[v] == m_shape.bv_pos[v]);
    if (!same) {
      release();
      throw std::runtime_error("wt_huff::load: the stored tree is not the Huffman tree of its symbol counts");
    }
    m_shape = ds;
  }

 private:
  void build(const unsigned char* host_text) {
    wtg::dbuf dt(m_size);
    dt.upload(host_text, m_size);
    wtg::check(wtg_huff_build(wtg::default_context(), (const uint8_t*)dt.get(), m_size, &m_dev));
    wtg::check(wtg_huff_get_shape(m_dev, &m_shape));
    m_sigma = m_shape.sigma;
    const uint64_t nbits = m_shape.bv_size;
    const uint64_t *dw = nullptr, *dd = nullptr;
    wtg::check(wtg_huff_data(m_dev, &dw, &dd));
    m_bv = bit_vector_type(nbits);
    if (nbits) wtg::check(wtg_memcpy_d2h(wtg::default_context(), m_bv.data(), dw, (nbits + 63) / 64 * 8));
    m_bv_rank = rank_1_type();
    m_bv_rank.assign_directory(dd, nbits);
    m_bv_rank.set_vector(&m_bv);
    m_bv_select1.assign(*wtg_huff_select1(m_dev));
    m_bv_select1.set_vector(&m_bv);
    m_bv_select0.assign(*wtg_huff_select0(m_dev));
    m_bv_select0.set_vector(&m_bv);
  }
  void release() {
    if (m_dev) wtg_huff_destroy(wtg::default_context(), m_dev);
    m_dev = nullptr;
  }
  void take(wt_huff& o) {  // move: the device copy changes owner
    copy(o);
    m_dev = o.m_dev;
    o.m_dev = nullptr;
  }
  void copy(const wt_huff& o) {
    m_size = o.m_size;
    m_sigma = o.m_sigma;
    m_shape = o.m_shape;
    m_bv = o.m_bv;
    m_bv_rank = o.m_bv_rank;
    m_bv_rank.set_vector(&m_bv);
    m_bv_select1 = o.m_bv_select1;
    m_bv_select1.set_vector(&m_bv);
    m_bv_select0 = o.m_bv_select0;
    m_bv_select0.set_vector(&m_bv);
    m_dev = nullptr;  // the copy answers host queries; batches need the original
  }
};

}  // namespace sdsl
#endif
