// sdsl_standin.hpp -- the few sdsl-lite container types the drop-in classes
// need, for builds without sdsl-lite (it is not available here).  When a real
// sdsl header is included first (its include guard INCLUDED_SDSL_INT_VECTOR),
// nothing is defined and the classes use sdsl's own bit_vector / int_vector.
//
// Layout contract shared with sdsl: bit i of a bit_vector is bit i%64 of word
// i/64; capacity() = 64*ceil(size/64); one zero padding word follows the data.
#pragma once
#ifndef INCLUDED_SDSL_INT_VECTOR
#include <array>
#include <cstdint>
#include <cstddef>
#include <istream>
#include <iterator>
#include <ostream>
#include <vector>

namespace sdsl {

typedef uint64_t size_type_standin;
struct structure_tree_node;

template <uint8_t t_width = 0>
class int_vector {
 public:
  typedef uint64_t size_type;
  typedef uint64_t value_type;
  int_vector() = default;
  explicit int_vector(size_type n, value_type v = 0, uint8_t w = (t_width ? t_width : 64))
      : width_(t_width ? t_width : w), v_(n, v) {}
  size_type size() const { return v_.size(); }
  bool empty() const { return v_.empty(); }
  uint8_t width() const { return width_; }
  value_type operator[](size_type i) const { return v_[i]; }
  value_type& operator[](size_type i) { return v_[i]; }
  const value_type* data() const { return v_.data(); }
  void resize(size_type n) { v_.resize(n); }

 private:
  uint8_t width_ = t_width ? t_width : 64;
  std::vector<value_type> v_;
};

template <>
class int_vector<1> {
 public:
  typedef uint64_t size_type;
  typedef uint64_t value_type;
  int_vector() : w_(1, 0) {}
  explicit int_vector(size_type n, value_type v = 0) : n_(n), w_((n + 63) / 64 + 1, v ? ~0ull : 0ull) {
    w_.back() = 0;
    if (v && (n & 63)) w_[n / 64] &= (1ull << (n & 63)) - 1;
  }
  size_type size() const { return n_; }
  size_type capacity() const { return (n_ + 63) / 64 * 64; }
  bool empty() const { return n_ == 0; }
  bool operator[](size_type i) const { return (w_[i >> 6] >> (i & 63)) & 1; }
  void set(size_type i, bool b) {
    if (b) w_[i >> 6] |= 1ull << (i & 63);
    else w_[i >> 6] &= ~(1ull << (i & 63));
  }
  const uint64_t* data() const { return w_.data(); }
  uint64_t* data() { return w_.data(); }
  size_type word_count() const { return (n_ + 63) / 64; }

 private:
  size_type n_ = 0;
  std::vector<uint64_t> w_;  // (n+63)/64 words + one zero padding word
};

typedef int_vector<1> bit_vector;
// closed index ranges [first, second] of the node interface (sdsl's wt_helper)
typedef std::array<uint64_t, 2> range_type;
typedef std::vector<range_type> range_vec_type;

// Read-only random-access iterator over any container with size() and a const
// operator[] (sdsl's random_access_const_iterator, which the wavelet trees'
// begin() / end() return): an index into the container, dereferenced by
// operator[].
template <class t_rac>
class random_access_const_iterator {
 public:
  typedef std::random_access_iterator_tag iterator_category;
  typedef typename t_rac::value_type value_type;
  typedef value_type const_reference;
  typedef value_type reference;
  typedef const value_type* pointer;
  typedef typename t_rac::size_type size_type;
  typedef std::ptrdiff_t difference_type;

  random_access_const_iterator() = default;
  random_access_const_iterator(const t_rac* rac, size_type idx) : m_rac(rac), m_idx(idx) {}
  const_reference operator*() const { return (*m_rac)[m_idx]; }
  const_reference operator[](difference_type i) const { return (*m_rac)[m_idx + i]; }
  random_access_const_iterator& operator++() { ++m_idx; return *this; }
  random_access_const_iterator operator++(int) { auto t = *this; ++m_idx; return t; }
  random_access_const_iterator& operator--() { --m_idx; return *this; }
  random_access_const_iterator operator--(int) { auto t = *this; --m_idx; return t; }
  random_access_const_iterator& operator+=(difference_type i) { m_idx += i; return *this; }
  random_access_const_iterator& operator-=(difference_type i) { m_idx -= i; return *this; }
  random_access_const_iterator operator+(difference_type i) const { auto t = *this; return t += i; }
  random_access_const_iterator operator-(difference_type i) const { auto t = *this; return t -= i; }
  difference_type operator-(const random_access_const_iterator& o) const {
    return (difference_type)m_idx - (difference_type)o.m_idx;
  }
  bool operator==(const random_access_const_iterator& o) const { return m_rac == o.m_rac && m_idx == o.m_idx; }
  bool operator!=(const random_access_const_iterator& o) const { return !(*this == o); }
  bool operator<(const random_access_const_iterator& o) const { return m_idx < o.m_idx; }
  bool operator>(const random_access_const_iterator& o) const { return m_idx > o.m_idx; }
  bool operator<=(const random_access_const_iterator& o) const { return m_idx <= o.m_idx; }
  bool operator>=(const random_access_const_iterator& o) const { return m_idx >= o.m_idx; }

 private:
  const t_rac* m_rac = nullptr;
  size_type m_idx = 0;
};

}  // namespace sdsl
#endif
