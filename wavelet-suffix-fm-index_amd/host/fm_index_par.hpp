// fm_index_par.hpp -- drop-in for the FM-index construction of the reference
// (fm-index/par/WT.h:52, fm-index/par/WT.C:33-45):
//   void FMIndex(unsigned char* s, long n, sdsl::int_vector<8>& bwt)
// computes the suffix array (on the GPU by prefix doubling below 2^31 bytes,
// host SA-IS above, in place of divsufsort; the suffix array is unique, so the
// BWT is the same), the BWT bwt[i] = s[SA[i]-1] (s[n-1] where SA[i] = 0) and the
// Huffman-shaped wavelet tree over it on the GPU.
// The reference discards the tree; fm_index below keeps it and answers
// backward-search counts on the host (count) or in device batches.
#ifndef WTG_FM_INDEX_PAR_HPP
#define WTG_FM_INDEX_PAR_HPP
#include <cstdint>
#include <vector>
#include "wt_huff_par.hpp"

namespace wtg {

class fm_index {
 public:
  fm_index() = default;
  // text of n bytes; exact counts need a text ending in a unique smallest byte
  fm_index(const unsigned char* s, uint64_t n) : m_n(n) {
    m_bwt.resize(n);
    if (n) {
      // suffix array and BWT on the GPU (prefix doubling) below 2^31 bytes; the
      // host SA-IS above, or when the device cannot hold the working arrays
      // (about 48n bytes), where the reference's divsufsort would still succeed
      const int rc = n < (1ull << 31) ? gpu_bwt(s, n, m_bwt.data()) : WTG_ERR_NOMEM;
      if (rc == WTG_ERR_NOMEM) {
        std::vector<int64_t> sa(n);
        check(wtg_suffix_array(s, n, sa.data()));
        check(wtg_bwt_from_sa(s, sa.data(), n, m_bwt.data()));
      } else {
        check(rc);
      }
    }
    m_wt = sdsl::wt_huff<>(m_bwt.data(), n);
  }
  const std::vector<unsigned char>& bwt() const { return m_bwt; }
  const sdsl::wt_huff<>& wt() const { return m_wt; }
  uint64_t size() const { return m_n; }

  // occurrences of P[0..m) by backward search over the wavelet tree
  uint64_t count(const unsigned char* P, uint64_t m) const {
    const wtg_huff_shape& S = m_wt.shape();
    uint64_t sp = 0, ep = m_n;
    for (uint64_t j = m; j-- > 0 && sp < ep;) {
      const unsigned char c = P[j];
      if (S.c_to_leaf[c] < 0) return 0;
      sp = S.C[c] + m_wt.rank(sp, c);
      ep = S.C[c] + m_wt.rank(ep, c);
    }
    return ep > sp ? ep - sp : 0;
  }
  // device batch: npat patterns of m bytes at d_pats (pattern q at q*m)
  void count_batch(const uint8_t* d_pats, uint32_t m, uint64_t npat, uint64_t* d_out) const {
    m_wt.count_batch(d_pats, m, npat, d_out);
  }

 private:
  static int gpu_bwt(const unsigned char* s, uint64_t n, unsigned char* out) {
    wtg_ctx* ctx = default_context();
    void *dt = nullptr, *dsa = nullptr, *db = nullptr;
    int rc = wtg_malloc(ctx, &dt, n);
    if (rc == WTG_OK) rc = wtg_malloc(ctx, &dsa, n * 8);
    if (rc == WTG_OK) rc = wtg_malloc(ctx, &db, n);
    if (rc == WTG_OK) rc = wtg_memcpy_h2d(ctx, dt, s, n);
    if (rc == WTG_OK) rc = wtg_suffix_array_gpu(ctx, (const uint8_t*)dt, n, (int64_t*)dsa);
    if (rc == WTG_OK) rc = wtg_bwt_from_sa_gpu(ctx, (const uint8_t*)dt, (const int64_t*)dsa, n, (uint8_t*)db);
    if (rc == WTG_OK) rc = wtg_memcpy_d2h(ctx, out, db, n);
    if (db) wtg_free(ctx, db);
    if (dsa) wtg_free(ctx, dsa);
    if (dt) wtg_free(ctx, dt);
    return rc;
  }

  uint64_t m_n = 0;
  std::vector<unsigned char> m_bwt;
  sdsl::wt_huff<> m_wt;
};

}  // namespace wtg

// the reference's entry point: fills bwt (n entries) and builds the tree
template <class t_int_vector>
inline void FMIndex(unsigned char* s, long n, t_int_vector& bwt) {
  wtg::fm_index idx(s, (uint64_t)n);
  for (long i = 0; i < n; i++) bwt[i] = idx.bwt()[i];
}
#endif
