// wtg_sa.cpp -- host-side suffix array (SA-IS, Nong-Zhang-Chan induced sorting)
// and the BWT of fm-index/par/WT.C:38-41, plus the seeded DNA text / pattern
// generators of the FM-index workload.  The reference builds its SA with
// libdivsufsort (an un-vendored submodule); a suffix array is unique, so any
// exact construction gives the same BWT.
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <vector>
#include "wtg.h"

namespace {

// SA-IS over an integer string s[0..n) with alphabet [0, K); s[n-1] must be the
// unique smallest symbol (the sentinel).  sa receives the suffix array.
template <typename C>
void sais(const C* s, int64_t* sa, int64_t n, int64_t K) {
  if (n == 1) { sa[0] = 0; return; }
  std::vector<uint8_t> t(n);  // 1 = S-type
  t[n - 1] = 1;
  for (int64_t i = n - 2; i >= 0; i--)
    t[i] = (s[i] < s[i + 1] || (s[i] == s[i + 1] && t[i + 1])) ? 1 : 0;
  auto is_lms = [&](int64_t i) { return i > 0 && t[i] && !t[i - 1]; };
  std::vector<int64_t> bkt(K + 1);
  auto buckets = [&](bool end) {
    std::fill(bkt.begin(), bkt.end(), 0);
    for (int64_t i = 0; i < n; i++) bkt[s[i] + 1]++;
    for (int64_t c = 1; c <= K; c++) bkt[c] += bkt[c - 1];
    if (end) for (int64_t c = 0; c < K; c++) bkt[c] = bkt[c + 1];
  };
  auto induce = [&]() {
    buckets(false);  // bucket heads
    for (int64_t i = 0; i < n; i++) {
      const int64_t j = sa[i] - 1;
      if (sa[i] > 0 && !t[j]) sa[bkt[s[j]]++] = j;
    }
    buckets(true);  // bucket tails
    for (int64_t i = n - 1; i >= 0; i--) {
      const int64_t j = sa[i] - 1;
      if (sa[i] > 0 && t[j]) sa[--bkt[s[j]]] = j;
    }
  };
  // stage 1: sort LMS substrings
  std::fill(sa, sa + n, -1);
  buckets(true);
  for (int64_t i = 1; i < n; i++)
    if (is_lms(i)) sa[--bkt[s[i]]] = i;
  induce();
  // compact sorted LMS positions and name the LMS substrings
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++)
    if (is_lms(sa[i])) sa[m++] = sa[i];
  std::fill(sa + m, sa + n, -1);
  int64_t name = 0, prev = -1;
  for (int64_t i = 0; i < m; i++) {
    const int64_t pos = sa[i];
    bool diff = false;
    for (int64_t d = 0; d < n; d++) {
      if (prev == -1 || s[pos + d] != s[prev + d] || t[pos + d] != t[prev + d]) { diff = true; break; }
      if (d > 0 && (is_lms(pos + d) || is_lms(prev + d))) break;
    }
    if (diff) { name++; prev = pos; }
    sa[m + pos / 2] = name - 1;
  }
  std::vector<int64_t> s1;
  s1.reserve(m);
  for (int64_t i = n - 1; i >= m; i--)
    if (sa[i] >= 0) s1.push_back(sa[i]);
  std::reverse(s1.begin(), s1.end());
  // stage 2: sort the reduced string
  std::vector<int64_t> sa1(m);
  if (name < m) sais<int64_t>(s1.data(), sa1.data(), m, name);
  else for (int64_t i = 0; i < m; i++) sa1[s1[i]] = i;
  // stage 3: induce the full SA from the sorted LMS suffixes
  std::vector<int64_t> lms;
  lms.reserve(m);
  for (int64_t i = 1; i < n; i++)
    if (is_lms(i)) lms.push_back(i);
  for (int64_t i = 0; i < m; i++) sa1[i] = lms[sa1[i]];
  std::fill(sa, sa + n, -1);
  buckets(true);
  for (int64_t i = m - 1; i >= 0; i--) {
    const int64_t j = sa1[i];
    sa[--bkt[s[j]]] = j;
  }
  induce();
}

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s;
  uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
  double unit() { return ((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
  // Gamma(a, 1), Marsaglia-Tsang (a < 1 via the a+1 boost)
  double gamma(double a) {
    if (a < 1) return gamma(a + 1) * pow(unit(), 1.0 / a);
    const double d = a - 1.0 / 3, c = 1.0 / sqrt(9 * d);
    for (;;) {
      double x, v;
      do {
        const double u1 = unit(), u2 = unit();
        x = sqrt(-2 * log(u1)) * cos(6.283185307179586 * u2);
        v = 1 + c * x;
      } while (v <= 0);
      v = v * v * v;
      const double u = unit();
      if (log(u) < 0.5 * x * x + d - d * v + d * log(v)) return d * v;
    }
  }
};
}  // namespace

extern "C" int wtg_suffix_array(const uint8_t* t, uint64_t n, int64_t* sa) {
  // lexicographic suffix order with a proper prefix first (divsufsort's order)
  if (!sa || (n && !t)) return WTG_ERR_ARG;
  if (n >= (1ull << 40)) return WTG_ERR_ARG;
  if (n == 0) return WTG_OK;
  bool sentinel = true;  // last byte unique and smallest: SA-IS runs on the text itself
  for (uint64_t i = 0; i + 1 < n && sentinel; i++) sentinel = t[i] > t[n - 1];
  if (sentinel) {
    sais<uint8_t>(t, sa, (int64_t)n, 256);
    return WTG_OK;
  }
  // otherwise append a virtual sentinel: symbols shifted by one, 0 at the end
  std::vector<uint16_t> s(n + 1);
  for (uint64_t i = 0; i < n; i++) s[i] = (uint16_t)(t[i] + 1);
  s[n] = 0;
  std::vector<int64_t> sa1(n + 1);
  sais<uint16_t>(s.data(), sa1.data(), (int64_t)n + 1, 257);
  std::copy(sa1.begin() + 1, sa1.end(), sa);  // sa1[0] is the sentinel
  return WTG_OK;
}

extern "C" int wtg_bwt_from_sa(const uint8_t* t, const int64_t* sa, uint64_t n, uint8_t* bwt) {
  // fm-index/par/WT.C:38-41: bwt[i] = s[SA[i]-1], or s[n-1] when SA[i] == 0
  if (!t || !sa || !bwt) return WTG_ERR_ARG;
  for (uint64_t i = 0; i < n; i++) bwt[i] = sa[i] == 0 ? t[n - 1] : t[sa[i] - 1];
  return WTG_OK;
}

extern "C" int wtg_gen_dna(uint64_t seed, uint64_t n, uint8_t* out) {
  // order-5 Markov chain over ACGT (1024 contexts, Dirichlet(0.5) rows); then
  // copies of seeded source segments of length U[1000,10000] to random
  // destinations until 5% of positions are overwritten; then 0.1% set to 'N';
  // out[n-1] = '\0' (the sentinel of the FM-index text)
  if (!out || n < 2) return WTG_ERR_ARG;
  Rng r{seed};
  static const uint8_t A[4] = {'A', 'C', 'G', 'T'};
  std::vector<double> P(1024 * 4);
  for (int ctx = 0; ctx < 1024; ctx++) {
    double g[4], sum = 0;
    for (int k = 0; k < 4; k++) { g[k] = r.gamma(0.5); sum += g[k]; }
    double acc = 0;
    for (int k = 0; k < 4; k++) { acc += g[k] / sum; P[ctx * 4 + k] = acc; }
    P[ctx * 4 + 3] = 1.0;
  }
  const uint64_t m = n - 1;
  uint32_t ctx = 0;
  for (uint64_t i = 0; i < m; i++) {
    const double u = r.unit();
    int k = 0;
    while (k < 3 && u > P[ctx * 4 + k]) k++;
    out[i] = A[k];
    ctx = ((ctx << 2) | (uint32_t)k) & 1023u;
  }
  uint64_t copied = 0;
  while (m > 20000 && copied < m / 20) {
    const uint64_t len = 1000 + r.next() % 9001;
    const uint64_t src = r.next() % (m - len), dst = r.next() % (m - len);
    memmove(out + dst, out + src, len);
    copied += len;
  }
  for (uint64_t k = 0; k < m / 1000; k++) out[r.next() % m] = 'N';
  out[m] = 0;
  return WTG_OK;
}

extern "C" int wtg_gen_patterns(uint64_t seed, const uint8_t* text, uint64_t n, uint32_t m,
                                uint64_t npat, uint8_t* out) {
  // npat patterns of length m: even index = exact substring at a uniform
  // position, odd index = a substring with 1-2 uniform substitutions by ACGT
  if (!text || !out || n < (uint64_t)m + 2) return WTG_ERR_ARG;
  Rng r{seed};
  static const uint8_t A[4] = {'A', 'C', 'G', 'T'};
  for (uint64_t q = 0; q < npat; q++) {
    const uint64_t p = r.next() % (n - 1 - m);
    uint8_t* o = out + q * m;
    memcpy(o, text + p, m);
    if (q & 1) {
      const int subs = 1 + (int)(r.next() & 1);
      for (int k = 0; k < subs; k++) o[r.next() % m] = A[r.next() & 3];
    }
  }
  return WTG_OK;
}
