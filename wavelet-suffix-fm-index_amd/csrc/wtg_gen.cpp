// wtg_gen.cpp -- seeded synthetic workloads (host side).  Counter-based
// splitmix64 streams: value i of stream `seed` depends only on (seed, i), so
// generation is parallel and reproducible.
#include <stdint.h>
#include <string.h>
#include <cstring>
#include <math.h>
#include <thread>
#include <vector>
#include <algorithm>
#include "wtg.h"

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t ctr_rand(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}
// uniform double in (0, 1]
static inline double ctr_unit(uint64_t seed, uint64_t i) {
  return ((ctr_rand(seed, i) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

extern "C" uint64_t wtg_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  return mix64(z);
}

template <typename F>
static void par_for(uint64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (1u << 16)) nt = 1;
  std::vector<std::thread> th;
  uint64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    uint64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=]() { for (uint64_t i = a; i < b; i++) f(i); });
  }
  for (auto& x : th) x.join();
}

static inline void put(void* out, int w, uint64_t i, uint64_t v) {
  switch (w) {
    case 1: ((uint8_t*)out)[i] = (uint8_t)v; break;
    case 2: ((uint16_t*)out)[i] = (uint16_t)v; break;
    case 4: ((uint32_t*)out)[i] = (uint32_t)v; break;
    default: ((uint64_t*)out)[i] = v; break;
  }
}

extern "C" int wtg_gen_uniform(uint64_t seed, uint64_t n, uint64_t range, int w, void* out) {
  if (!out || range == 0 || !(w == 1 || w == 2 || w == 4 || w == 8)) return WTG_ERR_ARG;
  par_for(n, [&](uint64_t i) {
    // multiply-shift maps a 64-bit draw onto [0, range)
    unsigned __int128 p = (unsigned __int128)ctr_rand(seed, i) * range;
    put(out, w, i, (uint64_t)(p >> 64));
  });
  return WTG_OK;
}

extern "C" int wtg_gen_geometric(uint64_t seed, uint64_t n, double p, uint64_t cap, int w, void* out) {
  if (!out || !(p > 0 && p < 1) || !(w == 1 || w == 2 || w == 4 || w == 8)) return WTG_ERR_ARG;
  const double lq = log1p(-p);
  par_for(n, [&](uint64_t i) {
    double k = floor(log(ctr_unit(seed, i)) / lq);
    uint64_t v = (k >= (double)cap) ? cap : (uint64_t)k;
    put(out, w, i, v);
  });
  return WTG_OK;
}

extern "C" int wtg_gen_zipf_bytes(uint64_t seed, uint64_t n, double s, uint8_t* out) {
  // Zipf(s) over ranks 1..256 by an inverse-CDF table, ranks mapped to byte
  // values through a seeded Fisher-Yates permutation of 0..255.
  if (!out || !(s > 0)) return WTG_ERR_ARG;
  double cdf[256], acc = 0;
  for (int r = 0; r < 256; r++) { acc += pow((double)(r + 1), -s); cdf[r] = acc; }
  for (int r = 0; r < 256; r++) cdf[r] /= acc;
  cdf[255] = 1.0;
  uint8_t perm[256];
  for (int i = 0; i < 256; i++) perm[i] = (uint8_t)i;
  uint64_t st = seed ^ 0xA5A5A5A5DEADBEEFull;
  for (int i = 255; i > 0; i--) {
    uint64_t j = wtg_splitmix64(&st) % (uint64_t)(i + 1);
    uint8_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
  }
  par_for(n, [&](uint64_t i) {
    double u = ctr_unit(seed, i);
    int r = (int)(std::lower_bound(cdf, cdf + 256, u) - cdf);
    if (r > 255) r = 255;
    out[i] = perm[r];
  });
  return WTG_OK;
}

extern "C" int wtg_gen_select_vectors(uint64_t seed, uint32_t nvec, uint64_t nbits, uint64_t* words,
                                      double* density) {
  // nvec bitvectors of nbits bits, vector b at words + b*ceil(nbits/64); density
  // d_b log-uniform in [0.01, 0.5].  Even b: Bernoulli(d_b) bits.  Odd b:
  // alternating runs, 0-runs Geometric(mean 64(1-d_b)/d_b), 1-runs Geometric(mean 64),
  // both >= 1, starting with a 0-run.
  if (!words || nbits == 0) return WTG_ERR_ARG;
  const uint64_t nw = (nbits + 63) / 64;
  std::vector<double> d(nvec);
  for (uint32_t b = 0; b < nvec; b++)
    d[b] = exp(log(0.01) + (log(0.5) - log(0.01)) * (ctr_unit(seed ^ 0xD5D5D5D5ull, b) - 1e-300));
  if (density)
    for (uint32_t b = 0; b < nvec; b++) density[b] = d[b];
  // even vectors: word-parallel Bernoulli
  for (uint32_t b = 0; b < nvec; b += 2) {
    const uint64_t sb = mix64(seed + 0x1000 + b);
    const double db = d[b];
    uint64_t* w = words + (uint64_t)b * nw;
    par_for(nw, [&](uint64_t k) {
      uint64_t x = 0;
      for (int j = 0; j < 64; j++) {
        const uint64_t i = k * 64 + j;
        if (i < nbits && ctr_unit(sb, i) <= db) x |= 1ull << j;
      }
      w[k] = x;
    });
  }
  // odd vectors: sequential runs, one thread per vector
  std::vector<uint32_t> odd;
  for (uint32_t b = 1; b < nvec; b += 2) odd.push_back(b);
  std::vector<std::thread> th;
  for (uint32_t b : odd) {
    th.emplace_back([=]() {
      uint64_t* w = words + (uint64_t)b * nw;
      std::memset(w, 0, nw * 8);
      const uint64_t sb = mix64(seed + 0x2000 + b);
      const double l1 = log1p(-1.0 / 64), l0 = log1p(-1.0 / (64 * (1 - d[b]) / d[b]));
      uint64_t pos = 0, k = 0;
      bool one = false;
      while (pos < nbits) {
        const double lq = one ? l1 : l0;
        const uint64_t len = 1 + (uint64_t)floor(log(ctr_unit(sb, k++)) / lq);
        if (one) {
          const uint64_t e = std::min(nbits, pos + len);
          for (uint64_t i = pos; i < e; i++) w[i >> 6] |= 1ull << (i & 63);
        }
        pos += len;
        one = !one;
      }
    });
    if (th.size() >= 16) { for (auto& x : th) x.join(); th.clear(); }
  }
  for (auto& x : th) x.join();
  return WTG_OK;
}

extern "C" int wtg_gen_batch(uint64_t seed, uint32_t nseq, double log2_min, double log2_max,
                             const uint32_t* alphabets, uint32_t nalpha, uint64_t* n, uint32_t* sigma,
                             uint8_t* sym) {
  // n_i = floor(2^U(log2_min, log2_max)), a_i uniform over alphabets, symbols
  // uniform in [0, a_i) (multiply-shift of a counter-based draw)
  if (!n || !sigma || !alphabets || nalpha == 0 || !(log2_max >= log2_min) || log2_min < 0) return WTG_ERR_ARG;
  std::vector<uint64_t> off(nseq + 1, 0);
  for (uint32_t i = 0; i < nseq; i++) {
    const double u = ctr_unit(seed ^ 0xB47C4ull, 2 * (uint64_t)i) - 1e-300;
    n[i] = (uint64_t)floor(exp2(log2_min + (log2_max - log2_min) * u));
    sigma[i] = alphabets[ctr_rand(seed ^ 0xB47C4ull, 2 * (uint64_t)i + 1) % nalpha];
    if (sigma[i] == 0 || sigma[i] > 256) return WTG_ERR_ARG;
    off[i + 1] = off[i] + n[i];
  }
  if (!sym) return WTG_OK;
  par_for(nseq, [&](uint64_t i) {
    const uint64_t si = mix64(seed + 0x3000 + i);
    for (uint64_t k = 0; k < n[i]; k++) {
      const unsigned __int128 p = (unsigned __int128)ctr_rand(si, k) * sigma[i];
      sym[off[i] + k] = (uint8_t)(uint64_t)(p >> 64);
    }
  });
  return WTG_OK;
}
