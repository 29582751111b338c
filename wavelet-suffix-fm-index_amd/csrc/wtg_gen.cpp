// wtg_gen.cpp -- seeded synthetic workloads (host side).  Counter-based
// splitmix64 streams: value i of stream `seed` depends only on (seed, i), so
// generation is parallel and reproducible.
#include <stdint.h>
#include <string.h>
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
