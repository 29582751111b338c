// wtg_lsel.hip -- learned select over a batch of bitvectors.
//
// Per bitvector: rank9 directory + select_support_mcl<1> samples (the plain
// method), then a continuous piecewise-linear model of j -> select(j) with
// knots at chosen arguments 1 = J_0 < J_1 < ... < J_{K-1} = ones:
//   pred(j) = alpha*j + beta + (1-f)*y[k] + f*y[k+1],   J_k <= j < J_{k+1},
//   f = (j - J_k) / (J_{k+1} - J_k)
// alpha, beta: the line through the first and last training sample (fp64);
// y: fp32 residual knots.
//
// Build:
//   1. placement (k_lsel_place): the arguments are cut into chunks of
//      `max_gap`, one thread per chunk.  A greedy cone walks the chunk's ones
//      in order from the last knot and sets a knot where no line from it stays
//      within `tol` bits of every one seen, so |pred(j) - select(j)| <= tol for
//      every j.  A run of ones is one segment; the gap between two runs costs
//      two knots.  Counted in one pass, written in a second.
//   2. training: forward + backward kernels (work items of at most 16 samples,
//      then a sum per knot interval) and Adam
//      on every knot, minimising the mean squared error on the training
//      samples select(T), select(2T), ... that the plain method answers.
//   3. the error window eps = max_j |round(pred(j)) - select(j)| is measured
//      over EVERY one (exhaustive kernel).  If training raised it above
//      eps_max, the placed knots are restored (their eps <= tol + 1).
//
// Query: a lookup table indexed by j >> lut_log gives the knot interval (a
// short forward probe finishes it); the prediction p names the 512-bit rank9
// block b; eps <= 512 puts select(j) in blocks b-1..b+1, whose directory
// entries give the block and, by their 9-bit word counts, the word; one word
// read and an in-word select finish.  The answer is exact whatever the model
// predicts: outside that window the plain select answers.  The plain sampled
// select stays available (wtg_lsel_plain_batch) as the judge.
#include "wtg_internal.hpp"
#include <functional>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, d);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), d);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    v = o > v ? o : v;
  }
  return v;
}

struct Knot {
  uint32_t j;   // argument (1-based rank)
  float y;      // residual against alpha*j + beta
};

struct LselDesc {
  const uint64_t* words;
  const uint64_t* dir;
  SelQ s1;
  uint64_t nbits, ones;
  uint64_t nblk;         // rank9 entries 0..nblk
  const Knot* knots;
  const uint32_t* lut;   // (ones >> lut_log) + 1 entries: knot interval of argument i << lut_log
  double alpha, beta;
  uint32_t lut_log, nknots;
  int32_t learned;
  int32_t _pad;
};

// the model line, unfused (the placement and the oracle compute it the same way)
__device__ __forceinline__ double lsel_line(double alpha, double beta, double j) {
#pragma clang fp contract(off)
  return alpha * j + beta;
}

// the model's prediction for argument j inside knot interval [a.j, b.j) (the
// last interval also takes j = b.j), clamped to the vector
__device__ __forceinline__ uint64_t lsel_eval(const LselDesc& d, const Knot& a, const Knot& b, uint64_t j) {
  const float f = (float)(j - a.j) / (float)(b.j - a.j);
  const float s = (1.0f - f) * a.y + f * b.y;
  const double p = lsel_line(d.alpha, d.beta, (double)j) + (double)s;
  const double hi = (double)(d.nbits - 1);
  return p <= 0.0 ? 0ull : p >= hi ? d.nbits - 1 : (uint64_t)llrint(p);
}

// move interval k (knots a, b) forward until it holds argument j
__device__ __forceinline__ void lsel_advance(const LselDesc& d, uint64_t j, uint32_t& k, Knot& a, Knot& b) {
  while (j >= b.j && k + 2 < d.nknots) {
    ++k;
    a = b;
    b = d.knots[k + 1];
  }
}

// the model's prediction for argument j (1 <= j <= ones), clamped to the vector
__device__ __forceinline__ uint64_t lsel_predict(const LselDesc& d, uint64_t j) {
  uint32_t k = d.lut[j >> d.lut_log];
  Knot a = d.knots[k], b = d.knots[k + 1];
  lsel_advance(d, j, k, a, b);
  return lsel_eval(d, a, b, j);
}

// exact select1(j) from a predicted position p: the rank9 entries of blocks
// b-1..b+2 (b = p/512) bracket the argument's block, its 9-bit fields the word
__device__ __forceinline__ uint64_t lsel_correct(const LselDesc& d, uint64_t j, uint64_t p) {
  const uint64_t b = p >> 9;
  const uint64_t b0 = b ? b - 1 : 0;
  uint64_t st[4], fl[3];
  if (b0 + 3 <= d.nblk) {
    const uint4* q = reinterpret_cast<const uint4*>(d.dir + 2 * b0);  // 16-byte entries
    uint4 e[3];
#pragma unroll
    for (int i = 0; i < 3; i++) e[i] = q[i];
    const uint64_t s3 = d.dir[2 * (b0 + 3)];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      st[i] = ((uint64_t)e[i].y << 32) | e[i].x;
      fl[i] = ((uint64_t)e[i].w << 32) | e[i].z;
    }
    st[3] = s3;
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint64_t B = b0 + i;
      st[i] = B <= d.nblk ? d.dir[2 * B] : d.ones;
      if (i < 3) fl[i] = B <= d.nblk ? d.dir[2 * B + 1] : 0ull;
    }
  }
  if (st[0] >= j || st[3] < j) return select_plain(d.s1, j);  // outside the window
  const uint32_t i = (uint32_t)(st[1] < j) + (uint32_t)(st[2] < j);
  const uint64_t B = b0 + i;
  const uint64_t s = i == 0 ? st[0] : i == 1 ? st[1] : st[2];
  const uint64_t F = i == 0 ? fl[0] : i == 1 ? fl[1] : fl[2];
  const uint32_t r = (uint32_t)(j - s);  // 1-based rank inside the block (<= 512)
  const uint64_t W = (d.nbits + 63) >> 6;
  const uint32_t mv = W - 8 * B < 8 ? (uint32_t)(W - 8 * B) : 8u;  // words of the block
  uint32_t k = 0, ck = 0;
#pragma unroll
  for (uint32_t m = 1; m < 8; m++) {
    const uint32_t c = (uint32_t)(F >> (63 - 9 * m)) & 0x1FFu;
    if (m < mv && c < r) {
      k = m;
      ck = c;
    }
  }
  const uint64_t w = 8 * B + k;
  return (w << 6) + select_in_word(d.words[w], r - ck - 1);
}

template <bool LEARNED>
__global__ __launch_bounds__(256) void k_lsel_query(const LselDesc* __restrict__ descs, uint32_t nvec,
                                                    const uint32_t* __restrict__ qv,
                                                    const uint64_t* __restrict__ qj, uint64_t nq,
                                                    uint64_t* __restrict__ out) {
  extern __shared__ LselDesc s_desc[];
  for (uint32_t k = threadIdx.x; k < nvec; k += blockDim.x) s_desc[k] = descs[k];
  __syncthreads();
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t b = qv[q];
    const uint64_t j = qj[q];
    uint64_t res = WTG_NONE;
    if (b < nvec) {
      const LselDesc& d = s_desc[b];
      if (j >= 1 && j <= d.ones) {
        if (LEARNED && d.learned) res = lsel_correct(d, j, lsel_predict(d, j));
        else res = select_plain(d.s1, j);
      }
    }
    out[q] = res;
  }
}

// ------------------------------------------------------ access counting
// The 64-byte sectors of HBM one query reads, counted by walking the same
// addresses as k_lsel_query<LEARNED> (a report, not on the query path).
// Plain (select_support_mcl_par.hpp:564-611): long_index[sb], then the long
// pool entry, or superblock[sb] + the mini sample + the scanned words.
// Learned: the lookup table entry, the knots the probe reads, the rank9
// entries of the three-block window and the answer's word.
__device__ __forceinline__ uint32_t sectors_span(const void* first, const void* last) {
  return (uint32_t)(((uintptr_t)last >> 6) - ((uintptr_t)first >> 6) + 1);
}

__device__ uint32_t plain_sectors(const SelQ& s, uint64_t j, uint64_t ans) {
  const uint64_t i = j - 1, sb = i >> 12, off = i & 0xFFF;
  if (s.long_index[sb] >= 0) return 2;
  if ((off & 63) == 0) return 3;
  const uint64_t base = s.superblock[sb] + s.mini[sb * 64 + (off >> 6)];
  return 3 + sectors_span(s.words + ((base + 1) >> 6), s.words + (ans >> 6));
}

__device__ uint32_t learned_sectors(const LselDesc& d, uint64_t j, uint64_t ans) {
  uint32_t k = d.lut[j >> d.lut_log];
  const uint32_t k0 = k;
  while (j >= d.knots[k + 1].j && k + 2 < d.nknots) ++k;
  uint32_t n = 1 + sectors_span(&d.knots[k0], &d.knots[k + 1]);
  const uint64_t b = lsel_predict(d, j) >> 9;
  const uint64_t b0 = b ? b - 1 : 0;
  const uint64_t last = b0 + 3 <= d.nblk ? b0 + 3 : d.nblk;
  n += b0 <= d.nblk ? sectors_span(d.dir + 2 * b0, d.dir + 2 * last) : 0;
  const uint64_t s0 = b0 <= d.nblk ? d.dir[2 * b0] : d.ones;
  const uint64_t s3 = b0 + 3 <= d.nblk ? d.dir[2 * (b0 + 3)] : d.ones;
  if (s0 >= j || s3 < j) return n + plain_sectors(d.s1, j, ans);  // outside the window
  return n + 1;
}

template <bool LEARNED>
__global__ __launch_bounds__(256) void k_lsel_sectors(const LselDesc* __restrict__ descs, uint32_t nvec,
                                                      const uint32_t* __restrict__ qv,
                                                      const uint64_t* __restrict__ qj,
                                                      const uint64_t* __restrict__ ans, uint64_t nq,
                                                      uint32_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint32_t b = qv[q];
  const uint64_t j = qj[q];
  uint32_t n = 0;
  if (b < nvec && j >= 1 && j <= descs[b].ones) {
    const LselDesc& d = descs[b];
    n = (LEARNED && d.learned) ? learned_sectors(d, j, ans[q]) : plain_sectors(d.s1, j, ans[q]);
  }
  out[q] = n;
}

// ------------------------------------------------------------- placement
struct PlaceVec {
  SelQ s1;
  const uint64_t* words;
  uint64_t nwords;
  uint64_t ones;
  uint64_t cbase;        // first flat chunk index of this vector
  double alpha, beta;
};

__device__ __forceinline__ uint32_t find_base(const uint64_t* base, uint32_t nv, uint64_t flat) {
  uint32_t lo = 0, hi = nv;  // last vector with base <= flat
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (base[mid] <= flat) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Greedy knot placement over one chunk of arguments [j0, j1], one wave per
// chunk.  The sequential walk it reproduces: from the last knot (x0, y0) keep
// the cone [lo, hi] of slopes whose line passes within tol of every one seen,
// lo = max over ones of (P - tol - y0) / (j - x0), hi = min of (P + tol - y0) /
// (j - x0); the first argument j that empties it puts a knot at the previous
// argument jp on the cone's middle slope, y0 + (lo + hi)/2 * (jp - x0), and
// the walk restarts from there with j.  Here the ones come 64 at a time (a
// window of 64 words: popcount scan; each non-empty word marks its first one's
// index in a per-wave LDS row and a running max over that row gives every one
// its word), lane i takes argument jn + i, and a running max / min over the
// lanes (a prefix scan: max and min are exact, so the values are the
// sequential ones) finds the first lane whose cone is empty; the knot goes in,
// and the lanes from that one on are redone from the new knot.  Same double
// arithmetic as the walk (no contraction), so the same knots.  The scans are
// DPP moves (no LDS round trips).
// Count pass (off == nullptr): cnt[flat] = the chunk's knots, and the first
// LSEL_CAP of them to scratch[flat * LSEL_CAP ..].  Emit pass: a chunk with at
// most LSEL_CAP knots copies them from scratch; a larger one walks again and
// writes knots[off[flat] ..] (a different count sets *err).
constexpr uint32_t LSEL_CAP = 64;
__global__ __launch_bounds__(256) void k_lsel_place(const PlaceVec* __restrict__ pv, const uint64_t* __restrict__ cbase,
                                                    uint32_t nv, uint64_t nflat, uint32_t gap_log, double tol,
                                                    uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                                                    Knot* __restrict__ knots, Knot* __restrict__ scratch,
                                                    uint32_t* __restrict__ err) {
#pragma clang fp contract(off)
  const uint32_t lane = (uint32_t)lane_id();
  const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  __shared__ int32_t s_first[4][64];  // per wave: the window word whose first one is one i
  int32_t* first = s_first[wv];
  const uint64_t flat = (uint64_t)blockIdx.x * 4 + wv;
  if (flat >= nflat) return;
  const PlaceVec& v = pv[find_base(cbase, nv, flat)];
  const uint64_t c = flat - v.cbase;
  const uint64_t j0 = (c << gap_log) + 1;
  const uint64_t j1 = ((c + 1) << gap_log) < v.ones ? ((c + 1) << gap_log) : v.ones;
  const uint32_t lim = off ? off[flat + 1] - off[flat] : 0u;
  if (off && lim <= LSEL_CAP) {  // the count pass kept them
    for (uint32_t i = lane; i < lim; i += 64) knots[off[flat] + i] = scratch[flat * LSEL_CAP + i];
    return;
  }
  Knot* out = off ? knots + off[flat] : scratch + flat * LSEL_CAP;
  const uint32_t cap = off ? lim : LSEL_CAP;
  uint32_t n = 0;  // wave-uniform
  auto put = [&](uint64_t jj, double val) {
    if (n < cap && lane == 0) out[n] = Knot{(uint32_t)jj, (float)(val - lsel_line(v.alpha, v.beta, (double)jj))};
    n++;
  };
  if (j0 <= j1) {
    const uint64_t pos = select_plain(v.s1, j0);
    uint64_t w = pos >> 6;
    uint64_t x = v.words[w] & (~0ull << (pos & 63));
    x &= x - 1;  // the one at pos itself
    double x0 = (double)j0, y0 = (double)pos, lo = -1e300, hi = 1e300;
    uint64_t jp = j0, jn = j0 + 1;
    put(j0, y0);
    while (jn <= j1) {
      // the window: word w (its unconsumed ones x) and the 63 words after it
      const uint64_t wl = w + lane;
      const uint64_t wd = lane == 0 ? x : (wl < v.nwords ? v.words[wl] : 0ull);
      const uint32_t pc = popc64(wd);
      const uint32_t incl = wave_incl_sum_dpp(pc);
      const uint32_t T = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
      const uint64_t need = j1 - jn + 1;
      const uint32_t m = (uint32_t)(need < (uint64_t)(T < 64u ? T : 64u) ? need : (T < 64u ? T : 64u));
      if (m == 0) {  // 64 words without a one
        if (w + 64 >= v.nwords) {  // no ones left: inconsistent with the count
          if (lane == 0) *err = 1u;
          break;
        }
        w += 64;
        x = v.words[w];
        continue;
      }
      // lane i (< m): the i-th one of the window, in word k = the last
      // non-empty word whose first one has index <= i (the first one of word
      // l has index incl - pc)
      first[lane] = -1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (pc && incl - pc < 64u) first[incl - pc] = (int32_t)lane;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int32_t kw = wave_incl_max_dpp(first[lane]);
      const uint32_t ks = kw < 0 ? 0u : (uint32_t)kw;
      const uint32_t ex_k = (uint32_t)__shfl((int)incl, (int)ks) - (uint32_t)__shfl((int)pc, (int)ks);
      const uint64_t wk = shfl_u64(wd, (int)ks);
      const uint32_t bit = lane < m ? select_in_word(wk, lane - ex_k) : 0u;
      const double P = (double)(((w + ks) << 6) + bit);
      const double jj = (double)(jn + lane);
      // the cone over lanes [start, m)
      uint32_t start = 0;
      for (uint32_t it = 0;; it++) {
        if (it > 64) {  // each round moves start forward: cannot happen
          if (lane == 0) *err = 1u;
          break;
        }
        const bool act = lane >= start && lane < m;
        double a = -1e300, b = 1e300;
        if (act) {
          const double dx = jj - x0;
          a = (P - tol - y0) / dx;
          b = (P + tol - y0) / dx;
        }
        wave_incl_maxmin_f64_dpp(a, b);
        a = lo > a ? lo : a;
        b = hi < b ? hi : b;
        const uint64_t fail = wave_ballot(act && a > b);
        if (fail == 0) {
          lo = readlane_f64(a, (int)(m - 1));
          hi = readlane_f64(b, (int)(m - 1));
          jp = jn + m - 1;
          break;
        }
        const uint32_t f = (uint32_t)__builtin_ctzll(fail);
        double plo = lo, phi = hi;
        uint64_t jq = jp;
        if (f > start) {
          plo = readlane_f64(a, (int)(f - 1));
          phi = readlane_f64(b, (int)(f - 1));
          jq = jn + f - 1;
        }
        const double ye = y0 + 0.5 * (plo + phi) * ((double)jq - x0);
        put(jq, ye);
        x0 = (double)jq;
        y0 = ye;
        lo = -1e300;
        hi = 1e300;
        jp = jq;
        start = f;
      }
      // consume the window's first m ones
      const uint32_t kl = (uint32_t)__builtin_amdgcn_readlane((int)ks, (int)(m - 1));
      const uint32_t bl = (uint32_t)__builtin_amdgcn_readlane((int)bit, (int)(m - 1));
      const uint64_t wlast = shfl_u64(wd, (int)kl);
      w += kl;
      x = wlast & ~((2ull << bl) - 1);
      jn += m;
    }
    if ((double)jp > x0) put(jp, y0 + 0.5 * (lo + hi) * ((double)jp - x0));
  }
  if (!off) cnt[flat] = n;
  else if (n != lim) *err = 1u;
}

// lookup table: thread per knot interval k, entries i with i << lut_log in
// the interval's arguments [J_k, J_{k+1}) (the first from 0, the last to ones)
// (every model in one launch: model vids[blockIdx.y], blocks stride over its knots)
__global__ __launch_bounds__(256) void k_lsel_lut(const LselDesc* __restrict__ descs,
                                                  const uint32_t* __restrict__ vids) {
  const LselDesc& d = descs[vids[blockIdx.y]];
  const Knot* K = d.knots;
  const uint32_t nk = d.nknots, lut_log = d.lut_log;
  const uint64_t ones = d.ones, nlut = (ones >> lut_log) + 1;
  uint32_t* lut = const_cast<uint32_t*>(d.lut);  // the table being built
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k + 1 < nk; k += gridDim.x * blockDim.x) {
    const uint64_t lo = k == 0 ? 0ull : K[k].j;
    const uint64_t hi = k + 2 == nk ? ones + 1 : K[k + 1].j;
    for (uint64_t i = (lo + (1ull << lut_log) - 1) >> lut_log; (i << lut_log) < hi && i < nlut; i++)
      lut[i] = k;
  }
}

// ------------------------------------------------------------- training
// One model's view for the training kernels.  Several models train together:
// their knots are consecutive in one array; thread g handles the interval
// starting at global knot g (none for a model's last knot).
struct TrainVec {
  const uint64_t* pos;   // ns training samples: pos[t] = select(T*(t+1))
  float* tgt;            // ns targets: pos[t] - line(T*(t+1)), rounded to float once
  uint64_t ns;
  uint64_t kfirst;       // global index of the model's first knot
  uint32_t nknots, tlog;
  double alpha, beta;
};

struct TrainBufs {
  Knot* K;
  float* m1;             // Adam moments
  float* m2;
  float* gl;             // per interval: sum e*(1-f), sum e*f, sum e^2
  float* gr;
  float* le;
  const uint64_t* kfirst;  // [nv] model first knots (for the vector search)
};

// the samples j = T*u of interval k (knots a, b): u in [t1, t2] (none when t1 > t2)
__device__ __forceinline__ void interval_samples(const TrainVec& v, const Knot& a, const Knot& b, uint64_t k,
                                                 uint64_t& t1, uint64_t& t2) {
  const uint64_t T = 1ull << v.tlog;
  t1 = ((uint64_t)a.j + T - 1) >> v.tlog;
  if (t1 < 1) t1 = 1;
  t2 = (k + 2 == v.nknots ? (uint64_t)b.j : (uint64_t)b.j - 1) >> v.tlog;
  if (t2 > v.ns) t2 = v.ns;
}

// Training in work items: interval g's samples are cut at the multiples of
// LSEL_ITEM of their target index into items of at most LSEL_ITEM samples
// (nitem[g] of them, the first at item_start[g]; each reads one aligned block),
// a thread per item sums its samples in order, and a thread per interval sums
// its items' partial sums in order (the same sums every run; a wave's threads
// do about equal work whatever the interval lengths).
constexpr uint32_t LSEL_ITEM = 16;
struct TrainItems {
  uint32_t* model_of;     // [nk] the model of each interval
  uint32_t* nitem;        // [nk] items per interval
  uint32_t* item_start;   // [nk + 1] exclusive prefix of nitem (item_start[nk] = items)
  uint32_t* item_g;       // [cap] the interval of each item
  float *pa, *pc, *pe;    // [cap] partial sums per item
  uint64_t cap;           // items upper bound (grid size)
};

__global__ __launch_bounds__(256) void k_lsel_items_count(const TrainVec* __restrict__ tv, uint32_t nv, TrainBufs tb,
                                                          uint64_t nk_total, TrainItems it) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk_total) return;
  const uint32_t m = find_base(tb.kfirst, nv, g);
  const TrainVec& v = tv[m];
  const uint64_t k = g - v.kfirst;
  uint32_t c = 0;
  if (k + 1 < v.nknots) {
    uint64_t t1, t2;
    interval_samples(v, tb.K[g], tb.K[g + 1], k, t1, t2);
    if (t2 >= t1) c = (uint32_t)((t2 - 1) / LSEL_ITEM - (t1 - 1) / LSEL_ITEM + 1);
  }
  it.model_of[g] = m;
  it.nitem[g] = c;
}

__global__ __launch_bounds__(256) void k_lsel_items_map(uint64_t nk_total, TrainItems it) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk_total) return;
  const uint32_t s0 = it.item_start[g], c = it.nitem[g];
  for (uint32_t q = 0; q < c; q++) it.item_g[s0 + q] = (uint32_t)g;
}

// the training targets of model blockIdx.y: the samples' offsets from the
// model's baseline line (they do not change while the knots train)
__global__ __launch_bounds__(256) void k_lsel_targets(const TrainVec* __restrict__ tv) {
  const TrainVec v = tv[blockIdx.y];
  const uint64_t T = 1ull << v.tlog;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < v.ns; t += (uint64_t)gridDim.x * blockDim.x)
    v.tgt[t] = (float)((double)v.pos[t] - lsel_line(v.alpha, v.beta, (double)((t + 1) * T)));
}

// forward + backward of one item: its samples' residuals, partial gradients, loss
__global__ __launch_bounds__(256) void k_lsel_items(const TrainVec* __restrict__ tv, TrainBufs tb,
                                                    uint64_t nk_total, TrainItems it) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= it.item_start[nk_total]) return;
  const uint32_t g = it.item_g[i];
  const TrainVec& v = tv[it.model_of[g]];
  const uint64_t k = g - v.kfirst;
  const Knot a = tb.K[g], b = tb.K[(uint64_t)g + 1];
  const uint64_t T = 1ull << v.tlog;
  uint64_t t1, t2;
  interval_samples(v, a, b, k, t1, t2);
  // the item's aligned block of LSEL_ITEM targets (indices x = u - 1), clipped to [t1, t2]
  const uint64_t xb = ((t1 - 1) / LSEL_ITEM + (uint64_t)(i - it.item_start[g])) * LSEL_ITEM;
  const uint64_t u0 = xb + 1 > t1 ? xb + 1 : t1;
  const uint64_t ue = xb + LSEL_ITEM < t2 ? xb + LSEL_ITEM : t2;
  const float span = (float)(b.j - a.j);
  // four aligned 16-byte loads (each model's targets start 256-byte aligned
  // and are padded to a multiple of 64)
  float tg[LSEL_ITEM];
  const float4* tb4 = reinterpret_cast<const float4*>(v.tgt + xb);
#pragma unroll
  for (uint32_t r = 0; r < LSEL_ITEM / 4; r++) {
    const float4 f4 = tb4[r];
    tg[4 * r] = f4.x;
    tg[4 * r + 1] = f4.y;
    tg[4 * r + 2] = f4.z;
    tg[4 * r + 3] = f4.w;
  }
  float sa = 0.f, sc = 0.f, e2 = 0.f;
#pragma unroll
  for (uint32_t q = 0; q < LSEL_ITEM; q++) {
    const uint64_t u = xb + 1 + q;  // sample u, target tg[q]
    if (u >= u0 && u <= ue) {
      const uint64_t j = u * T;
      const float f = (float)(j - a.j) / span;
      const float e = (1.0f - f) * a.y + f * b.y - tg[q];
      sa += e * (1.0f - f);
      sc += e * f;
      e2 += e * e;
    }
  }
  it.pa[i] = sa;
  it.pc[i] = sc;
  it.pe[i] = e2;
}

// per interval: the sums of its items, in item order
__global__ __launch_bounds__(256) void k_lsel_items_sum(TrainBufs tb, uint64_t nk_total, TrainItems it) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk_total) return;
  const uint32_t s0 = it.item_start[g], s1 = it.item_start[g + 1];
  float sa = 0.f, sc = 0.f, e2 = 0.f;
  for (uint32_t q = s0; q < s1; q++) {
    sa += it.pa[q];
    sc += it.pc[q];
    e2 += it.pe[q];
  }
  tb.gl[g] = sa;
  tb.gr[g] = sc;
  tb.le[g] = e2;
}

// d loss / d y_k = 2/ns * (gl[k] + gr[k-1])
__device__ __forceinline__ float knot_grad(const TrainVec& v, const TrainBufs& tb, uint64_t g) {
  const uint64_t k = g - v.kfirst;
  float s = 0.f;
  if (k + 1 < v.nknots) s += tb.gl[g];
  if (k >= 1) s += tb.gr[g - 1];
  return s * (2.0f / (float)(v.ns ? v.ns : 1));
}

// Adam step of knot g; its gradient from the item sums of the intervals on
// either side (summed in item order, as k_lsel_items_sum does)
__global__ __launch_bounds__(256) void k_lsel_adam(const TrainVec* __restrict__ tv, uint32_t nv, TrainBufs tb,
                                                   uint64_t nk_total, TrainItems it, float lr, float b1, float b2,
                                                   float bc1, float bc2) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk_total) return;
  const TrainVec& v = tv[it.model_of[g]];
  const uint64_t k = g - v.kfirst;
  float sa = 0.f, sc = 0.f;
  if (k + 1 < v.nknots)
    for (uint32_t q = it.item_start[g], q1 = it.item_start[g + 1]; q < q1; q++) sa += it.pa[q];
  if (k >= 1)
    for (uint32_t q = it.item_start[g - 1], q1 = it.item_start[g]; q < q1; q++) sc += it.pc[q];
  float sg = 0.f;
  if (k + 1 < v.nknots) sg += sa;
  if (k >= 1) sg += sc;
  const float gr = sg * (2.0f / (float)(v.ns ? v.ns : 1));
  const float m = b1 * tb.m1[g] + (1.f - b1) * gr;
  const float s = b2 * tb.m2[g] + (1.f - b2) * gr * gr;
  tb.m1[g] = m;
  tb.m2[g] = s;
  tb.K[g].y -= lr * (m / bc1) / (sqrtf(s / bc2) + 1e-6f);
}

// the gradient alone (for checks), one model
__global__ void k_lsel_grad(const TrainVec* __restrict__ tv, TrainBufs tb, float* __restrict__ grad) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= tv[0].nknots) return;
  grad[g] = knot_grad(tv[0], tb, g);
}

// sum of squared residuals of model b (training loss numerator)
__global__ __launch_bounds__(256) void k_lsel_loss(const TrainVec* __restrict__ tv, TrainBufs tb,
                                                   double* __restrict__ out) {
  const uint32_t b = blockIdx.x;  // one block per model
  const TrainVec& v = tv[b];
  double acc = 0;
  for (uint64_t k = threadIdx.x; k + 1 < v.nknots; k += blockDim.x) acc += (double)tb.le[v.kfirst + k];
  __shared__ double s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = s[0];
}

// exact error window eps = max over every one j of |pred(j) - select(j)| of
// model vids[blockIdx.y], in two passes.
// k_lsel_eps_lb: the exact errors at the model's training samples (each is a
// one: select(T*(t+1)) at pos[t]); their maximum, a lower bound of eps, goes
// to eps[b].
// k_lsel_eps: a thread per run of LSEL_EPS_W consecutive words (the run looks
// its knot interval up once and moves it forward, the interval lsel_predict
// finds for each argument, and the same arithmetic, lsel_eval).  Inside one
// knot interval the prediction is linear in j (up to the rounding of its float
// and double steps, within `margin`), and the i-th of a word's c ones lies in
// [pfirst + i, plast - (c-1-i)]: no error of the word exceeds the larger of its
// end errors, qlast - pfirst - (c-1) and plast - qfirst - (c-1), plus the
// margin.  Every word's end errors are exact and kept; a word of more than two
// ones whose bound exceeds the lower bound (or whose ones span two intervals)
// can raise it: the whole wave evaluates its ones.
__global__ __launch_bounds__(256) void k_lsel_eps_lb(const LselDesc* __restrict__ descs,
                                                     const uint32_t* __restrict__ vids,
                                                     const TrainVec* __restrict__ tv,
                                                     unsigned long long* __restrict__ eps) {
  const uint32_t b = vids[blockIdx.y];
  const LselDesc d = descs[b];
  const TrainVec v = tv[blockIdx.y];
  uint64_t mx = 0;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < v.ns; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = lsel_predict(d, (t + 1) << v.tlog), p = v.pos[t];
    const uint64_t e = q > p ? q - p : p - q;
    mx = e > mx ? e : mx;
  }
  // one atomic per block (eps[b] only grows: a block whose maximum is not
  // above the value it reads cannot raise it)
  __shared__ uint64_t s_mx[4];
  mx = wave_max_u64(mx);
  if (lane_id() == 0) s_mx[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t m = s_mx[0];
    for (uint32_t k = 1; k < blockDim.x / 64; k++) m = s_mx[k] > m ? s_mx[k] : m;
    if (m > __atomic_load_n(&eps[b], __ATOMIC_RELAXED)) atomicMax(&eps[b], (unsigned long long)m);
  }
}

constexpr uint32_t LSEL_EPS_W = 16;
__global__ __launch_bounds__(256) void k_lsel_eps(const LselDesc* __restrict__ descs,
                                                  const uint32_t* __restrict__ vids,
                                                  unsigned long long* __restrict__ eps) {
  const uint32_t b = vids[blockIdx.y];
  const LselDesc d = descs[b];
  const uint64_t W = (d.nbits + 63) >> 6;
  const uint64_t lb = (uint64_t)eps[b];
  const uint32_t lane = (uint32_t)lane_id();
  uint64_t mx = 0;
  auto dist = [](uint64_t q, uint64_t p) { return q > p ? q - p : p - q; };
  auto rl64 = [](uint64_t v, int l) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  };
  // wave-uniform loop (the candidate words below are evaluated by the whole wave)
  for (uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * LSEL_EPS_W;;
       w0 += (uint64_t)gridDim.x * blockDim.x * LSEL_EPS_W) {
    const bool live = w0 < W;
    if (!wave_ballot(live)) break;
    uint64_t xs[LSEL_EPS_W];
#pragma unroll
    for (uint32_t q = 0; q < LSEL_EPS_W; q++) xs[q] = w0 + q < W ? d.words[w0 + q] : 0ull;
    uint64_t j = live ? rank9_dev(d.dir, d.words, w0 << 6) : 0ull;  // ones before the run
    uint32_t k = 0;
    Knot ka{0u, 0.f}, kb{0u, 0.f};
    bool have = false;
#pragma unroll
    for (uint32_t q = 0; q < LSEL_EPS_W; q++) {
      const uint64_t x = xs[q];
      const uint64_t w = w0 + q, jb = j;
      bool cand = false;
      uint32_t kf = 0;
      Knot fa{0u, 0.f}, fb{0u, 0.f};
      if (x) {
        const uint32_t c = popc64(x);
        const uint64_t jf = j + 1, jl = j + c;
        if (!have) {
          k = d.lut[jf >> d.lut_log];
          ka = d.knots[k];
          kb = d.knots[k + 1];
          have = true;
        }
        lsel_advance(d, jf, k, ka, kb);
        kf = k;
        fa = ka;
        fb = kb;
        const uint64_t pf = (w << 6) + (uint64_t)__builtin_ctzll(x);
        const uint64_t pl = (w << 6) + 63 - (uint64_t)__builtin_clzll(x);
        const uint64_t qf = lsel_eval(d, ka, kb, jf);
        uint64_t ql = qf;
        if (c > 1) {
          lsel_advance(d, jl, k, ka, kb);
          ql = lsel_eval(d, ka, kb, jl);
        }
        {
          // the word's end ones: exact errors, kept directly (a word of one or
          // two ones is then done)
          const uint64_t e0 = dist(qf, pf) > dist(ql, pl) ? dist(qf, pf) : dist(ql, pl);
          mx = e0 > mx ? e0 : mx;
          cand = k != kf && c > 2;
          if (!cand && c > 2) {
            // the i-th one (i = 0..c-1) sits in [pf + i, pl - (c-1-i)]: its error
            // q_i - p_i is at most linear in i, so at most qf - pf or
            // ql - pf - (c-1); p_i - q_i at most pl - ql or pl - (c-1) - qf
            const int64_t margin = 2 + (int64_t)((fabsf(fa.y) + fabsf(fb.y)) * 0x1p-19f);
            const int64_t cm = (int64_t)c - 1;
            const int64_t u1 = (int64_t)ql - (int64_t)pf - cm, u2 = (int64_t)pl - (int64_t)qf - cm;
            const int64_t u0 = (int64_t)e0;
            const int64_t ub = (u0 > u1 ? (u0 > u2 ? u0 : u2) : (u1 > u2 ? u1 : u2)) + margin;
            cand = ub > (int64_t)lb;
          }
        }
        j += c;
      }
      {
        // each candidate word by the whole wave: lane i takes the word's bit i
        // (its rank from the word's rank base, its interval moved forward
        // from the one of the word's first one: the walk's arithmetic)
        for (uint64_t cmask = wave_ballot(cand); cmask; cmask &= cmask - 1) {
          const int l = (int)__builtin_ctzll(cmask);
          const uint64_t xw = rl64(x, l), jw = rl64(jb, l), ww = rl64(w, l);
          uint32_t k2 = (uint32_t)__builtin_amdgcn_readlane((int)kf, l);
          Knot a2{(uint32_t)__builtin_amdgcn_readlane((int)fa.j, l),
                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fa.y), l))};
          Knot b2{(uint32_t)__builtin_amdgcn_readlane((int)fb.j, l),
                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fb.y), l))};
          if ((xw >> lane) & 1ull) {
            const uint64_t jj = jw + popc64(xw & ((1ull << lane) - 1ull)) + 1;
            lsel_advance(d, jj, k2, a2, b2);
            const uint64_t e = dist(lsel_eval(d, a2, b2, jj), (ww << 6) + lane);
            mx = e > mx ? e : mx;
          }
        }
      }
    }
  }
  mx = wave_max_u64(mx);
  if (lane == 0 && mx > lb) atomicMax(&eps[b], (unsigned long long)mx);
}

// training samples of every vector in one launch: pos[u-1] = select1(u*T)
// for u = 1..ns, found by walking the words (thread per run of LSEL_EPS_W
// words from its rank9 count: the samples whose argument falls in a word are
// selected inside it); the first and last also to ends[2b], ends[2b+1]
struct SampleVec {
  uint64_t* pos;
  uint64_t ns;
};
__global__ __launch_bounds__(256) void k_lsel_samples(const LselDesc* __restrict__ descs,
                                                      const uint32_t* __restrict__ vids,
                                                      const SampleVec* __restrict__ sv, uint32_t tlog,
                                                      uint64_t* __restrict__ ends) {
  const uint32_t b = vids[blockIdx.y];
  const LselDesc d = descs[b];
  const SampleVec s = sv[blockIdx.y];
  const uint64_t W = (d.nbits + 63) >> 6, T = 1ull << tlog;
  for (uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * LSEL_EPS_W; w0 < W;
       w0 += (uint64_t)gridDim.x * blockDim.x * LSEL_EPS_W) {
    uint64_t xs[LSEL_EPS_W];  // the run's words, all loads in flight together
#pragma unroll
    for (uint32_t q = 0; q < LSEL_EPS_W; q++) xs[q] = w0 + q < W ? d.words[w0 + q] : 0ull;
    uint64_t j = rank9_dev(d.dir, d.words, w0 << 6);  // ones before the word
#pragma unroll
    for (uint32_t q = 0; q < LSEL_EPS_W; q++) {
      const uint64_t w = w0 + q, x = xs[q];
      const uint64_t c = (uint64_t)__builtin_popcountll(x);
      // arguments u*T in (j, j + c]
      for (uint64_t u = (j >> tlog) + 1; u * T <= j + c && u <= s.ns; u++) {
        const uint64_t p = (w << 6) + select_in_word(x, (uint32_t)(u * T - j - 1));
        s.pos[u - 1] = p;
        if (u == 1) ends[2 * b] = p;
        if (u == s.ns) ends[2 * b + 1] = p;
      }
      j += c;
    }
  }
}

// out[b] = the count at byte off of workspace b (uint64 or uint32)
__global__ __launch_bounds__(256) void k_lsel_gather(const char* __restrict__ ws, uint64_t stride, uint64_t off,
                                                     uint32_t n, uint32_t bytes, uint64_t* __restrict__ out) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const char* p = ws + b * stride + off;
  out[b] = bytes == 8 ? *(const uint64_t*)p : (uint64_t) * (const uint32_t*)p;
}

}  // namespace

struct wtg_lsel {
  uint32_t nvec = 0;
  std::vector<LselDesc> desc;          // host copy
  LselDesc* d_desc = nullptr;
  std::vector<uint64_t*> dirs;         // rank directories (in one of the blocks)
  std::vector<wtg_select> sel;         // select1 structures (their content in the blocks)
  std::vector<wtg_lsel_vec_info> info;
  std::vector<void*> blocks;           // owned model / sample memory
};

static void lsel_release(wtg_ctx* ctx, wtg_lsel* ls) {
  if (!ls) return;
  wtg_free(ctx, ls->d_desc);
  for (auto* p : ls->blocks) wtg_free(ctx, p);
  delete ls;
}

extern "C" int wtg_lsel_destroy(wtg_ctx* ctx, wtg_lsel* ls) {
  if (!ls) return WTG_OK;
  WTG_REQUIRE(ctx, "null ctx");
  lsel_release(ctx, ls);
  return WTG_OK;
}

static int lsel_upload_desc(wtg_ctx* ctx, wtg_lsel* ls) {
  WTG_CHECK_HIP(hipMemcpyAsync(ls->d_desc, ls->desc.data(), ls->nvec * sizeof(LselDesc),
                               hipMemcpyHostToDevice, ctx->stream));
  return WTG_OK;
}

static int alloc_block(wtg_ctx* ctx, wtg_lsel* ls, size_t bytes, void** p) {
  WTG_TRY(wtg_malloc(ctx, p, bytes));
  ls->blocks.push_back(*p);
  return WTG_OK;
}

static int sync_copy(wtg_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (kind == hipMemcpyDeviceToHost) return wtg_stage_d2h(ctx, dst, src, bytes);
  WTG_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return WTG_OK;
}

// the work items of the training kernels (their intervals' sample counts do
// not change: training moves the knots' y only); scratch from `mem`
static size_t train_items_bytes(uint64_t nk, uint64_t cap) {
  return wtg_align256(nk * 4) * 2 + wtg_align256((nk + 1) * 4) + wtg_align256(cap * 4) * 4 +
         wtg_align256((wtg_ceil_div(nk + 1, 4096) + 2) * 8) + 256;
}
static int train_items_setup(wtg_ctx* ctx, const TrainVec* d_tv, uint32_t nv, const TrainBufs& tb, uint64_t nk,
                             uint64_t cap, uint64_t nsmax, char* mem, TrainItems& it) {
  char* p = mem;
  auto take = [&](size_t bytes) { char* r = p; p += wtg_align256(bytes); return r; };
  it.model_of = (uint32_t*)take(nk * 4);
  it.nitem = (uint32_t*)take(nk * 4);
  it.item_start = (uint32_t*)take((nk + 1) * 4);
  it.item_g = (uint32_t*)take(cap * 4);
  it.pa = (float*)take(cap * 4);
  it.pc = (float*)take(cap * 4);
  it.pe = (float*)take(cap * 4);
  it.cap = cap;
  const uint64_t tiles = wtg_ceil_div(nk + 1, 4096);
  uint64_t* states = (uint64_t*)take((tiles + 2) * 8);
  uint32_t* ctr = (uint32_t*)p;
  WTG_TRY(wtg_fill_bytes_zero(ctx, states, (char*)ctr + 256 - (char*)states));
  const unsigned gk = (unsigned)wtg_ceil_div(nk, 256);
  const unsigned gt = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(wtg_ceil_div(nsmax, 256), 1024));
  WTG_LAUNCH(ctx, "lsel_targets", k_lsel_targets, dim3(gt, nv), dim3(256), 0, d_tv);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_items", k_lsel_items_count, dim3(gk), dim3(256), 0, d_tv, nv, tb, nk, it);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_items", k_scan_u32<uint32_t>, dim3((unsigned)wtg_ceil_div(nk, 4096)), dim3(256), 0,
             it.nitem, it.item_start, nk, states, ctr, ctx->d_flag);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_items", k_lsel_items_map, dim3(gk), dim3(256), 0, nk, it);
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}
// gl / gr / le of every interval
static int train_grad(wtg_ctx* ctx, const TrainVec* d_tv, const TrainBufs& tb, uint64_t nk, const TrainItems& it) {
  WTG_LAUNCH(ctx, "lsel_fwd_bwd", k_lsel_items, dim3((unsigned)wtg_ceil_div(it.cap, 256)), dim3(256), 0, d_tv, tb,
             nk, it);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_fwd_bwd", k_lsel_items_sum, dim3((unsigned)wtg_ceil_div(nk, 256)), dim3(256), 0, tb, nk, it);
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}

// knot placement, lookup tables, training and the eps check for the vectors
// in `act` (all memory goes to ls->blocks, released by the caller on failure)
static int lsel_models(wtg_ctx* ctx, wtg_lsel* ls, const wtg_lsel_params& P, const uint64_t* nbits,
                       const std::vector<uint64_t*>& pos, const std::vector<uint64_t>& ns,
                       const std::vector<uint32_t>& act) {
  const uint32_t nvec = ls->nvec;
  const uint32_t tlog = (uint32_t)__builtin_ctz(P.train_stride);
  const uint32_t gap_log = (uint32_t)__builtin_ctz(P.max_gap);
  const uint32_t na = (uint32_t)act.size();
  // 3. knot placement: count, scan (host), emit
  std::vector<PlaceVec> pv(na);
  std::vector<uint64_t> cbase(na + 1, 0);
  for (uint32_t a = 0; a < na; a++) {
    const LselDesc& d = ls->desc[act[a]];
    pv[a] = PlaceVec{d.s1, d.words, (d.nbits + 63) >> 6, d.ones, cbase[a], d.alpha, d.beta};
    cbase[a + 1] = cbase[a] + wtg_ceil_div(d.ones, P.max_gap);
  }
  const uint64_t nflat = cbase[na];
  PlaceVec* d_pv = nullptr;
  uint64_t* d_cbase = nullptr;
  uint32_t *d_cnt = nullptr, *d_err = nullptr;
  WTG_TRY(alloc_block(ctx, ls, na * sizeof(PlaceVec), (void**)&d_pv));
  WTG_TRY(alloc_block(ctx, ls, (na + 1) * 8, (void**)&d_cbase));
  WTG_TRY(alloc_block(ctx, ls, (nflat + 1) * 4, (void**)&d_cnt));
  WTG_TRY(alloc_block(ctx, ls, 256, (void**)&d_err));
  WTG_CHECK_HIP(hipMemcpyAsync(d_pv, pv.data(), na * sizeof(PlaceVec), hipMemcpyHostToDevice, ctx->stream));
  WTG_CHECK_HIP(hipMemcpyAsync(d_cbase, cbase.data(), (na + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  WTG_TRY(wtg_fill_u64(ctx, d_err, 1, 0));
  Knot* d_scr = nullptr;  // the count pass keeps each chunk's first LSEL_CAP knots
  WTG_TRY(wtg_malloc(ctx, (void**)&d_scr, nflat * LSEL_CAP * sizeof(Knot) + 256));
  const unsigned gp = (unsigned)wtg_ceil_div(nflat, 4);  // a wave per chunk
  WTG_LAUNCH(ctx, "lsel_place", k_lsel_place, dim3(gp), dim3(256), 0, d_pv, d_cbase, na, nflat, gap_log,
             (double)P.tol, d_cnt, (const uint32_t*)nullptr, (Knot*)nullptr, d_scr, d_err);
  WTG_LAUNCH_CHECK();
  std::vector<uint32_t> off(nflat + 1, 0);
  WTG_TRY(sync_copy(ctx, off.data(), d_cnt, nflat * 4, hipMemcpyDeviceToHost));
  uint64_t tot = 0;
  for (uint64_t c = 0; c < nflat; c++) {
    const uint32_t k = off[c];
    off[c] = (uint32_t)tot;
    tot += k;
  }
  WTG_REQUIRE(tot < (1ull << 32), "wtg_lsel_build: too many knots");
  off[nflat] = (uint32_t)tot;
  const uint64_t nk_total = tot;
  Knot *d_K = nullptr, *d_K0 = nullptr;
  float *m1 = nullptr, *m2 = nullptr, *gl = nullptr, *gr = nullptr, *le = nullptr;
  uint64_t* d_kfirst = nullptr;
  WTG_TRY(alloc_block(ctx, ls, nk_total * sizeof(Knot) + 256, (void**)&d_K));
  WTG_TRY(alloc_block(ctx, ls, nk_total * sizeof(Knot) + 256, (void**)&d_K0));
  WTG_TRY(alloc_block(ctx, ls, nk_total * 20 + 256, (void**)&m1));
  m2 = m1 + nk_total;
  gl = m2 + nk_total;
  gr = gl + nk_total;
  le = gr + nk_total;
  WTG_TRY(alloc_block(ctx, ls, (na + 1) * 8, (void**)&d_kfirst));
  WTG_CHECK_HIP(hipMemcpyAsync(d_cnt, off.data(), (nflat + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
  WTG_LAUNCH(ctx, "lsel_place", k_lsel_place, dim3(gp), dim3(256), 0, d_pv, d_cbase, na, nflat, gap_log,
             (double)P.tol, (uint32_t*)nullptr, (const uint32_t*)d_cnt, d_K, d_scr, d_err);
  WTG_LAUNCH_CHECK();
  uint32_t perr = 0;
  const int prc = sync_copy(ctx, &perr, d_err, 4, hipMemcpyDeviceToHost);
  wtg_free(ctx, d_scr);
  WTG_TRY(prc);
  WTG_REQUIRE(perr == 0, "wtg_lsel_build: knot placement passes disagree");
  WTG_CHECK_HIP(hipMemcpyAsync(d_K0, d_K, nk_total * sizeof(Knot), hipMemcpyDeviceToDevice, ctx->stream));
  WTG_TRY(wtg_fill_u64(ctx, m1, nk_total * 5 / 2 + 1, 0));
  // 4. per vector: knots, lookup table
  std::vector<TrainVec> tv(na);
  std::vector<uint64_t> kfirst(na + 1, 0);
  uint64_t ns_total = 0, nsmax = 0, lut_total = 0;
  uint32_t nkmax = 2;
  std::vector<uint32_t> lut_logs(na);
  for (uint32_t a = 0; a < na; a++) {
    ns_total += wtg_align256(ns[act[a]] * 4) / 4;
    nsmax = std::max<uint64_t>(nsmax, ns[act[a]]);
    const uint64_t nka = off[cbase[a + 1]] - off[cbase[a]];
    const uint64_t ones = ls->desc[act[a]].ones;
    const uint64_t avg = ones / (nka > 1 ? nka - 1 : 1);
    uint32_t lg = 0;
    while (lg < 20 && (2ull << lg) <= avg) lg++;
    lut_logs[a] = lg;
    lut_total += wtg_align256(((ones >> lg) + 1) * 4) / 4;
    nkmax = std::max<uint32_t>(nkmax, (uint32_t)nka);
  }
  float* tgt = nullptr;
  WTG_TRY(alloc_block(ctx, ls, ns_total * 4, (void**)&tgt));
  uint32_t* lut = nullptr;
  WTG_TRY(alloc_block(ctx, ls, lut_total * 4, (void**)&lut));
  WTG_TRY(wtg_fill_bytes_zero(ctx, lut, lut_total * 4));  // every entry is written below; 0 keeps reads in range
  for (uint32_t a = 0; a < na; a++) {
    const uint32_t b = act[a];
    LselDesc& d = ls->desc[b];
    const uint64_t k0 = off[cbase[a]], k1 = off[cbase[a + 1]];
    kfirst[a] = k0;
    tv[a] = TrainVec{pos[b], tgt, ns[b], k0, (uint32_t)(k1 - k0), tlog, d.alpha, d.beta};
    tgt += wtg_align256(ns[b] * 4) / 4;
    d.knots = d_K + k0;
    d.nknots = (uint32_t)(k1 - k0);
    const uint32_t lg = lut_logs[a];
    d.lut_log = lg;
    d.lut = lut;
    lut += wtg_align256(((d.ones >> lg) + 1) * 4) / 4;
    d.learned = 1;
    ls->info[b].nknots = d.nknots;
    ls->info[b].lut_log = lg;
  }
  kfirst[na] = nk_total;
  // model tables: TrainVec[na] | eps[nvec] | loss[na] | act[na]
  const size_t o_eps = wtg_align256(na * sizeof(TrainVec)), o_loss = o_eps + wtg_align256(nvec * 8);
  const size_t o_act = o_loss + wtg_align256(na * 8);
  char* tabs = nullptr;
  WTG_TRY(alloc_block(ctx, ls, o_act + wtg_align256(na * 4), (void**)&tabs));
  TrainVec* d_tv = (TrainVec*)tabs;
  unsigned long long* d_eps = (unsigned long long*)(tabs + o_eps);
  double* d_loss = (double*)(tabs + o_loss);
  uint32_t* d_act = (uint32_t*)(tabs + o_act);
  WTG_CHECK_HIP(hipMemcpyAsync(d_tv, tv.data(), na * sizeof(TrainVec), hipMemcpyHostToDevice, ctx->stream));
  WTG_CHECK_HIP(hipMemcpyAsync(d_kfirst, kfirst.data(), (na + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  WTG_CHECK_HIP(hipMemcpyAsync(d_act, act.data(), na * 4, hipMemcpyHostToDevice, ctx->stream));
  WTG_TRY(lsel_upload_desc(ctx, ls));
  WTG_LAUNCH(ctx, "lsel_lut", k_lsel_lut, dim3((unsigned)std::min<uint64_t>(wtg_ceil_div(nkmax, 256), 1024), na),
             dim3(256), 0, ls->d_desc, d_act);
  WTG_LAUNCH_CHECK();
  TrainBufs tb{d_K, m1, m2, gl, gr, le, d_kfirst};
  const unsigned gk = (unsigned)wtg_ceil_div(nk_total, 256);
  TrainItems items{};
  {
    uint64_t cap = nk_total + 1;
    cap += nk_total;  // an interval's aligned items: at most one more than its samples / LSEL_ITEM + 1
    for (uint32_t a = 0; a < na; a++) cap += ns[act[a]] / LSEL_ITEM + 1;
    char* mem = nullptr;
    WTG_TRY(alloc_block(ctx, ls, train_items_bytes(nk_total, cap), (void**)&mem));
    WTG_TRY(train_items_setup(ctx, d_tv, na, tb, nk_total, cap, nsmax, mem, items));
  }
  // eps (exhaustive) and training loss of the knots now in d_K
  uint64_t wmax = 1;
  for (uint32_t a = 0; a < na; a++) wmax = std::max<uint64_t>(wmax, (nbits[act[a]] + 63) / 64);
  const unsigned ge_x = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(wtg_ceil_div(wmax, 256 * LSEL_EPS_W), 256));
  const unsigned ge_s = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(wtg_ceil_div(nsmax, 256), 256));
  auto measure = [&](std::vector<uint64_t>& eps, std::vector<double>& loss) -> int {
    WTG_TRY(wtg_fill_u64(ctx, d_eps, nvec, 0));
    WTG_TRY(train_grad(ctx, d_tv, tb, nk_total, items));
    // every model in one launch each: blocks of a vector stride over its words
    WTG_LAUNCH(ctx, "lsel_eps", k_lsel_eps_lb, dim3(ge_s, na), dim3(256), 0, ls->d_desc, d_act, d_tv, d_eps);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "lsel_eps", k_lsel_eps, dim3(ge_x, na), dim3(256), 0, ls->d_desc, d_act, d_eps);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "lsel_loss", k_lsel_loss, dim3(na), dim3(256), 0, d_tv, tb, d_loss);
    WTG_LAUNCH_CHECK();
    std::vector<unsigned long long> e(nvec);
    std::vector<double> l(na);
    WTG_CHECK_HIP(hipMemcpyAsync(e.data(), d_eps, nvec * 8, hipMemcpyDeviceToHost, ctx->stream));
    WTG_CHECK_HIP(hipMemcpyAsync(l.data(), d_loss, na * 8, hipMemcpyDeviceToHost, ctx->stream));
    WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t a = 0; a < na; a++) {
      eps[a] = e[act[a]];
      loss[a] = ns[act[a]] ? l[a] / (double)ns[act[a]] : 0.0;
    }
    return WTG_OK;
  };
  std::vector<uint64_t> eps0(na), eps1(na);
  std::vector<double> loss0(na), loss1(na);
  WTG_TRY(measure(eps0, loss0));
  // 5. Adam on every knot of every model
  const float b1 = 0.9f, b2 = 0.999f;
  float p1 = 1.f, p2 = 1.f;
  for (uint32_t e = 0; e < P.epochs; e++) {
    WTG_LAUNCH(ctx, "lsel_fwd_bwd", k_lsel_items, dim3((unsigned)wtg_ceil_div(items.cap, 256)), dim3(256), 0, d_tv,
               tb, nk_total, items);
    WTG_LAUNCH_CHECK();
    p1 *= b1;
    p2 *= b2;
    WTG_LAUNCH(ctx, "lsel_adam", k_lsel_adam, dim3(gk), dim3(256), 0, d_tv, na, tb, nk_total, items, P.lr, b1, b2,
               1.f - p1, 1.f - p2);
    WTG_LAUNCH_CHECK();
  }
  WTG_TRY(measure(eps1, loss1));
  // 6. keep the trained knots where eps <= eps_max, else restore the placed ones
  for (uint32_t a = 0; a < na; a++) {
    const uint32_t b = act[a];
    wtg_lsel_vec_info& I = ls->info[b];
    I.eps_init = eps0[a];
    I.loss_init = loss0[a];
    I.loss_final = loss1[a];
    I.trained = eps1[a] <= P.eps_max ? 1u : 0u;
    I.eps = I.trained ? eps1[a] : eps0[a];
    if (!I.trained) {
      WTG_CHECK_HIP(hipMemcpyAsync(d_K + tv[a].kfirst, d_K0 + tv[a].kfirst, tv[a].nknots * sizeof(Knot),
                                   hipMemcpyDeviceToDevice, ctx->stream));
    }
    ls->desc[b].learned = I.eps <= P.eps_max ? 1 : 0;
    I.learned = (uint32_t)ls->desc[b].learned;
  }
  WTG_TRY(lsel_upload_desc(ctx, ls));
  return WTG_OK;
}

// rank9 directories and select1 structures of all vectors, in three host
// reads: none between the vectors' rank passes, one for the ones counts, one
// for the long-superblock counts (after_ones runs between the first two, its
// device work queued ahead of the select counts)
static int lsel_plain_build(wtg_ctx* ctx, wtg_lsel* ls, const uint64_t* const* d_words, const uint64_t* nbits,
                            const std::function<int(const std::vector<uint64_t>&)>& after_ones) {
  const uint32_t nvec = ls->nvec;
  uint64_t dw = 0;
  size_t stride = 0, worst = 0;
  for (uint32_t b = 0; b < nvec; b++) {
    dw += wtg_align256(wtg_rank9_words(nbits[b]) * 8) / 8;
    stride = std::max(stride, wtg_align256(wtg_rank_select_ws_bytes(nbits[b])));
    worst += wtg_select_arena_worst(nbits[b]);
  }
  uint64_t* dirs = nullptr;
  WTG_TRY(alloc_block(ctx, ls, dw * 8, (void**)&dirs));
  for (uint32_t b = 0; b < nvec; b++) {
    ls->dirs[b] = dirs;
    dirs += wtg_align256(wtg_rank9_words(nbits[b]) * 8) / 8;
  }
  // per-vector workspaces, then the gathered counts
  WTG_TRY(wtg_ws_reserve(ctx, stride * nvec + (size_t)nvec * 8 + 256));
  char* ws = (char*)ctx->ws;
  uint64_t* d_cnt = (uint64_t*)(ws + stride * nvec);
  wtg_arena ar[3]{};
  WTG_TRY(alloc_block(ctx, ls, worst, (void**)&ar[0].base));
  ar[0].size = worst;
  std::vector<wtg_rs_pending> pend(nvec);
  for (uint32_t b = 0; b < nvec; b++)
    WTG_TRY(wtg_rank_select_begin(ctx, d_words[b], nbits[b], ls->dirs[b], true, &ls->sel[b], nullptr,
                                  WTG_ONES_DEFER, &pend[b], ws + stride * b, &ar[0]));
  std::vector<uint64_t> cnt(nvec);
  auto gather = [&](size_t off, uint32_t bytes) -> int {
    WTG_LAUNCH(ctx, "lsel_gather", k_lsel_gather, dim3((unsigned)wtg_ceil_div(nvec, 256)), dim3(256), 0, ws, stride,
               off, nvec, bytes, d_cnt);
    WTG_LAUNCH_CHECK();
    return sync_copy(ctx, cnt.data(), d_cnt, (size_t)nvec * 8, hipMemcpyDeviceToHost);
  };
  WTG_TRY(gather(WTG_RS_ONES_OFF, 8));
  WTG_TRY(after_ones(cnt));  // work that needs only the directories and the counts
  size_t after = 0;
  for (uint32_t b = 0; b < nvec; b++) after += wtg_select_arena_count(cnt[b]);
  WTG_TRY(alloc_block(ctx, ls, after, (void**)&ar[1].base));
  ar[1].size = after;
  for (uint32_t b = 0; b < nvec; b++) {
    pend[b].ar = &ar[1];
    WTG_TRY(wtg_rank_select_count(ctx, &pend[b], cnt[b]));
  }
  WTG_TRY(gather(WTG_RS_NL_OFF + 4, 4));  // the select1 structure's long count
  size_t pool = 0;
  for (uint32_t b = 0; b < nvec; b++) pool += wtg_select_arena_pool((uint32_t)cnt[b]);
  if (pool) {
    WTG_TRY(alloc_block(ctx, ls, pool, (void**)&ar[2].base));
    ar[2].size = pool;
  }
  for (uint32_t b = 0; b < nvec; b++) {
    pend[b].ar = &ar[2];
    const uint32_t nl[2] = {0, (uint32_t)cnt[b]};
    WTG_TRY(wtg_rank_select_end(ctx, &pend[b], nl));
  }
  return WTG_OK;
}

extern "C" int wtg_lsel_build(wtg_ctx* ctx, uint32_t nvec, const uint64_t* const* d_words,
                              const uint64_t* nbits, const wtg_lsel_params* prm, wtg_lsel** out) {
  WTG_REQUIRE(ctx && out && d_words && nbits && nvec >= 1 && nvec <= 256, "wtg_lsel_build: bad argument (1 <= nvec <= 256)");
  wtg_lsel_params P{16, 4096, 480, 50, 4.0f, 512};
  if (prm) P = *prm;
  auto pow2 = [](uint32_t x) { return x && !(x & (x - 1)); };
  WTG_REQUIRE(pow2(P.train_stride) && pow2(P.max_gap) && P.max_gap >= 2 && P.max_gap <= (1u << 24),
              "wtg_lsel_build: train_stride and max_gap must be powers of two, 2 <= max_gap <= 2^24");
  WTG_REQUIRE(P.eps_max >= 2 && P.eps_max <= 512 && P.tol >= 1 && P.tol + 1 < P.eps_max,
              "wtg_lsel_build: need 2 <= eps_max <= 512 and 1 <= tol < eps_max - 1");
  for (uint32_t b = 0; b < nvec; b++)
    WTG_REQUIRE(d_words[b] && nbits[b] >= 1 && nbits[b] < (1ull << 40), "wtg_lsel_build: empty vector");
  WTG_TRY(wtg_set_device(ctx));
  *out = nullptr;
  wtg_lsel* ls = new wtg_lsel();
  ls->nvec = nvec;
  ls->desc.resize(nvec);
  ls->dirs.assign(nvec, nullptr);
  ls->sel.resize(nvec);
  ls->info.resize(nvec);
  int rc = WTG_OK;
  const uint32_t tlog = (uint32_t)__builtin_ctz(P.train_stride);
  std::vector<uint64_t*> pos(nvec, nullptr);
  std::vector<uint64_t> ns(nvec, 0);
  std::vector<uint32_t> act;  // vectors that get a model
  // 2. (queued inside 1, once the ones counts are known) the training samples
  //    of every vector, one launch; after 1, one host read of the first and
  //    last samples gives the baseline lines
  uint64_t* d_ends = nullptr;
  std::vector<LselDesc> desc0;  // the first upload's source (alive past the plain build's next host read)
  auto start_samples = [&](const std::vector<uint64_t>& ones) -> int {
    for (uint32_t b = 0; b < nvec; b++) {
      LselDesc& d = ls->desc[b];
      std::memset(&d, 0, sizeof(d));
      d.words = d_words[b];
      d.dir = ls->dirs[b];
      d.nbits = nbits[b];
      d.ones = ones[b];
      d.nblk = ((nbits[b] + 63) / 64) >> 3;
    }
    WTG_TRY(wtg_malloc(ctx, (void**)&ls->d_desc, nvec * sizeof(LselDesc)));
    desc0 = ls->desc;  // (the select fields follow after 1)
    WTG_CHECK_HIP(hipMemcpyAsync(ls->d_desc, desc0.data(), nvec * sizeof(LselDesc), hipMemcpyHostToDevice,
                                 ctx->stream));
    WTG_TRY(alloc_block(ctx, ls, nvec * 16, (void**)&d_ends));
    std::vector<SampleVec> sv;
    uint64_t wmax = 1;
    for (uint32_t b = 0; b < nvec; b++) {
      ns[b] = ones[b] >> tlog;
      if (ns[b] < 2 || ones[b] >= (1ull << 32)) continue;  // too few ones (or u32 knots): plain select
      WTG_TRY(alloc_block(ctx, ls, ns[b] * 8, (void**)&pos[b]));
      act.push_back(b);
      sv.push_back(SampleVec{pos[b], ns[b]});
      wmax = std::max<uint64_t>(wmax, (nbits[b] + 63) / 64);
    }
    if (act.empty()) return WTG_OK;
    SampleVec* d_sv = nullptr;
    uint32_t* d_vids = nullptr;
    WTG_TRY(alloc_block(ctx, ls, sv.size() * sizeof(SampleVec), (void**)&d_sv));
    WTG_TRY(alloc_block(ctx, ls, act.size() * 4, (void**)&d_vids));
    WTG_CHECK_HIP(hipMemcpyAsync(d_sv, sv.data(), sv.size() * sizeof(SampleVec), hipMemcpyHostToDevice, ctx->stream));
    WTG_CHECK_HIP(hipMemcpyAsync(d_vids, act.data(), act.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    const unsigned gx = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(wtg_ceil_div(wmax, 256 * LSEL_EPS_W), 256));
    WTG_LAUNCH(ctx, "lsel_samples", k_lsel_samples, dim3(gx, (unsigned)act.size()), dim3(256), 0, ls->d_desc, d_vids,
               d_sv, tlog, d_ends);
    WTG_LAUNCH_CHECK();
    return WTG_OK;
  };
  // 1. the plain method: rank9 + select1 of every vector.  Each vector has its
  //    own workspace, so one host read brings the ones counts of all vectors
  //    and one the long counts; directories and select content sit in blocks.
  for (uint32_t b = 0; b < nvec; b++) std::memset(&ls->sel[b], 0, sizeof(wtg_select));
  if ((rc = lsel_plain_build(ctx, ls, d_words, nbits, start_samples)) != WTG_OK) goto fail;
  for (uint32_t b = 0; b < nvec; b++) {
    LselDesc& d = ls->desc[b];
    const wtg_select& s = ls->sel[b];
    d.s1 = SelQ{s.superblock, s.mini, s.long_index, s.long_pool, d_words[b], (nbits[b] + 63) / 64 + 1, 1};
    ls->info[b].nbits = nbits[b];
    ls->info[b].ones = s.arg_cnt;
  }
  if ((rc = lsel_upload_desc(ctx, ls)) != WTG_OK) goto fail;
  {
    std::vector<uint64_t> ends(2 * nvec, 0);
    if (!act.empty() && (rc = sync_copy(ctx, ends.data(), d_ends, nvec * 16, hipMemcpyDeviceToHost)) != WTG_OK)
      goto fail;
    for (const uint32_t b : act) {
      const double j0 = (double)P.train_stride, j1 = (double)(ns[b] << tlog);
      ls->desc[b].alpha = ((double)ends[2 * b + 1] - (double)ends[2 * b]) / (j1 - j0);
      ls->desc[b].beta = (double)ends[2 * b] - ls->desc[b].alpha * j0;
    }
  }
  if (!act.empty() && (rc = lsel_models(ctx, ls, P, nbits, pos, ns, act)) != WTG_OK) goto fail;
  if ((rc = wtg_finish(ctx)) != WTG_OK) goto fail;
  *out = ls;
  return WTG_OK;
fail:
  lsel_release(ctx, ls);
  return rc;
}

extern "C" int wtg_lsel_info(const wtg_lsel* ls, uint32_t vec, wtg_lsel_vec_info* out) {
  WTG_REQUIRE(ls && out && vec < ls->nvec, "wtg_lsel_info: bad argument");
  *out = ls->info[vec];
  return WTG_OK;
}

extern "C" int wtg_lsel_model(const wtg_lsel* ls, uint32_t vec, wtg_lsel_model_view* out) {
  WTG_REQUIRE(ls && out && vec < ls->nvec, "wtg_lsel_model: bad argument");
  const LselDesc& d = ls->desc[vec];
  std::memset(out, 0, sizeof(*out));
  out->learned = (uint32_t)d.learned;
  if (d.knots) {
    out->d_knots = d.knots;
    out->nknots = d.nknots;
    out->d_lut = d.lut;
    out->lut_log = d.lut_log;
    out->lut_size = (d.ones >> d.lut_log) + 1;
  }
  out->alpha = d.alpha;
  out->beta = d.beta;
  return WTG_OK;
}

static int lsel_query(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec, const uint64_t* d_j,
                      uint64_t nq, uint64_t* d_out, bool learned) {
  WTG_REQUIRE(ctx && ls && (nq == 0 || (d_vec && d_j && d_out)), "null argument");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  const unsigned g = (unsigned)std::min<uint64_t>(wtg_ceil_div(nq, 256), 8192);
  const size_t lds = ls->nvec * sizeof(LselDesc);
  if (learned)
    WTG_LAUNCH(ctx, "lsel_query", k_lsel_query<true>, dim3(g), dim3(256), lds, ls->d_desc, ls->nvec, d_vec, d_j,
               nq, d_out);
  else
    WTG_LAUNCH(ctx, "lsel_plain", k_lsel_query<false>, dim3(g), dim3(256), lds, ls->d_desc, ls->nvec, d_vec,
               d_j, nq, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

extern "C" int wtg_lsel_select_batch(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec,
                                     const uint64_t* d_j, uint64_t nq, uint64_t* d_out) {
  return lsel_query(ctx, ls, d_vec, d_j, nq, d_out, true);
}
extern "C" int wtg_lsel_plain_batch(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec,
                                    const uint64_t* d_j, uint64_t nq, uint64_t* d_out) {
  return lsel_query(ctx, ls, d_vec, d_j, nq, d_out, false);
}

extern "C" int wtg_lsel_sectors(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec, const uint64_t* d_j,
                                const uint64_t* d_ans, uint64_t nq, int learned, uint32_t* d_out) {
  WTG_REQUIRE(ctx && ls && (nq == 0 || (d_vec && d_j && d_ans && d_out)), "wtg_lsel_sectors: null argument");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  const dim3 g((unsigned)wtg_ceil_div(nq, 256));
  if (learned)
    WTG_LAUNCH(ctx, "lsel_sectors", k_lsel_sectors<true>, g, dim3(256), 0, ls->d_desc, ls->nvec, d_vec, d_j, d_ans,
               nq, d_out);
  else
    WTG_LAUNCH(ctx, "lsel_sectors", k_lsel_sectors<false>, g, dim3(256), 0, ls->d_desc, ls->nvec, d_vec, d_j, d_ans,
               nq, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

extern "C" int wtg_lsel_loss_grad(wtg_ctx* ctx, const void* d_knots, uint32_t nknots, uint32_t stride,
                                  double alpha, double beta, const uint64_t* d_pos, uint64_t ns, double* loss,
                                  float* d_grad) {
  auto pow2 = [](uint32_t x) { return x && !(x & (x - 1)); };
  WTG_REQUIRE(ctx && d_knots && d_pos && d_grad && loss && nknots >= 2 && pow2(stride),
              "wtg_lsel_loss_grad: bad argument");
  WTG_TRY(wtg_set_device(ctx));
  // scratch: TrainVec | kfirst | gl | gr | le | loss | targets | work items
  const uint64_t cap = 2 * (uint64_t)nknots + ns / LSEL_ITEM + 2;
  const size_t o_tv = 0, o_kf = wtg_align256(sizeof(TrainVec)), o_gl = o_kf + 256;
  const size_t o_gr = wtg_align256(o_gl + (size_t)nknots * 4), o_le = wtg_align256(o_gr + (size_t)nknots * 4);
  const size_t o_l = wtg_align256(o_le + (size_t)nknots * 4), o_tg = o_l + 256;
  const size_t o_it = wtg_align256(o_tg + (size_t)ns * 4);
  const size_t total = o_it + train_items_bytes(nknots, cap);
  WTG_TRY(wtg_ws_reserve(ctx, total));
  char* ws = (char*)ctx->ws;
  const TrainVec v{d_pos, (float*)(ws + o_tg), ns, 0, nknots, (uint32_t)__builtin_ctz(stride), alpha, beta};
  const uint64_t kf = 0;
  WTG_CHECK_HIP(hipMemcpyAsync(ws + o_tv, &v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
  WTG_CHECK_HIP(hipMemcpyAsync(ws + o_kf, &kf, 8, hipMemcpyHostToDevice, ctx->stream));
  const TrainVec* d_tv = (const TrainVec*)(ws + o_tv);
  TrainBufs tb{(Knot*)const_cast<void*>(d_knots), nullptr, nullptr, (float*)(ws + o_gl), (float*)(ws + o_gr),
               (float*)(ws + o_le), (const uint64_t*)(ws + o_kf)};
  const unsigned g = (unsigned)wtg_ceil_div(nknots, 256);
  TrainItems items{};
  WTG_TRY(train_items_setup(ctx, d_tv, 1u, tb, nknots, cap, ns, ws + o_it, items));
  WTG_TRY(train_grad(ctx, d_tv, tb, nknots, items));
  WTG_LAUNCH(ctx, "lsel_grad", k_lsel_grad, dim3(g), dim3(256), 0, d_tv, tb, d_grad);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_loss", k_lsel_loss, dim3(1), dim3(256), 0, d_tv, tb, (double*)(ws + o_l));
  WTG_LAUNCH_CHECK();
  double l = 0;
  WTG_CHECK_HIP(hipMemcpyAsync(&l, ws + o_l, 8, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *loss = ns ? l / (double)ns : 0;
  return wtg_finish(ctx);
}
