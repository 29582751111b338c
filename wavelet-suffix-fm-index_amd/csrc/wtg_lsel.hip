// wtg_lsel.hip -- learned select over a batch of bitvectors.
//
// Per bitvector: rank9 directory + select_support_mcl<1> samples (the plain
// method), then a piecewise-linear model of j -> select(j):
//   pred(j) = alpha*j + beta + (1-f)*y[k] + f*y[k+1],  k = j / S, f = (j mod S) / S
// alpha, beta: the line through the first and last training sample (fp64);
// y: fp32 residual knots every S arguments, trained on the GPU (forward +
// backward kernel, Adam) to minimise the mean squared error on the training
// samples select(T), select(2T), ... that the plain method answers.
// The error window eps = max_j |round(pred(j)) - select(j)| is then measured
// over EVERY one of the vector (exhaustive kernel).  Knot spacings are tried
// from the largest down; the first with eps <= eps_max is kept, otherwise the
// vector answers with the plain sampled select.
//
// Query: p = clamp(round(pred(j))); r = rank9(p); the answer is the (j-r)-th
// one at/after p, or the (r-j+1)-th one before p -- at most eps/64 + 1 words
// away.  The result is exact whatever the model predicts; eps only bounds the
// scan.  The plain sampled select stays available (wtg_lsel_plain_batch) as
// the judge.
#include "wtg_internal.hpp"
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

struct LselDesc {
  const uint64_t* words;
  const uint64_t* dir;
  SelQ s1;
  uint64_t nbits, ones;
  const float* y;        // nknots residual knots (learned vectors)
  double alpha, beta;
  uint32_t slog;         // log2 of the knot spacing
  uint32_t nknots;
  int32_t learned;
  int32_t _pad;
};

// the model's prediction as the position it starts from
__device__ __forceinline__ uint64_t lsel_predict(const LselDesc& d, uint64_t j) {
  const uint64_t k = j >> d.slog;
  const float f = (float)(j & ((1ull << d.slog) - 1)) / (float)(1ull << d.slog);
  const float s = (1.0f - f) * d.y[k] + f * d.y[k + 1];
  const double p = fma(d.alpha, (double)j, d.beta) + (double)s;
  const double hi = (double)(d.nbits - 1);
  return p <= 0.0 ? 0ull : p >= hi ? d.nbits - 1 : (uint64_t)llrint(p);
}

// exact select1(j) from a start position p: rank9 at p, then a word scan
// forward or backward to the j-th one
__device__ __forceinline__ uint64_t lsel_correct(const LselDesc& d, uint64_t j, uint64_t p) {
  const uint64_t r = rank9_dev(d.dir, d.words, p);
  const uint64_t W = (d.nbits + 63) >> 6;
  uint64_t w = p >> 6;
  const uint32_t o = (uint32_t)(p & 63);
  if (r < j) {
    uint64_t k = j - r;
    uint64_t x = d.words[w] & (~0ull << o);
    for (;;) {
      const uint32_t c = popc64(x);
      if (c >= k) return (w << 6) + select_in_word(x, (uint32_t)(k - 1));
      k -= c;
      if (++w >= W) return WTG_NONE;
      x = d.words[w];
    }
  } else {
    uint64_t k = r - j + 1;
    uint64_t x = o ? d.words[w] & ((1ull << o) - 1) : 0ull;
    for (;;) {
      const uint32_t c = popc64(x);
      if (c >= k) return (w << 6) + select_in_word(x, (uint32_t)(c - k));
      k -= c;
      if (w == 0) return WTG_NONE;
      x = d.words[--w];
    }
  }
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

// ------------------------------------------------------------- training
// One spline model's view for the training kernels (several models are
// trained together; `intervals` indexes a flat space of knot intervals).
struct TrainVec {
  const uint64_t* pos;   // ns training samples: pos[t] = select(T*(t+1))
  uint64_t ns;
  float* y;              // nknots parameters
  float* m1;             // Adam moments
  float* m2;
  float* gl;             // per interval: sum e*(1-f), sum e*f, sum e^2
  float* gr;
  float* le;
  uint64_t ibase;        // first flat interval index of this model
  double alpha, beta;
  uint32_t slog, tlog, nknots;
  uint32_t _pad;
};

__device__ __forceinline__ uint32_t find_vec(const TrainVec* tv, uint32_t nv, uint64_t flat) {
  uint32_t lo = 0, hi = nv;  // last model with ibase <= flat
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (tv[mid].ibase <= flat) lo = mid; else hi = mid;
  }
  return lo;
}

// forward + backward: thread per knot interval k; samples j in [kS, (k+1)S)
__global__ __launch_bounds__(256) void k_lsel_fwd_bwd(const TrainVec* __restrict__ tv, uint32_t nv,
                                                      uint64_t nint) {
  const uint64_t flat = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (flat >= nint) return;
  const uint32_t b = find_vec(tv, nv, flat);
  const TrainVec& v = tv[b];
  const uint64_t k = flat - v.ibase;
  if (k + 1 >= v.nknots) return;
  const uint64_t S = 1ull << v.slog, T = 1ull << v.tlog;
  // samples t with T*(t+1) in [kS, (k+1)S) and 1 <= t+1 <= ns
  uint64_t t1 = (k * S) >> v.tlog;
  if (t1 < 1) t1 = 1;
  uint64_t t2 = (((k + 1) * S) >> v.tlog) - 1;
  if (t2 > v.ns) t2 = v.ns;
  const float y0 = v.y[k], y1 = v.y[k + 1];
  float a = 0.f, c = 0.f, e2 = 0.f;
  for (uint64_t u = t1; u <= t2; u++) {
    const uint64_t j = u * T;
    const float f = (float)(j - k * S) / (float)S;
    const float target = (float)((double)v.pos[u - 1] - fma(v.alpha, (double)j, v.beta));
    const float e = (1.0f - f) * y0 + f * y1 - target;
    a += e * (1.0f - f);
    c += e * f;
    e2 += e * e;
  }
  v.gl[k] = a;
  v.gr[k] = c;
  v.le[k] = e2;
}

// Adam step on every knot: grad = 2/ns * (gl[k] + gr[k-1])
__global__ __launch_bounds__(256) void k_lsel_adam(const TrainVec* __restrict__ tv, uint32_t nv,
                                                   uint64_t nknots_total, const uint64_t* __restrict__ kbase,
                                                   float lr, float b1, float b2, float bc1, float bc2) {
  const uint64_t flat = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (flat >= nknots_total) return;
  uint32_t lo = 0, hi = nv;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (kbase[mid] <= flat) lo = mid; else hi = mid;
  }
  const TrainVec& v = tv[lo];
  const uint64_t k = flat - kbase[lo];
  if (k >= v.nknots) return;
  float g = 0.f;
  if (k + 1 < v.nknots) g += v.gl[k];
  if (k >= 1) g += v.gr[k - 1];
  g *= 2.0f / (float)(v.ns ? v.ns : 1);
  const float m = b1 * v.m1[k] + (1.f - b1) * g;
  const float s = b2 * v.m2[k] + (1.f - b2) * g * g;
  v.m1[k] = m;
  v.m2[k] = s;
  v.y[k] -= lr * (m / bc1) / (sqrtf(s / bc2) + 1e-6f);
}

// the gradient alone (for checks): grad[k] = 2/ns * (gl[k] + gr[k-1])
__global__ void k_lsel_grad(const TrainVec* __restrict__ tv, float* __restrict__ grad) {
  const TrainVec& v = tv[0];
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= v.nknots) return;
  float g = 0.f;
  if (k + 1 < v.nknots) g += v.gl[k];
  if (k >= 1) g += v.gr[k - 1];
  grad[k] = g * 2.0f / (float)(v.ns ? v.ns : 1);
}

// initial knots: the residual at the sample nearest to j = kS
__global__ void k_lsel_init(const TrainVec* __restrict__ tv, uint32_t nv, uint64_t nknots_total,
                            const uint64_t* __restrict__ kbase) {
  const uint64_t flat = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (flat >= nknots_total) return;
  uint32_t lo = 0, hi = nv;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (kbase[mid] <= flat) lo = mid; else hi = mid;
  }
  const TrainVec& v = tv[lo];
  const uint64_t k = flat - kbase[lo];
  if (k >= v.nknots) return;
  float y = 0.f;
  if (v.ns) {
    uint64_t t = (k << v.slog) >> v.tlog;  // sample index + 1 at j = kS
    if (t < 1) t = 1;
    if (t > v.ns) t = v.ns;
    const uint64_t j = t << v.tlog;
    y = (float)((double)v.pos[t - 1] - fma(v.alpha, (double)j, v.beta));
  }
  v.y[k] = y;
  v.m1[k] = 0.f;
  v.m2[k] = 0.f;
}

// exhaustive error window: thread per word, every one of the word
__global__ __launch_bounds__(256) void k_lsel_eps(const LselDesc* __restrict__ descs, uint32_t b,
                                                  unsigned long long* __restrict__ eps) {
  const LselDesc d = descs[b];
  const uint64_t W = (d.nbits + 63) >> 6;
  uint64_t mx = 0;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = d.words[w];
    if (!x) continue;
    uint64_t j = rank9_dev(d.dir, d.words, w << 6);
    while (x) {
      const uint32_t o = (uint32_t)__builtin_ctzll(x);
      x &= x - 1;
      j++;
      const uint64_t p = (w << 6) + o;
      const uint64_t q = lsel_predict(d, j);
      const uint64_t e = q > p ? q - p : p - q;
      mx = e > mx ? e : mx;
    }
  }
  mx = wave_max_u64(mx);
  if (lane_id() == 0 && mx) atomicMax(&eps[b], (unsigned long long)mx);
}

// sum of squared residuals of one model (training loss numerator)
__global__ __launch_bounds__(256) void k_lsel_loss(const TrainVec* __restrict__ tv, uint32_t b,
                                                   double* __restrict__ out) {
  const TrainVec& v = tv[b];
  double acc = 0;
  for (uint64_t k = threadIdx.x; k + 1 < v.nknots; k += blockDim.x) acc += (double)v.le[k];
  __shared__ double s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = s[0];
}

// training samples: pos[t] = select1(T*(t+1)) by the plain sampled select
__global__ __launch_bounds__(256) void k_lsel_samples(const LselDesc* __restrict__ descs, uint32_t b,
                                                      uint32_t tlog, uint64_t ns, uint64_t* __restrict__ pos) {
  const LselDesc& d = descs[b];
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ns;
       t += (uint64_t)gridDim.x * blockDim.x)
    pos[t] = select_plain(d.s1, (t + 1) << tlog);
}

}  // namespace

struct wtg_lsel {
  uint32_t nvec = 0;
  std::vector<LselDesc> desc;          // host copy
  LselDesc* d_desc = nullptr;
  std::vector<uint64_t*> dirs;         // owned rank directories
  std::vector<wtg_select> sel;         // owned select1 structures
  std::vector<wtg_lsel_vec_info> info;
  std::vector<void*> blocks;           // owned model / sample memory
};

static void lsel_release(wtg_ctx* ctx, wtg_lsel* ls) {
  if (!ls) return;
  wtg_free(ctx, ls->d_desc);
  for (auto* d : ls->dirs) wtg_free(ctx, d);
  for (auto& s : ls->sel) wtg_select_free(ctx, &s);
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

extern "C" int wtg_lsel_build(wtg_ctx* ctx, uint32_t nvec, const uint64_t* const* d_words,
                              const uint64_t* nbits, const wtg_lsel_params* prm, wtg_lsel** out) {
  WTG_REQUIRE(ctx && out && d_words && nbits && nvec >= 1 && nvec <= 256, "wtg_lsel_build: bad argument (1 <= nvec <= 256)");
  wtg_lsel_params P{16, 4096, 64, 200, 4.0f, 512};
  if (prm) P = *prm;
  auto pow2 = [](uint32_t x) { return x && !(x & (x - 1)); };
  WTG_REQUIRE(pow2(P.train_stride) && pow2(P.spacing_max) && pow2(P.spacing_min) &&
                  P.spacing_min >= 2 * P.train_stride && P.spacing_max >= P.spacing_min,
              "wtg_lsel_build: spacings and stride must be powers of two, spacing_min >= 2*stride");
  for (uint32_t b = 0; b < nvec; b++) WTG_REQUIRE(d_words[b] && nbits[b] >= 1, "wtg_lsel_build: empty vector");
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
  unsigned long long* d_eps = nullptr;
  double* d_loss = nullptr;
  TrainVec* d_tv = nullptr;
  uint64_t* d_kbase = nullptr;
  // 1. the plain method: rank9 + select1 per vector
  for (uint32_t b = 0; b < nvec && rc == WTG_OK; b++) {
    std::memset(&ls->sel[b], 0, sizeof(wtg_select));
    if ((rc = wtg_malloc(ctx, (void**)&ls->dirs[b], wtg_rank9_words(nbits[b]) * 8)) != WTG_OK) break;
    if ((rc = wtg_rank_select_run(ctx, d_words[b], nbits[b], ls->dirs[b], true, &ls->sel[b], nullptr)) != WTG_OK)
      break;
    LselDesc& d = ls->desc[b];
    std::memset(&d, 0, sizeof(d));
    d.words = d_words[b];
    d.dir = ls->dirs[b];
    const wtg_select& s = ls->sel[b];
    d.s1 = SelQ{s.superblock, s.mini, s.long_index, s.long_pool, d_words[b], (nbits[b] + 63) / 64 + 1, 1};
    d.nbits = nbits[b];
    d.ones = s.arg_cnt;
    ls->info[b].nbits = nbits[b];
    ls->info[b].ones = s.arg_cnt;
  }
  if (rc != WTG_OK) goto fail;
  if ((rc = wtg_malloc(ctx, (void**)&ls->d_desc, nvec * sizeof(LselDesc))) != WTG_OK) goto fail;
  if ((rc = lsel_upload_desc(ctx, ls)) != WTG_OK) goto fail;
  // 2. training samples and the baseline line
  for (uint32_t b = 0; b < nvec; b++) {
    ns[b] = ls->desc[b].ones >> tlog;
    if (ns[b] < 2) continue;
    if ((rc = alloc_block(ctx, ls, ns[b] * 8, (void**)&pos[b])) != WTG_OK) goto fail;
    const unsigned g = (unsigned)std::min<uint64_t>(wtg_ceil_div(ns[b], 256), 4096);
    WTG_LAUNCH(ctx, "lsel_samples", k_lsel_samples, dim3(g), dim3(256), 0, ls->d_desc, b, tlog, ns[b], pos[b]);
    WTG_LAUNCH_CHECK();
    uint64_t first = 0, last = 0;
    if (hipMemcpyAsync(&first, pos[b], 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(&last, pos[b] + ns[b] - 1, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      wtg_set_error("wtg_lsel_build: sample copy failed");
      rc = WTG_ERR_HIP;
      goto fail;
    }
    const double j0 = (double)P.train_stride, j1 = (double)(ns[b] << tlog);
    ls->desc[b].alpha = ((double)last - (double)first) / (j1 - j0);
    ls->desc[b].beta = (double)first - ls->desc[b].alpha * j0;
  }
  if ((rc = alloc_block(ctx, ls, nvec * 8, (void**)&d_eps)) != WTG_OK) goto fail;
  if ((rc = alloc_block(ctx, ls, nvec * 8, (void**)&d_loss)) != WTG_OK) goto fail;
  if ((rc = alloc_block(ctx, ls, nvec * sizeof(TrainVec), (void**)&d_tv)) != WTG_OK) goto fail;
  if ((rc = alloc_block(ctx, ls, (nvec + 1) * 8, (void**)&d_kbase)) != WTG_OK) goto fail;
  {
    // 3. spacing search: train every still-open vector at spacing S, measure eps
    std::vector<char> done(nvec, 0);
    for (uint32_t b = 0; b < nvec; b++) done[b] = ns[b] < 2;  // too few ones: plain select
    for (uint32_t S = P.spacing_max; S >= P.spacing_min; S >>= 1) {
      std::vector<uint32_t> act;
      for (uint32_t b = 0; b < nvec; b++)
        if (!done[b]) act.push_back(b);
      if (act.empty()) break;
      const uint32_t slog = (uint32_t)__builtin_ctz(S);
      std::vector<TrainVec> tv(act.size());
      std::vector<uint64_t> kbase(act.size() + 1, 0);
      uint64_t nint = 0, nk = 0;
      for (size_t a = 0; a < act.size(); a++) {
        const uint32_t b = act[a];
        TrainVec& v = tv[a];
        std::memset(&v, 0, sizeof(v));
        v.nknots = (uint32_t)((ls->desc[b].ones >> slog) + 2);
        v.ibase = nint;
        kbase[a] = nk;
        nint += v.nknots - 1;
        nk += v.nknots;
      }
      kbase[act.size()] = nk;
      // one block per round: y | m1 | m2 | gl | gr | le
      float* blk = nullptr;
      if ((rc = alloc_block(ctx, ls, (size_t)nk * 6 * 4 + 256, (void**)&blk)) != WTG_OK) goto fail;
      for (size_t a = 0; a < act.size(); a++) {
        const uint32_t b = act[a];
        TrainVec& v = tv[a];
        float* base = blk + kbase[a] * 6;
        v.y = base;
        v.m1 = base + v.nknots;
        v.m2 = base + 2 * (uint64_t)v.nknots;
        v.gl = base + 3 * (uint64_t)v.nknots;
        v.gr = base + 4 * (uint64_t)v.nknots;
        v.le = base + 5 * (uint64_t)v.nknots;
        v.pos = pos[b];
        v.ns = ns[b];
        v.alpha = ls->desc[b].alpha;
        v.beta = ls->desc[b].beta;
        v.slog = slog;
        v.tlog = tlog;
      }
      const uint32_t na = (uint32_t)act.size();
      WTG_CHECK_HIP(hipMemcpyAsync(d_tv, tv.data(), na * sizeof(TrainVec), hipMemcpyHostToDevice, ctx->stream));
      WTG_CHECK_HIP(hipMemcpyAsync(d_kbase, kbase.data(), (na + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
      const unsigned gk = (unsigned)wtg_ceil_div(nk, 256), gi = (unsigned)wtg_ceil_div(nint, 256);
      WTG_LAUNCH(ctx, "lsel_init", k_lsel_init, dim3(gk), dim3(256), 0, d_tv, na, nk, d_kbase);
      WTG_LAUNCH_CHECK();
      // eps and loss of the initial (untrained) knots
      auto measure = [&](std::vector<uint64_t>& eps, std::vector<double>& loss) -> int {
        for (size_t a = 0; a < act.size(); a++) {
          LselDesc& d = ls->desc[act[a]];
          d.y = tv[a].y;
          d.slog = slog;
          d.nknots = tv[a].nknots;
        }
        WTG_TRY(lsel_upload_desc(ctx, ls));
        WTG_TRY(wtg_fill_u64(ctx, d_eps, nvec, 0));
        WTG_LAUNCH(ctx, "lsel_fwd_bwd", k_lsel_fwd_bwd, dim3(gi), dim3(256), 0, d_tv, na, nint);
        WTG_LAUNCH_CHECK();
        for (size_t a = 0; a < act.size(); a++) {
          const uint32_t b = act[a];
          const unsigned ge = (unsigned)std::min<uint64_t>(wtg_ceil_div((nbits[b] + 63) / 64, 256), 2048);
          WTG_LAUNCH(ctx, "lsel_eps", k_lsel_eps, dim3(ge), dim3(256), 0, ls->d_desc, b, d_eps);
          WTG_LAUNCH_CHECK();
          WTG_LAUNCH(ctx, "lsel_loss", k_lsel_loss, dim3(1), dim3(256), 0, d_tv, (uint32_t)a, d_loss);
          WTG_LAUNCH_CHECK();
          WTG_CHECK_HIP(hipMemcpyAsync(&loss[a], d_loss + a, 8, hipMemcpyDeviceToHost, ctx->stream));
        }
        std::vector<unsigned long long> e(nvec);
        WTG_CHECK_HIP(hipMemcpyAsync(e.data(), d_eps, nvec * 8, hipMemcpyDeviceToHost, ctx->stream));
        WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t a = 0; a < act.size(); a++) eps[a] = e[act[a]];
        return WTG_OK;
      };
      std::vector<uint64_t> eps0(na), eps1(na);
      std::vector<double> loss0(na), loss1(na);
      if ((rc = measure(eps0, loss0)) != WTG_OK) goto fail;
      const float b1 = 0.9f, b2 = 0.999f;
      float p1 = 1.f, p2 = 1.f;
      for (uint32_t e = 0; e < P.epochs; e++) {
        WTG_LAUNCH(ctx, "lsel_fwd_bwd", k_lsel_fwd_bwd, dim3(gi), dim3(256), 0, d_tv, na, nint);
        WTG_LAUNCH_CHECK();
        p1 *= b1;
        p2 *= b2;
        WTG_LAUNCH(ctx, "lsel_adam", k_lsel_adam, dim3(gk), dim3(256), 0, d_tv, na, nk, d_kbase, P.lr, b1, b2,
                   1.f - p1, 1.f - p2);
        WTG_LAUNCH_CHECK();
      }
      if ((rc = measure(eps1, loss1)) != WTG_OK) goto fail;
      for (size_t a = 0; a < act.size(); a++) {
        const uint32_t b = act[a];
        wtg_lsel_vec_info& I = ls->info[b];
        const bool last_try = (S >> 1) < P.spacing_min;
        if (eps1[a] <= P.eps_max || last_try) {
          I.eps = eps1[a];
          I.eps_init = eps0[a];
          I.loss_init = ns[b] ? loss0[a] / (double)ns[b] : 0;
          I.loss_final = ns[b] ? loss1[a] / (double)ns[b] : 0;
          I.nknots = tv[a].nknots;
          I.spacing = eps1[a] <= P.eps_max ? S : 0;
          ls->desc[b].learned = eps1[a] <= P.eps_max ? 1 : 0;
          done[b] = 1;
        } else {
          ls->desc[b].learned = 0;
        }
      }
    }
    for (uint32_t b = 0; b < nvec; b++)
      if (!ls->desc[b].learned) ls->info[b].spacing = 0;
    if ((rc = lsel_upload_desc(ctx, ls)) != WTG_OK) goto fail;
  }
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

extern "C" int wtg_lsel_loss_grad(wtg_ctx* ctx, const float* d_y, uint32_t nknots, uint32_t spacing,
                                  uint32_t stride, double alpha, double beta, const uint64_t* d_pos, uint64_t ns,
                                  double* loss, float* d_grad) {
  auto pow2 = [](uint32_t x) { return x && !(x & (x - 1)); };
  WTG_REQUIRE(ctx && d_y && d_pos && d_grad && loss && nknots >= 2 && pow2(spacing) && pow2(stride) &&
                  spacing >= stride,
              "wtg_lsel_loss_grad: bad argument");
  WTG_REQUIRE(((ns * (uint64_t)stride) >> __builtin_ctz(spacing)) + 2 <= nknots,
              "wtg_lsel_loss_grad: knots do not cover the samples");
  WTG_TRY(wtg_set_device(ctx));
  // scratch: TrainVec | gl | gr | le | loss
  const size_t o_tv = 0, o_gl = wtg_align256(sizeof(TrainVec));
  const size_t o_gr = wtg_align256(o_gl + (size_t)nknots * 4), o_le = wtg_align256(o_gr + (size_t)nknots * 4);
  const size_t o_l = wtg_align256(o_le + (size_t)nknots * 4), total = o_l + 256;
  WTG_TRY(wtg_ws_reserve(ctx, total));
  char* ws = (char*)ctx->ws;
  TrainVec v{};
  v.pos = d_pos;
  v.ns = ns;
  v.y = const_cast<float*>(d_y);
  v.gl = (float*)(ws + o_gl);
  v.gr = (float*)(ws + o_gr);
  v.le = (float*)(ws + o_le);
  v.ibase = 0;
  v.alpha = alpha;
  v.beta = beta;
  v.slog = (uint32_t)__builtin_ctz(spacing);
  v.tlog = (uint32_t)__builtin_ctz(stride);
  v.nknots = nknots;
  WTG_CHECK_HIP(hipMemcpyAsync(ws + o_tv, &v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
  const TrainVec* d_tv = (const TrainVec*)(ws + o_tv);
  WTG_LAUNCH(ctx, "lsel_fwd_bwd", k_lsel_fwd_bwd, dim3((unsigned)wtg_ceil_div(nknots - 1, 256)), dim3(256), 0,
             d_tv, 1u, (uint64_t)(nknots - 1));
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_grad", k_lsel_grad, dim3((unsigned)wtg_ceil_div(nknots, 256)), dim3(256), 0, d_tv, d_grad);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "lsel_loss", k_lsel_loss, dim3(1), dim3(256), 0, d_tv, 0u, (double*)(ws + o_l));
  WTG_LAUNCH_CHECK();
  double l = 0;
  WTG_CHECK_HIP(hipMemcpyAsync(&l, ws + o_l, 8, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *loss = ns ? l / (double)ns : 0;
  return wtg_finish(ctx);
}
