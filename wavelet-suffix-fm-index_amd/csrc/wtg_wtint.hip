// wtg_wtint.hip -- sdsl::wt_int<> on the GPU: construction
// (wt_int/wt_int_par.hpp:368-406: tree bits, rank_support_v, select_support_mcl
// for 1 and 0 over the concatenated tree) and batched queries, one lane per
// query (rank :553-577, access :519-541, inverse_select :588-611,
// select :622-660).  Queries are chains of dependent random accesses, so the
// kernels keep many independent waves in flight rather than using LDS.
#include "wtg_internal.hpp"
#include <cstring>

struct wtg_wtint {
  uint64_t n = 0, sigma = 0, nbits = 0;
  uint32_t L = 0;
  uint64_t* words = nullptr;
  uint64_t* dir = nullptr;
  wtg_select s1{}, s0{};
};

struct WtQ {
  const uint64_t* words;
  const uint64_t* D;
  SelQ s1, s0;
  uint64_t n;
  uint32_t L;
};

__global__ void k_wtint_rank(WtQ w, const uint64_t* __restrict__ qi, const uint64_t* __restrict__ qc,
                             uint64_t nq, uint64_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  uint64_t i = qi[q];
  const uint64_t c = qc[q];
  if (i > w.n) { out[q] = WTG_NONE; return; }
  if (w.L < 64 && (1ull << w.L) <= c) { out[q] = 0; return; }
  uint64_t offset = 0, mask = 1ull << (w.L - 1), node_size = w.n;
  for (uint32_t k = 0; k < w.L && i; k++) {
    const uint64_t o = rank9_dev(w.D, w.words, offset);
    const uint64_t oi = rank9_dev(w.D, w.words, offset + i) - o;
    const uint64_t oe = rank9_dev(w.D, w.words, offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; i = oi; }
    else { node_size -= oe; i -= oi; }
    offset += w.n;
    mask >>= 1;
  }
  out[q] = i;
}

// lex_count traversal (wt_int/wt_int_par.hpp:712-741): rank(i,c) and the number
// of values below / above c in [i, j)
__device__ __forceinline__ void wtint_lex(const WtQ& w, uint64_t i, uint64_t j, uint64_t c, uint64_t& rk,
                                          uint64_t& sm, uint64_t& gr) {
  sm = gr = 0;
  if (w.L < 64 && (1ull << w.L) <= c) { rk = 0; sm = j - i; return; }
  uint64_t offset = 0, mask = 1ull << (w.L - 1), node_size = w.n;
  for (uint32_t k = 0; k < w.L; k++) {
    const uint64_t o = rank9_dev(w.D, w.words, offset);
    const uint64_t oi = rank9_dev(w.D, w.words, offset + i) - o;
    const uint64_t oj = rank9_dev(w.D, w.words, offset + j) - o;
    const uint64_t oe = rank9_dev(w.D, w.words, offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; sm += (j - i) - (oj - oi); i = oi; j = oj; }
    else { node_size -= oe; gr += oj - oi; i -= oi; j -= oj; }
    offset += w.n;
    mask >>= 1;
  }
  rk = i;
}

__global__ void k_wtint_lex_count(WtQ w, const uint64_t* __restrict__ qi, const uint64_t* __restrict__ qj,
                                  const uint64_t* __restrict__ qc, uint64_t nq, uint64_t* __restrict__ out_rank,
                                  uint64_t* __restrict__ out_smaller, uint64_t* __restrict__ out_greater) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint64_t i = qi[q], j = qj[q];
  uint64_t rk = WTG_NONE, sm = WTG_NONE, gr = WTG_NONE;
  if (i <= j && j <= w.n) wtint_lex(w, i, j, qc[q], rk, sm, gr);
  out_rank[q] = rk;
  out_smaller[q] = sm;
  out_greater[q] = gr;
}

// lex_smaller_count (wt_int/wt_int_par.hpp:753-780): rank(i,c) and the number of
// values below c in [0, i)
__global__ void k_wtint_lex_smaller(WtQ w, const uint64_t* __restrict__ qi, const uint64_t* __restrict__ qc,
                                    uint64_t nq, uint64_t* __restrict__ out_rank, uint64_t* __restrict__ out_smaller) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  uint64_t i = qi[q];
  const uint64_t c = qc[q];
  if (i > w.n) { out_rank[q] = out_smaller[q] = WTG_NONE; return; }
  if (w.L < 64 && (1ull << w.L) <= c) { out_rank[q] = 0; out_smaller[q] = i; return; }
  uint64_t offset = 0, mask = 1ull << (w.L - 1), node_size = w.n, res = 0;
  for (uint32_t k = 0; k < w.L && i; k++) {
    const uint64_t o = rank9_dev(w.D, w.words, offset);
    const uint64_t oi = rank9_dev(w.D, w.words, offset + i) - o;
    const uint64_t oe = rank9_dev(w.D, w.words, offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; res += i - oi; i = oi; }
    else { node_size -= oe; i -= oi; }
    offset += w.n;
    mask >>= 1;
  }
  out_rank[q] = i;
  out_smaller[q] = res;
}

// range_search_2d count (wt_int/wt_int_par.hpp:801-870, report = false): the
// positions in [lb, rb] with values in [vlb, vrb], as the difference of two
// lex_count "smaller" counts over [lb, rb + 1).  The reference clamps vrb to 2^L
// and its leaf test does not compare vlb, so vlb = 2^L counts the value 2^L - 1:
// vlb is clamped to 2^L - 1 the same way.
__global__ void k_wtint_range_count(WtQ w, const uint64_t* __restrict__ qlb, const uint64_t* __restrict__ qrb,
                                    const uint64_t* __restrict__ qvlb, const uint64_t* __restrict__ qvrb,
                                    uint64_t nq, uint64_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint64_t lb = qlb[q], rb = qrb[q];
  uint64_t vlb = qvlb[q], vrb = qvrb[q];
  if (rb >= w.n) { out[q] = WTG_NONE; return; }
  const uint64_t top = w.L < 64 ? (1ull << w.L) : ~0ull;
  if (vrb > top) vrb = top;
  if (vlb > vrb || lb > rb) { out[q] = 0; return; }
  if (w.L < 64 && vlb > top - 1) vlb = top - 1;
  uint64_t rk, gr, below_hi, below_lo;
  if (w.L < 64 ? vrb >= top - 1 : vrb == ~0ull) below_hi = rb + 1 - lb;   // every value is <= vrb
  else wtint_lex(w, lb, rb + 1, vrb + 1, rk, below_hi, gr);
  wtint_lex(w, lb, rb + 1, vlb, rk, below_lo, gr);
  out[q] = below_hi - below_lo;
}

__global__ void k_wtint_access(WtQ w, const uint64_t* __restrict__ qi, uint64_t nq,
                               uint64_t* __restrict__ out_sym, uint64_t* __restrict__ out_rank) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  uint64_t i = qi[q];
  if (i >= w.n) {
    out_sym[q] = WTG_NONE;
    if (out_rank) out_rank[q] = WTG_NONE;
    return;
  }
  uint64_t offset = 0, res = 0, node_size = w.n;
  for (uint32_t k = 0; k < w.L; k++) {
    res <<= 1;
    const uint64_t o = rank9_dev(w.D, w.words, offset);
    const uint64_t oi = rank9_dev(w.D, w.words, offset + i) - o;
    const uint64_t oe = rank9_dev(w.D, w.words, offset + node_size) - o;
    if (bit_dev(w.words, offset + i)) { offset += node_size - oe; node_size = oe; i = oi; res |= 1; }
    else { node_size -= oe; i -= oi; }
    offset += w.n;
  }
  out_sym[q] = res;
  if (out_rank) out_rank[q] = i;
}

template <uint32_t MAXL>  // the tree's levels are at most MAXL
__global__ void k_wtint_select(WtQ w, const uint64_t* __restrict__ qi, const uint64_t* __restrict__ qc,
                               uint64_t nq, uint64_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  uint64_t i = qi[q];
  const uint64_t c = qc[q];
  uint64_t path_off[MAXL + 1], path_rank[MAXL + 1];
  uint64_t offset = 0, mask = 1ull << (w.L - 1), node_size = w.n;
  path_off[0] = path_rank[0] = 0;
  for (uint32_t k = 0; k < w.L && node_size; k++) {
    const uint64_t o = rank9_dev(w.D, w.words, offset);
    path_rank[k] = o;
    const uint64_t oe = rank9_dev(w.D, w.words, offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; }
    else node_size -= oe;
    offset += w.n;
    path_off[k + 1] = offset;
    mask >>= 1;
  }
  if (i == 0 || node_size == 0 || node_size < i || (w.L < 64 && (c >> w.L) != 0)) {
    out[q] = WTG_NONE;  // the reference throws std::logic_error here
    return;
  }
  mask = 1;
  for (uint32_t k = w.L; k > 0; k--) {
    offset = path_off[k - 1];
    const uint64_t o = path_rank[k - 1];
    if (c & mask) i = select_plain(w.s1, o + i) - offset + 1;
    else i = select_plain(w.s0, offset - o + i) - offset + 1;
    mask <<= 1;
  }
  out[q] = i - 1;
}

static WtQ make_q(const wtg_wtint* wt) {
  WtQ w{};
  w.words = wt->words;
  w.D = wt->dir;
  const uint64_t W = (wt->nbits + 63) / 64;
  w.s1 = SelQ{wt->s1.superblock, wt->s1.mini, wt->s1.long_index, wt->s1.long_pool, wt->words, W, 1};
  w.s0 = SelQ{wt->s0.superblock, wt->s0.mini, wt->s0.long_index, wt->s0.long_pool, wt->words, W, 0};
  w.n = wt->n;
  w.L = wt->L;
  return w;
}

extern "C" int wtg_wtint_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n,
                               uint32_t max_level, wtg_wtint** out) {
  WTG_REQUIRE(ctx && out && (n == 0 || d_sym), "wtg_wtint_build: null argument");
  WTG_REQUIRE(sym_bytes == 1 || sym_bytes == 2 || sym_bytes == 4 || sym_bytes == 8, "bad sym_bytes");
  WTG_TRY(wtg_set_device(ctx));
  *out = nullptr;
  wtg_wtint* wt = new wtg_wtint();
  wt->n = n;
  int rc = WTG_OK;
  uint32_t L = max_level;
  if (n == 0) {  // wt_int_par.hpp:373-374: empty tree
    wt->L = max_level;
    *out = wt;
    return WTG_OK;
  }
  if (L == 0) {  // :384-392: max_level = hi(max(1, max value)) + 1
    uint64_t mx = 0;
    if ((rc = wtg_max_value(ctx, d_sym, sym_bytes, n, &mx)) != WTG_OK) { delete wt; return rc; }
    if (mx < 1) mx = 1;
    L = (uint32_t)wtg_hi_host(mx) + 1;
  }
  wt->L = L;
  wt->nbits = n * L;
  const uint64_t nw = wtg_levels_words(n, L);
  unsigned long long* d_distinct = (unsigned long long*)((char*)ctx->d_flag + 128);
  if ((rc = wtg_malloc(ctx, (void**)&wt->words, nw * 8)) != WTG_OK) goto fail;
  if ((rc = wtg_malloc(ctx, (void**)&wt->dir, wtg_rank9_words(wt->nbits) * 8)) != WTG_OK) goto fail;
  if ((rc = wtg_levels_build_async(ctx, d_sym, sym_bytes, n, L, wt->words, nullptr, 0, d_distinct)) != WTG_OK)
    goto fail;
  // the distinct count rides the stream into pinned memory; the support build
  // right behind the tree reads its count of ones with a sync, after which the
  // count is there (no sync of its own, no idle gap between tree and supports)
  if (hipMemcpyAsync(ctx->h_stage + WTG_STAGE_AUX, d_distinct, 8, hipMemcpyDeviceToHost, ctx->stream) !=
      hipSuccess) {
    wtg_set_error("wtg_wtint_build: copy failed");
    rc = WTG_ERR_HIP;
    goto fail;
  }
  if ((rc = wtg_rank_select_run(ctx, wt->words, wt->nbits, wt->dir, true, &wt->s1, &wt->s0)) != WTG_OK)
    goto fail;
  if ((rc = wtg_finish(ctx)) != WTG_OK) goto fail;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {  // (a no-op wait: the supports synced)
    wtg_set_error("wtg_wtint_build: sync failed");
    rc = WTG_ERR_HIP;
    goto fail;
  }
  wt->sigma = *reinterpret_cast<volatile unsigned long long*>(ctx->h_stage + WTG_STAGE_AUX);
  *out = wt;
  return WTG_OK;
fail:
  wtg_wtint_destroy(ctx, wt);
  return rc;
}

extern "C" int wtg_wtint_from_tree(wtg_ctx* ctx, const uint64_t* d_tree, uint64_t n, uint32_t levels,
                                   uint64_t sigma, wtg_wtint** out) {
  WTG_REQUIRE(ctx && out && (n == 0 || d_tree), "wtg_wtint_from_tree: null argument");
  WTG_REQUIRE(levels <= 64, "wtg_wtint_from_tree: levels > 64");
  WTG_TRY(wtg_set_device(ctx));
  *out = nullptr;
  wtg_wtint* wt = new wtg_wtint();
  wt->n = n;
  wt->L = levels;
  wt->sigma = sigma;
  if (n == 0 || levels == 0) {
    *out = wt;
    return WTG_OK;
  }
  wt->nbits = n * levels;
  const uint64_t nw = wtg_levels_words(n, levels);
  int rc = WTG_OK;
  if ((rc = wtg_malloc(ctx, (void**)&wt->words, (nw + 1) * 8)) != WTG_OK) goto fail;
  if ((rc = wtg_malloc(ctx, (void**)&wt->dir, wtg_rank9_words(wt->nbits) * 8)) != WTG_OK) goto fail;
  if ((rc = wtg_fill_u64(ctx, wt->words + nw, 1, 0)) != WTG_OK) goto fail;
  if (hipMemcpyAsync(wt->words, d_tree, nw * 8, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
    wtg_set_error("wtg_wtint_from_tree: copy failed");
    rc = WTG_ERR_HIP;
    goto fail;
  }
  if ((rc = wtg_rank_select_run(ctx, wt->words, wt->nbits, wt->dir, true, &wt->s1, &wt->s0)) != WTG_OK)
    goto fail;
  if ((rc = wtg_finish(ctx)) != WTG_OK) goto fail;
  *out = wt;
  return WTG_OK;
fail:
  wtg_wtint_destroy(ctx, wt);
  return rc;
}

extern "C" int wtg_wtint_destroy(wtg_ctx* ctx, wtg_wtint* wt) {
  if (!wt) return WTG_OK;
  WTG_REQUIRE(ctx, "null ctx");
  wtg_free(ctx, wt->words);
  wtg_free(ctx, wt->dir);
  wtg_select_free(ctx, &wt->s1);
  wtg_select_free(ctx, &wt->s0);
  delete wt;
  return WTG_OK;
}

extern "C" int wtg_wtint_info(const wtg_wtint* wt, uint64_t* n, uint32_t* levels, uint64_t* sigma,
                              const uint64_t** d_words, const uint64_t** d_dir) {
  WTG_REQUIRE(wt, "null wt");
  if (n) *n = wt->n;
  if (levels) *levels = wt->L;
  if (sigma) *sigma = wt->sigma;
  if (d_words) *d_words = wt->words;
  if (d_dir) *d_dir = wt->dir;
  return WTG_OK;
}
extern "C" const wtg_select* wtg_wtint_select1(const wtg_wtint* wt) { return wt ? &wt->s1 : nullptr; }
extern "C" const wtg_select* wtg_wtint_select0(const wtg_wtint* wt) { return wt ? &wt->s0 : nullptr; }

static int q_check(wtg_ctx* ctx, const wtg_wtint* wt, uint64_t nq, const void* a, const void* b) {
  WTG_REQUIRE(ctx && wt && (nq == 0 || (a && b)), "null argument");
  if (wt->n == 0 && nq) {
    wtg_set_error("query on an empty wt_int");
    return WTG_ERR_ARG;
  }
  return WTG_OK;
}

extern "C" int wtg_wtint_rank_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                    const uint64_t* d_c, uint64_t nq, uint64_t* d_out) {
  WTG_TRY(q_check(ctx, wt, nq, d_i, d_out));
  WTG_REQUIRE(nq == 0 || d_c, "null d_c");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "wtint_rank", k_wtint_rank, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             make_q(wt), d_i, d_c, nq, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_wtint_access_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                      uint64_t nq, uint64_t* d_out) {
  WTG_TRY(q_check(ctx, wt, nq, d_i, d_out));
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "wtint_access", k_wtint_access, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             make_q(wt), d_i, nq, d_out, (uint64_t*)nullptr);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_wtint_inverse_select_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                              uint64_t nq, uint64_t* d_rank, uint64_t* d_sym) {
  WTG_TRY(q_check(ctx, wt, nq, d_i, d_sym));
  WTG_REQUIRE(nq == 0 || d_rank, "null d_rank");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "wtint_access", k_wtint_access, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             make_q(wt), d_i, nq, d_sym, d_rank);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_wtint_lex_count_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                         const uint64_t* d_j, const uint64_t* d_c, uint64_t nq, uint64_t* d_rank,
                                         uint64_t* d_smaller, uint64_t* d_greater) {
  WTG_TRY(q_check(ctx, wt, nq, d_i, d_rank));
  WTG_REQUIRE(nq == 0 || (d_j && d_c && d_smaller && d_greater), "null argument");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "wtint_lex_count", k_wtint_lex_count, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             make_q(wt), d_i, d_j, d_c, nq, d_rank, d_smaller, d_greater);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_wtint_lex_smaller_count_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                                 const uint64_t* d_c, uint64_t nq, uint64_t* d_rank,
                                                 uint64_t* d_smaller) {
  WTG_TRY(q_check(ctx, wt, nq, d_i, d_rank));
  WTG_REQUIRE(nq == 0 || (d_c && d_smaller), "null argument");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "wtint_lex_smaller", k_wtint_lex_smaller, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             make_q(wt), d_i, d_c, nq, d_rank, d_smaller);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_wtint_range_count_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_lb,
                                           const uint64_t* d_rb, const uint64_t* d_vlb, const uint64_t* d_vrb,
                                           uint64_t nq, uint64_t* d_out) {
  WTG_TRY(q_check(ctx, wt, nq, d_lb, d_out));
  WTG_REQUIRE(nq == 0 || (d_rb && d_vlb && d_vrb), "null argument");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "wtint_range_count", k_wtint_range_count, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             make_q(wt), d_lb, d_rb, d_vlb, d_vrb, nq, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_wtint_select_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                      const uint64_t* d_c, uint64_t nq, uint64_t* d_out) {
  WTG_TRY(q_check(ctx, wt, nq, d_i, d_out));
  WTG_REQUIRE(nq == 0 || d_c, "null d_c");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  if (wt->L <= 24) {
    WTG_LAUNCH(ctx, "wtint_select", k_wtint_select<24>, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
               make_q(wt), d_i, d_c, nq, d_out);
  } else {
    WTG_LAUNCH(ctx, "wtint_select", k_wtint_select<64>, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
               make_q(wt), d_i, d_c, nq, d_out);
  }
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
