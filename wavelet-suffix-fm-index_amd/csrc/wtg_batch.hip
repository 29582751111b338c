// wtg_batch.hip -- batched level-wise wavelet trees of many uneven byte
// sequences (the per-sequence call WT(s_i, n_i, a_i) of wavelet/levelWT/WT.h:43
// for thousands of i at once).
//
// Sequences are grouped by their level count L = max(1, log2Up(a_i)), at most
// 2^(16-L) per group.  Within a group the sequences are concatenated and each
// symbol becomes the 16-bit key (group-local sequence id << L) | symbol.  The keys are then
// already ordered by their top bits, so the level pipeline starts at level
// `sbits` (the sequence-id bits) with the sequence lengths as that level's node
// counts: levels sbits .. sbits+L-1 of the keys are exactly levels 0 .. L-1 of
// every sequence, side by side.  The group's output is level-major: level l of
// sequence i is the bit range [l*N_g + P_i, l*N_g + P_i + n_i) of the group's
// words (N_g = group total, P_i = offset of i in the group).
//
// The first wtg_batch_run records the whole batch (key expansion + every
// level of every group) as one hipGraph; later runs replay it with one launch.
#include "wtg_internal.hpp"
#include <algorithm>
#include <cstring>
#include <vector>

namespace {

// Keys of one group: key(p) = (sequence << L) | symbol for the group's
// positions p (sequences concatenated in group order).  A thread builds 16
// consecutive keys: inside one sequence (the common case) from five aligned
// dword loads of the source bytes (funnel-shifted), stored as 16-byte vectors;
// across a sequence boundary or at the block's end one key at a time.  The
// block's sequence boundaries are cached in LDS.  Also counts, per sequence,
// the ones of its top symbol bit (level 0 of every sequence = the first level
// the pipeline builds).
constexpr int BK_SEQ_CACHE = 64;
template <typename K>
__global__ __launch_bounds__(256) void k_batch_keys(const uint8_t* __restrict__ sym, uint64_t sym_end, bool fast,
                                                    const uint64_t* __restrict__ seqoff,
                                                    const uint64_t* __restrict__ pref, uint32_t m, uint64_t n,
                                                    uint32_t L, K* __restrict__ keys, uint32_t* __restrict__ ones,
                                                    uint32_t* __restrict__ err) {
  __shared__ uint32_t s_first, s_cnt;
  __shared__ uint64_t s_pref[BK_SEQ_CACHE + 1], s_off[BK_SEQ_CACHE];
  __shared__ uint32_t s_ones[BK_SEQ_CACHE];
  const uint64_t b0 = (uint64_t)blockIdx.x * 4096;
  const uint64_t b1 = b0 + 4096 < n ? b0 + 4096 : n;
  if (threadIdx.x == 0) {  // sequences covering [b0, b1)
    uint32_t lo = 0, hi = m;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (pref[mid] <= b0) lo = mid; else hi = mid;
    }
    uint32_t l2 = lo, h2 = m;
    while (h2 - l2 > 1) {
      const uint32_t mid = (l2 + h2) / 2;
      if (pref[mid] < b1) l2 = mid; else h2 = mid;
    }
    s_first = lo;
    s_cnt = l2 - lo + 1;
  }
  __syncthreads();
  const uint32_t f = s_first, cnt = s_cnt;
  const bool cached = cnt <= BK_SEQ_CACHE;
  if (cached) {
    for (uint32_t k = threadIdx.x; k <= cnt; k += blockDim.x) s_pref[k] = k + f < m ? pref[f + k] : n;
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) { s_off[k] = seqoff[f + k]; s_ones[k] = 0; }
  }
  __syncthreads();
  auto seq_start = [&](uint32_t s) -> uint64_t { return cached ? s_pref[s] : pref[f + s]; };
  auto seq_end = [&](uint32_t s) -> uint64_t { return cached ? s_pref[s + 1] : (f + s + 1 < m ? pref[f + s + 1] : n); };
  auto seq_off = [&](uint32_t s) -> uint64_t { return cached ? s_off[s] : seqoff[f + s]; };
  auto add_ones = [&](uint32_t s, uint32_t c) {
    if (!c) return;
    if (cached) atomicAdd(&s_ones[s], c); else atomicAdd(&ones[f + s], c);
  };
  bool bad = false;
  const uint64_t p0 = b0 + (uint64_t)threadIdx.x * 16;
  if (p0 < b1) {
    // sequence of p0 (relative to f): last s with start <= p0
    uint32_t s = 0, hi = cnt;
    while (hi - s > 1) {
      const uint32_t mid = (s + hi) / 2;
      if (seq_start(mid) <= p0) s = mid; else hi = mid;
    }
    const uint64_t ps = seq_start(s), pe = seq_end(s);
    const uint64_t src = seq_off(s) + (p0 - ps);
    const uint32_t S = (f + s) << L;
    const uintptr_t ab = ((uintptr_t)(sym + src)) & ~(uintptr_t)3;
    if (fast && p0 + 16 <= pe && p0 + 16 <= b1 && (uint64_t)(ab - (uintptr_t)sym) + 20 <= sym_end) {
      const uint32_t* q = reinterpret_cast<const uint32_t*>(ab);
      const uint32_t sh = (uint32_t)(((uintptr_t)(sym + src)) & 3);
      uint32_t d[5];
#pragma unroll
      for (int k = 0; k < 5; k++) d[k] = q[k];
      uint32_t wv[4];
#pragma unroll
      for (int k = 0; k < 4; k++) wv[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], sh);
      const uint32_t hib = (0xFFu << L) & 0xFFu;  // symbol bits at or above L
      const uint32_t badm = hib * 0x01010101u;
      uint32_t c = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        bad |= (wv[k] & badm) != 0;
        c += (uint32_t)__popc((wv[k] >> (L - 1)) & 0x01010101u);
      }
      if constexpr (sizeof(K) == 1) {
        uint4 o;
        o.x = wv[0] | S * 0x01010101u;
        o.y = wv[1] | S * 0x01010101u;
        o.z = wv[2] | S * 0x01010101u;
        o.w = wv[3] | S * 0x01010101u;
        *reinterpret_cast<uint4*>(keys + p0) = o;
      } else if constexpr (sizeof(K) == 2) {
        const uint32_t SS = S | (S << 16);
        uint4 o[2];
        uint32_t* ov = reinterpret_cast<uint32_t*>(o);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          ov[2 * k] = SS | (wv[k] & 0xFFu) | ((wv[k] & 0xFF00u) << 8);
          ov[2 * k + 1] = SS | ((wv[k] >> 16) & 0xFFu) | ((wv[k] >> 8) & 0xFF0000u);
        }
        uint4* kp = reinterpret_cast<uint4*>(keys + p0);
        kp[0] = o[0];
        kp[1] = o[1];
      } else {
        uint4* kp = reinterpret_cast<uint4*>(keys + p0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          uint4 o;
          o.x = S | (wv[k] & 0xFFu);
          o.y = S | ((wv[k] >> 8) & 0xFFu);
          o.z = S | ((wv[k] >> 16) & 0xFFu);
          o.w = S | (wv[k] >> 24);
          kp[k] = o;
        }
      }
      add_ones(s, c);
    } else {
      uint32_t run = 0;
      uint64_t pend = pe;
      for (int k = 0; k < 16; k++) {
        const uint64_t p = p0 + k;
        if (p >= b1) break;
        while (p >= pend) {  // next sequence (those of length 0 are skipped)
          add_ones(s, run);
          run = 0;
          s++;
          pend = seq_end(s);
        }
        const uint32_t x = sym[seq_off(s) + (p - seq_start(s))];
        bad |= (x >> L) != 0;
        keys[p] = (K)((((uint64_t)(f + s)) << L) | x);
        run += (x >> (L - 1)) & 1u;
      }
      add_ones(s, run);
    }
  }
  if (bad) atomicOr(err, WTG_FLAG_BADSYM);
  if (cached) {
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x)
      if (s_ones[k]) atomicAdd(&ones[f + k], s_ones[k]);
  }
}

// copies `cnt` <= 64 bits starting at bit `b` of w (bits past the end are 0)
__device__ __forceinline__ uint64_t bits_at(const uint64_t* __restrict__ w, uint64_t b, uint32_t cnt) {
  const uint64_t i = b >> 6;
  const uint32_t sh = (uint32_t)(b & 63);
  uint64_t v = w[i] >> sh;
  if (sh && sh + cnt > 64) v |= w[i + 1] << (64 - sh);
  return cnt == 64 ? v : v & ((1ull << cnt) - 1);
}

// one sequence's levels in the WT() layout (level l at l*n_i)
__global__ void k_batch_extract(const uint64_t* __restrict__ src, uint64_t first, uint64_t stride, uint64_t ni,
                                uint32_t L, uint64_t* __restrict__ out) {
  const uint64_t total = ni * L;
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k * 64 >= total) return;
  uint64_t val = 0;
  uint64_t b = k * 64;
  const uint64_t e = (b + 64 < total) ? b + 64 : total;
  while (b < e) {
    const uint64_t l = b / ni, i = b - l * ni;
    const uint64_t rem = ni - i < e - b ? ni - i : e - b;
    const uint32_t cnt = (uint32_t)rem;
    val |= bits_at(src, first + l * stride + i, cnt) << (b - k * 64);
    b += cnt;
  }
  out[k] = val;
}

struct Group {
  uint32_t L = 0, sbits = 0;
  int key_bytes = 4;
  uint64_t n = 0;
  std::vector<uint32_t> seqs;   // batch indices in group order
  uint64_t* d_seqoff = nullptr;
  uint64_t* d_pref = nullptr;
  uint32_t* d_cnt = nullptr;
  uint32_t* d_ones = nullptr;   // ones of each sequence's top symbol bit (2^sbits)
  void* d_keys = nullptr;
  uint64_t word_base = 0;
};
}  // namespace

struct wtg_batch {
  const uint8_t* d_sym = nullptr;
  uint64_t sym_end = 0;         // max(off_i + n_i): bytes of d_sym the sequences span
  bool sym_fast = false;        // d_sym is 4-byte aligned: aligned dword loads of the symbols
  uint32_t nseq = 0;
  std::vector<uint64_t> n, off;
  std::vector<uint32_t> L, grp, idx;
  std::vector<uint64_t> pref;   // offset of each sequence inside its group
  std::vector<Group> groups;
  uint64_t* words = nullptr;
  uint64_t nwords = 0;
  size_t ws_need = 0;
  bool graph = true;
  bool mld = true;              // multi-level passes (else the level-by-level split)
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  void* graph_ws = nullptr;     // workspace the recorded graph addresses
};

static void batch_release(wtg_ctx* ctx, wtg_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->g) (void)hipGraphDestroy(b->g);
  for (auto& G : b->groups) {
    wtg_free(ctx, G.d_seqoff);
    wtg_free(ctx, G.d_pref);
    wtg_free(ctx, G.d_cnt);
    wtg_free(ctx, G.d_ones);
    wtg_free(ctx, G.d_keys);
  }
  wtg_free(ctx, b->words);
  delete b;
}

extern "C" int wtg_batch_destroy(wtg_ctx* ctx, wtg_batch* b) {
  if (!b) return WTG_OK;
  WTG_REQUIRE(ctx, "null ctx");
  batch_release(ctx, b);
  return WTG_OK;
}

static int log2up_u64(uint64_t x) { return wtg_log2up_host(x); }

// default group size: 16-bit keys (at most 2^(16-L) sequences per group)
static uint32_t default_cap_log(uint32_t L) { return 16 - L; }

extern "C" int wtg_batch_create_planned(wtg_ctx* ctx, const uint8_t* d_sym, uint32_t nseq, const uint64_t* h_off,
                                        const uint64_t* h_n, const uint32_t* h_sigma, const wtg_batch_plan* bplan,
                                        wtg_batch** out) {
  WTG_REQUIRE(ctx && out && nseq >= 1 && h_off && h_n && h_sigma, "wtg_batch_create: null argument");
  uint32_t caplog[9];
  for (uint32_t L = 1; L <= 8; L++) {
    caplog[L] = bplan && bplan->cap_log[L] != 0xFF ? bplan->cap_log[L] : default_cap_log(L);
    WTG_REQUIRE(caplog[L] + L <= 24, "wtg_batch_create: group size 2^cap_log needs cap_log + L <= 24");
  }
  WTG_TRY(wtg_set_device(ctx));
  *out = nullptr;
  wtg_batch* b = new wtg_batch();
  b->mld = ctx->level_method == 0 && wtg_mld_enabled();
  b->d_sym = d_sym;
  b->sym_fast = ((uintptr_t)d_sym & 3) == 0;
  b->nseq = nseq;
  b->n.assign(h_n, h_n + nseq);
  b->off.assign(h_off, h_off + nseq);
  b->L.resize(nseq);
  b->grp.resize(nseq);
  b->idx.resize(nseq);
  b->pref.resize(nseq);
  int rc = WTG_OK;
  uint64_t total = 0;
  for (uint32_t i = 0; i < nseq; i++) {
    if (h_sigma[i] < 1 || h_sigma[i] > 256) {
      wtg_set_error("wtg_batch_create: alphabet size must be in [1, 256]");
      batch_release(ctx, b);
      return WTG_ERR_ARG;
    }
    b->L[i] = (uint32_t)std::max(1, log2up_u64(h_sigma[i]));
    total += h_n[i];
    if (h_n[i]) b->sym_end = std::max(b->sym_end, h_off[i] + h_n[i]);
  }
  if (total && !d_sym) {
    wtg_set_error("wtg_batch_create: null symbols");
    batch_release(ctx, b);
    return WTG_ERR_ARG;
  }
  // groups by L, sequences in batch order inside a group, at most 2^cap_log[L]
  // sequences per group (default 2^(16-L): every key fits 16 bits)
  std::vector<Group> plan;
  for (uint32_t L = 1; L <= 8; L++) {
    std::vector<uint32_t> all;
    for (uint32_t i = 0; i < nseq; i++)
      if (b->L[i] == L) all.push_back(i);
    const size_t cap = (size_t)1 << caplog[L];
    for (size_t a0 = 0; a0 < all.size(); a0 += cap) {
      Group G;
      G.L = L;
      G.seqs.assign(all.begin() + a0, all.begin() + std::min(all.size(), a0 + cap));
      plan.push_back(G);
    }
  }
  for (Group& G : plan) {
    const uint32_t L = G.L;
    const uint32_t m = (uint32_t)G.seqs.size();
    G.sbits = (uint32_t)log2up_u64(m);
    G.key_bytes = G.sbits + L <= 8 ? 1 : G.sbits + L <= 16 ? 2 : 4;
    std::vector<uint64_t> so(m), pr(m + 1, 0);
    std::vector<uint32_t> cnt(1ull << G.sbits, 0);
    for (uint32_t t = 0; t < m; t++) {
      const uint32_t i = G.seqs[t];
      so[t] = b->off[i];
      pr[t + 1] = pr[t] + b->n[i];
      cnt[t] = (uint32_t)b->n[i];
      b->grp[i] = (uint32_t)b->groups.size();
      b->idx[i] = t;
      b->pref[i] = pr[t];
    }
    G.n = pr[m];
    if (G.n >= (1ull << 32) || G.sbits + L > 24) {
      wtg_set_error("wtg_batch_create: a group of sequences exceeds 2^32 - 1 symbols");
      b->groups.push_back(G);
      batch_release(ctx, b);
      return WTG_ERR_UNSUPPORTED;
    }
    G.word_base = b->nwords;
    b->nwords += wtg_levels_words(G.n, L) + 1;
    if ((rc = wtg_malloc(ctx, (void**)&G.d_seqoff, m * 8)) != WTG_OK ||
        (rc = wtg_malloc(ctx, (void**)&G.d_pref, (m + 1) * 8)) != WTG_OK ||
        (rc = wtg_malloc(ctx, (void**)&G.d_cnt, cnt.size() * 4)) != WTG_OK ||
        (rc = wtg_malloc(ctx, (void**)&G.d_ones, cnt.size() * 4)) != WTG_OK ||
        (!b->mld && (rc = wtg_malloc(ctx, &G.d_keys, std::max<uint64_t>(G.n, 1) * G.key_bytes)) != WTG_OK)) {
      b->groups.push_back(G);
      batch_release(ctx, b);
      return rc;
    }
    if ((rc = wtg_memcpy_h2d(ctx, G.d_seqoff, so.data(), m * 8)) != WTG_OK ||
        (rc = wtg_memcpy_h2d(ctx, G.d_pref, pr.data(), (m + 1) * 8)) != WTG_OK ||
        (rc = wtg_memcpy_h2d(ctx, G.d_cnt, cnt.data(), cnt.size() * 4)) != WTG_OK) {
      b->groups.push_back(G);
      batch_release(ctx, b);
      return rc;
    }
    b->ws_need = std::max(b->ws_need, b->mld ? wtg_mld_ws_bytes(G.n, G.sbits + L, G.sbits)
                                             : wtg_levels_ws_bytes(G.n, G.sbits + L, G.key_bytes));
    b->groups.push_back(G);
  }
  if ((rc = wtg_malloc(ctx, (void**)&b->words, std::max<uint64_t>(b->nwords, 1) * 8)) != WTG_OK) {
    batch_release(ctx, b);
    return rc;
  }
  if ((rc = wtg_fill_u64(ctx, b->words, std::max<uint64_t>(b->nwords, 1), 0)) != WTG_OK ||
      (rc = wtg_ws_reserve(ctx, b->ws_need)) != WTG_OK || (rc = wtg_finish(ctx)) != WTG_OK) {
    batch_release(ctx, b);
    return rc;
  }
  *out = b;
  return WTG_OK;
}

// every kernel of the batch, in stream order (no allocation, no host sync)
static int batch_enqueue(wtg_ctx* ctx, wtg_batch* b) {
  for (auto& G : b->groups) {
    if (G.n == 0) continue;
    if (b->mld) {
      // up to 8 levels per pass from level sbits, straight from the byte symbols:
      // a sequence is a level-sbits node, its keys its own bytes (their low L bits)
      if (G.sbits) {
        WTG_TRY(wtg_mld_build_from(ctx, b->d_sym, 1, G.n, G.sbits + G.L, G.sbits, G.d_cnt, G.d_seqoff,
                                   b->words + G.word_base));
      } else {
        WTG_TRY(wtg_mld_build_from(ctx, b->d_sym + b->off[G.seqs[0]], 1, G.n, G.L, 0, nullptr, nullptr,
                                   b->words + G.word_base));
      }
      continue;
    }
    const unsigned grid = (unsigned)wtg_ceil_div(G.n, 4096);
    const uint32_t m = (uint32_t)G.seqs.size();
    WTG_TRY(wtg_fill_bytes_zero(ctx, G.d_ones, (1ull << G.sbits) * 4));
    if (G.key_bytes == 1)
      WTG_LAUNCH(ctx, "batch_keys", k_batch_keys<uint8_t>, dim3(grid), dim3(256), 0, b->d_sym, b->sym_end,
                 b->sym_fast, G.d_seqoff, G.d_pref, m, G.n, G.L, (uint8_t*)G.d_keys, G.d_ones, ctx->d_flag);
    else if (G.key_bytes == 2)
      WTG_LAUNCH(ctx, "batch_keys", k_batch_keys<uint16_t>, dim3(grid), dim3(256), 0, b->d_sym, b->sym_end,
                 b->sym_fast, G.d_seqoff, G.d_pref, m, G.n, G.L, (uint16_t*)G.d_keys, G.d_ones, ctx->d_flag);
    else
      WTG_LAUNCH(ctx, "batch_keys", k_batch_keys<uint32_t>, dim3(grid), dim3(256), 0, b->d_sym, b->sym_end,
                 b->sym_fast, G.d_seqoff, G.d_pref, m, G.n, G.L, (uint32_t*)G.d_keys, G.d_ones, ctx->d_flag);
    WTG_LAUNCH_CHECK();
    WTG_TRY(wtg_levels_build_from(ctx, G.d_keys, G.key_bytes, G.n, G.sbits + G.L, G.sbits,
                                  G.sbits ? G.d_cnt : nullptr, b->words + G.word_base,
                                  G.sbits ? G.d_ones : nullptr));
  }
  return WTG_OK;
}

extern "C" int wtg_batch_set_graph(wtg_batch* b, int use_graph) {
  WTG_REQUIRE(b, "null batch");
  b->graph = use_graph != 0;
  return WTG_OK;
}

extern "C" int wtg_batch_run(wtg_ctx* ctx, wtg_batch* b) {
  WTG_REQUIRE(ctx && b, "wtg_batch_run: null argument");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(wtg_ws_reserve(ctx, b->ws_need));
  if (!b->graph) {
    WTG_TRY(batch_enqueue(ctx, b));
    return wtg_finish(ctx);
  }
  if (b->exec && b->graph_ws != ctx->ws) {  // the workspace moved: record again
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->g);
    b->exec = nullptr;
    b->g = nullptr;
  }
  if (!b->exec) {
    const bool kt = ctx->ktime;
    ctx->ktime = false;  // no timing events inside the recorded graph
    WTG_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(ctx, b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->ktime = kt;
    if (rc != WTG_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    WTG_CHECK_HIP(e);
    b->g = g;
    WTG_CHECK_HIP(hipGraphInstantiate(&b->exec, b->g, nullptr, nullptr, 0));
    b->graph_ws = ctx->ws;
  }
  {
    wtg_kscope ks(ctx, "batch_graph");
    WTG_CHECK_HIP(hipGraphLaunch(b->exec, ctx->stream));
  }
  return wtg_finish(ctx);
}

extern "C" int wtg_batch_words(const wtg_batch* b, const uint64_t** d_words, uint64_t* nwords) {
  WTG_REQUIRE(b, "null batch");
  if (d_words) *d_words = b->words;
  if (nwords) *nwords = b->nwords;
  return WTG_OK;
}

extern "C" int wtg_batch_seq(const wtg_batch* b, uint32_t seq, uint32_t* levels, uint64_t* first_bit,
                             uint64_t* level_stride) {
  WTG_REQUIRE(b && seq < b->nseq, "wtg_batch_seq: bad sequence");
  const Group& G = b->groups[b->grp[seq]];
  if (levels) *levels = b->L[seq];
  if (first_bit) *first_bit = G.word_base * 64 + b->pref[seq];
  if (level_stride) *level_stride = G.n;
  return WTG_OK;
}

extern "C" int wtg_batch_extract(wtg_ctx* ctx, const wtg_batch* b, uint32_t seq, uint64_t* d_out) {
  WTG_REQUIRE(ctx && b && seq < b->nseq && d_out, "wtg_batch_extract: bad argument");
  WTG_TRY(wtg_set_device(ctx));
  const Group& G = b->groups[b->grp[seq]];
  const uint64_t ni = b->n[seq];
  const uint32_t L = b->L[seq];
  const uint64_t nw = wtg_levels_words(ni, L);
  if (!nw) return WTG_OK;
  WTG_LAUNCH(ctx, "batch_extract", k_batch_extract, dim3((unsigned)wtg_ceil_div(nw, 256)), dim3(256), 0,
             b->words + G.word_base, b->pref[seq], G.n, ni, L, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

extern "C" int wtg_batch_create(wtg_ctx* ctx, const uint8_t* d_sym, uint32_t nseq, const uint64_t* h_off,
                                const uint64_t* h_n, const uint32_t* h_sigma, wtg_batch** out) {
  return wtg_batch_create_planned(ctx, d_sym, nseq, h_off, h_n, h_sigma, nullptr, out);
}

// Autotuner: for each alphabet width L present, time every candidate group
// size on the batch's own sequences of that width (one recorded run, then the
// best of `reps` graph replays) and keep the fastest.  Candidates: 2^(8-L)
// sequences per group (8-bit keys), 2^(14-L) and 2^(16-L) (16-bit keys),
// 2^(20-L) and 2^(24-L) (32-bit keys; one group holds every sequence of the
// width when they fit).
extern "C" int wtg_batch_tune(wtg_ctx* ctx, const uint8_t* d_sym, uint32_t nseq, const uint64_t* h_off,
                              const uint64_t* h_n, const uint32_t* h_sigma, uint32_t reps, wtg_batch_plan* out,
                              float* ms_out) {
  WTG_REQUIRE(ctx && out && nseq >= 1 && h_off && h_n && h_sigma && reps >= 1, "wtg_batch_tune: bad argument");
  std::memset(out->cap_log, 0xFF, sizeof(out->cap_log));
  if (ms_out)
    for (int k = 0; k < 9 * 5; k++) ms_out[k] = -1.0f;
  for (uint32_t L = 1; L <= 8; L++) {
    std::vector<uint64_t> off, n;
    std::vector<uint32_t> sig;
    for (uint32_t i = 0; i < nseq; i++) {
      WTG_REQUIRE(h_sigma[i] >= 1 && h_sigma[i] <= 256, "wtg_batch_tune: alphabet size must be in [1, 256]");
      if ((uint32_t)std::max(1, log2up_u64(h_sigma[i])) != L) continue;
      off.push_back(h_off[i]);
      n.push_back(h_n[i]);
      sig.push_back(h_sigma[i]);
    }
    if (off.empty()) continue;
    const int cand[5] = {(int)(8 - L), (int)(16 - L) - 2, (int)(16 - L), (int)(20 - L), (int)(24 - L)};
    float best = 0.f;
    for (int c = 0; c < 5; c++) {
      if (cand[c] < 0) continue;
      // a candidate no coarser than needed for the width's sequence count
      const uint64_t m = off.size();
      if (c > 0 && ((uint64_t)1 << cand[c - 1]) >= m && cand[c - 1] >= 0) continue;
      if (wtg_ceil_div(m, (uint64_t)1 << cand[c]) > 64) continue;  // at most 64 groups per width
      wtg_batch_plan p;
      std::memset(p.cap_log, 0xFF, sizeof(p.cap_log));
      p.cap_log[L] = (uint8_t)cand[c];
      wtg_batch* b = nullptr;
      WTG_TRY(wtg_batch_create_planned(ctx, d_sym, (uint32_t)m, off.data(), n.data(), sig.data(), &p, &b));
      int rc = wtg_batch_run(ctx, b);  // records the graph
      float ms = 0.f, bestc = 1e30f;
      for (uint32_t r = 0; r < reps && rc == WTG_OK; r++) {
        if ((rc = wtg_timer_start(ctx)) == WTG_OK && (rc = wtg_batch_run(ctx, b)) == WTG_OK &&
            (rc = wtg_timer_stop(ctx, &ms)) == WTG_OK)
          bestc = std::min(bestc, ms);
      }
      wtg_batch_destroy(ctx, b);
      if (rc != WTG_OK) return rc;
      if (ms_out) ms_out[L * 5 + c] = bestc;
      if (out->cap_log[L] == 0xFF || bestc < best) {
        best = bestc;
        out->cap_log[L] = (uint8_t)cand[c];
      }
    }
  }
  return WTG_OK;
}
