// wtg_deep.hip -- level-wise wavelet tree for 25..64 levels (wt_int over wide
// values, wt_int/wt_int_par.hpp:384-395 accepts any max_level up to 64).
//
// Reference semantics (wavelet/levelWT/WT.C:43-227 == serialWT/WT.C:32-103):
// level l holds P_l = the stable sort of the symbols by their top l bits; the
// bit at global position l*n + i is bit L-1-l of P_l[i].
//
// The dense schedules index their node tables by prefix (2^l entries), which
// stops at 24 levels.  Here nodes are numbered by their order in the level:
// P_l is sorted by prefix, so a node is a run of equal prefixes and its id is
// the number of runs before it.  Per level, over the keys in P_l order:
//   k_deep_flags  per key: run head (prefix differs from the key before) and
//                 zero (level bit 0), packed zero | head << 32; the level's
//                 bits from wave ballots;
//   exclusive scan of the packed flags (rocPRIM): zeros before i (low half)
//                 and heads before i (high half);
//   k_deep_nodes  per run: start S, zeros before it Zs, zeros before its end Ze
//                 (compact arrays, one entry per non-empty node);
//   k_deep_split  the stable split of every node: a key at i in node v with
//                 zi zeros before it goes to S + (zi - Zs) (bit 0) or
//                 S + (Ze - Zs) + (i - S) - (zi - Zs) (bit 1): P_{l+1}.
// Work and memory are linear in n per level whatever the alphabet; this path
// is chosen only above 24 levels (the multi-level passes are faster below).
// With the multi-level passes on, the top 24 levels come from them (levels
// 0..23 depend only on the keys' top 24 bits), P_24 from a stable radix sort
// of the keys by those 24 bits, and only the levels from 24 on take the
// per-level steps above.
#include "wtg_internal.hpp"
#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace {

__device__ __forceinline__ uint64_t deep_prefix(uint64_t x, uint32_t L, uint32_t l) {
  return l == 0 ? 0ull : x >> (L - l);  // top l bits (L - l in 1..63 for l >= 1)
}

template <typename K>
__global__ __launch_bounds__(256) void k_deep_flags(const K* __restrict__ keys, uint64_t n, uint32_t L, uint32_t l,
                                                    uint64_t* __restrict__ flags, unsigned long long* __restrict__ words,
                                                    uint32_t* __restrict__ err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < n;
  uint64_t x = 0;
  bool one = false;
  if (act) {
    x = (uint64_t)keys[i];
    if (l == 0 && L < 64 && (x >> L) != 0) atomicOr(err, WTG_FLAG_BADSYM);
    const bool head = i == 0 || deep_prefix(x, L, l) != deep_prefix((uint64_t)keys[i - 1], L, l);
    one = (x >> (L - 1 - l)) & 1u;
    flags[i] = (uint64_t)(!one) | ((uint64_t)head << 32);
  }
  // the wave's 64 consecutive level bits at global bit l*n + i0 (not word
  // aligned: OR-ed into the one or two words they cover)
  const uint64_t m = wave_ballot(act && one);
  const uint64_t i0 = i - (uint64_t)lane_id();
  if (lane_id() == 0 && m) {
    const uint64_t b = (uint64_t)l * n + i0;
    const uint32_t o = (uint32_t)(b & 63);
    atomicOr(&words[b >> 6], (unsigned long long)(m << o));
    if (o && (m >> (64 - o))) atomicOr(&words[(b >> 6) + 1], (unsigned long long)(m >> (64 - o)));
  }
}

// per run (node) of the level: start, zeros before it, zeros before its end
__global__ __launch_bounds__(256) void k_deep_nodes(const uint64_t* __restrict__ flags,
                                                    const uint64_t* __restrict__ scan, uint64_t n,
                                                    uint32_t* __restrict__ S, uint32_t* __restrict__ Zs,
                                                    uint32_t* __restrict__ Ze) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t f = flags[i], c = scan[i];
  const uint32_t zi = (uint32_t)c, hb = (uint32_t)(c >> 32);
  if (f >> 32) {  // a head: node hb starts here, node hb-1 ends here
    S[hb] = (uint32_t)i;
    Zs[hb] = zi;
    if (hb) Ze[hb - 1] = zi;
  }
  if (i == n - 1) Ze[hb + (uint32_t)(f >> 32) - 1] = zi + (uint32_t)(f & 1u);
}

template <typename K>
__global__ __launch_bounds__(256) void k_deep_split(const K* __restrict__ in, K* __restrict__ out, uint64_t n,
                                                    const uint64_t* __restrict__ flags,
                                                    const uint64_t* __restrict__ scan, const uint32_t* __restrict__ S,
                                                    const uint32_t* __restrict__ Zs, const uint32_t* __restrict__ Ze) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t f = flags[i], c = scan[i];
  const uint32_t zi = (uint32_t)c;
  const uint32_t v = (uint32_t)(c >> 32) + (uint32_t)(f >> 32) - 1;  // this key's node
  const uint32_t s = S[v], zs = Zs[v];
  const uint64_t dest = (f & 1u) ? (uint64_t)s + (zi - zs)
                                 : (uint64_t)s + (Ze[v] - zs) + ((uint32_t)i - s) - (zi - zs);
  out[dest] = in[i];
}

// distinct symbols: runs of equal keys in P_L (the keys sorted by all L bits)
template <typename K>
__global__ __launch_bounds__(256) void k_deep_distinct(const K* __restrict__ keys, uint64_t n,
                                                       unsigned long long* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool head = i < n && (i == 0 || keys[i] != keys[i - 1]);
  const uint64_t m = wave_ballot(head);
  if (lane_id() == 0 && m) atomicAdd(out, (unsigned long long)popc64(m));
}

constexpr uint32_t DEEP_TOP = 24;  // levels built by the multi-level passes first
// the P_24 sort: onesweep (or one block) at every size.  rocPRIM's merge-sort
// path, which it otherwise takes up to 2^20 keys, masks a bit range that ends
// at the key's top bit with 1 << width (undefined: here an empty mask)
using DeepSortCfg = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                               rocprim::default_config, 0>;

template <typename K>
int deep_build_t(wtg_ctx* ctx, const K* d_sym, uint64_t n, uint32_t L, uint64_t* d_words,
                 unsigned long long* d_distinct) {
  const bool top = ctx->level_method == 0 && wtg_mld_enabled();
  constexpr uint32_t KB = 8 * sizeof(K);
  // P_24 = the keys stably sorted by bits [L - 24, L) (bits past the key type are 0)
  const uint32_t sort_lo = L - DEEP_TOP, sort_hi = L < KB ? L : KB;
  const bool sort = top && sort_lo < sort_hi;
  size_t scan_tmp = 0, sort_tmp = 0;
  {
    uint64_t* p = nullptr;
    if (rocprim::exclusive_scan(nullptr, scan_tmp, p, p, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(),
                                ctx->stream) != hipSuccess) {
      wtg_set_error("levels (deep): scan size query failed");
      return WTG_ERR_HIP;
    }
    K* q = nullptr;
    if (sort && rocprim::radix_sort_keys<DeepSortCfg>(nullptr, sort_tmp, (const K*)q, q, (size_t)n, sort_lo,
                                                      sort_hi, ctx->stream) !=
                    hipSuccess) {
      wtg_set_error("levels (deep): sort size query failed");
      return WTG_ERR_HIP;
    }
  }
  const size_t o_k0 = 0;
  const size_t o_k1 = wtg_align256(o_k0 + n * sizeof(K));
  const size_t o_fl = wtg_align256(o_k1 + n * sizeof(K));
  const size_t o_sc = wtg_align256(o_fl + n * 8);
  const size_t o_S = wtg_align256(o_sc + n * 8);
  const size_t o_Zs = wtg_align256(o_S + n * 4);
  const size_t o_Ze = wtg_align256(o_Zs + n * 4);
  const size_t o_tmp = wtg_align256(o_Ze + n * 4);
  const size_t need = o_tmp + std::max(scan_tmp, sort_tmp) + 256;
  // one reservation for both steps (the multi-level passes then find it large enough)
  WTG_TRY(wtg_ws_reserve(ctx, top ? std::max(need, wtg_mld_ws_bytes(n, DEEP_TOP, 0)) : need));
  uint32_t l0 = 0;
  if (top) {
    // levels 0..23 (and their words cleared), then the remaining words cleared
    WTG_TRY(wtg_mld_build_top(ctx, d_sym, (int)sizeof(K), n, L, DEEP_TOP, d_words));
    const uint64_t wt = wtg_levels_words(n, DEEP_TOP);
    WTG_TRY(wtg_fill_bytes_zero(ctx, d_words + wt, (wtg_levels_words(n, L) - wt) * 8));
    l0 = DEEP_TOP;
  } else {
    WTG_TRY(wtg_fill_bytes_zero(ctx, d_words, wtg_levels_words(n, L) * 8));
  }
  char* ws = (char*)ctx->ws;
  K* kb[2] = {(K*)(ws + o_k0), (K*)(ws + o_k1)};
  uint64_t* fl = (uint64_t*)(ws + o_fl);
  uint64_t* sc = (uint64_t*)(ws + o_sc);
  uint32_t* S = (uint32_t*)(ws + o_S);
  uint32_t* Zs = (uint32_t*)(ws + o_Zs);
  uint32_t* Ze = (uint32_t*)(ws + o_Ze);
  void* tmp = ws + o_tmp;
  const unsigned grid = (unsigned)wtg_ceil_div(n, 256);
  const K* cur = d_sym;
  if (sort) {
    // into the buffer the first split below does not write
    K* out = kb[(l0 & 1) ^ 1];
    size_t tb = sort_tmp;
    if (rocprim::radix_sort_keys<DeepSortCfg>(tmp, tb, d_sym, out, (size_t)n, sort_lo, sort_hi, ctx->stream) !=
        hipSuccess) {
      wtg_set_error("levels (deep): sort failed");
      return WTG_ERR_HIP;
    }
    cur = out;
  }
  for (uint32_t l = l0; l < L; l++) {
    WTG_LAUNCH(ctx, "deep_flags", k_deep_flags<K>, dim3(grid), dim3(256), 0, cur, n, L, l, fl,
               (unsigned long long*)d_words, ctx->d_flag);
    WTG_LAUNCH_CHECK();
    if (l + 1 == L) break;  // the last level's order is not needed
    size_t tb = scan_tmp;
    if (rocprim::exclusive_scan(tmp, tb, fl, sc, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), ctx->stream) !=
        hipSuccess) {
      wtg_set_error("levels (deep): scan failed");
      return WTG_ERR_HIP;
    }
    WTG_LAUNCH(ctx, "deep_nodes", k_deep_nodes, dim3(grid), dim3(256), 0, fl, sc, n, S, Zs, Ze);
    WTG_LAUNCH_CHECK();
    K* nxt = kb[l & 1];
    WTG_LAUNCH(ctx, "deep_split", k_deep_split<K>, dim3(grid), dim3(256), 0, cur, nxt, n, fl, sc, S, Zs, Ze);
    WTG_LAUNCH_CHECK();
    cur = nxt;
  }
  if (d_distinct) {
    // P_L: one more split of the last level orders the keys by all L bits
    WTG_TRY(wtg_fill_u64(ctx, d_distinct, 1, 0));
    size_t tb = scan_tmp;
    K* nxt = cur == kb[0] ? kb[1] : kb[0];
    if (rocprim::exclusive_scan(tmp, tb, fl, sc, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), ctx->stream) !=
        hipSuccess) {
      wtg_set_error("levels (deep): scan failed");
      return WTG_ERR_HIP;
    }
    WTG_LAUNCH(ctx, "deep_nodes", k_deep_nodes, dim3(grid), dim3(256), 0, fl, sc, n, S, Zs, Ze);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "deep_split", k_deep_split<K>, dim3(grid), dim3(256), 0, cur, nxt, n, fl, sc, S, Zs, Ze);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "deep_distinct", k_deep_distinct<K>, dim3(grid), dim3(256), 0, nxt, n, d_distinct);
    WTG_LAUNCH_CHECK();
  }
  return WTG_OK;
}

}  // namespace

int wtg_deep_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint64_t* d_words,
                   unsigned long long* d_distinct) {
  switch (sym_bytes) {
    case 1: return deep_build_t<uint8_t>(ctx, (const uint8_t*)d_sym, n, L, d_words, d_distinct);
    case 2: return deep_build_t<uint16_t>(ctx, (const uint16_t*)d_sym, n, L, d_words, d_distinct);
    case 4: return deep_build_t<uint32_t>(ctx, (const uint32_t*)d_sym, n, L, d_words, d_distinct);
    default: return deep_build_t<uint64_t>(ctx, (const uint64_t*)d_sym, n, L, d_words, d_distinct);
  }
}
