// wtg_sa_gpu.hip -- suffix array of a byte text on the GPU by prefix doubling
// (the reference calls divsufsort, fm-index/par/WT.C:35-36; any exact suffix
// sorter gives the same array).  The sigma distinct byte values are coded
// 1..sigma in byte order in c = bits(sigma) bits, 0 past the text end (so a
// suffix that ends sorts before its extensions); round 0 sorts all suffixes by
// their first q = floor(63 / c) bytes (7 for 256 values, 21 for DNA).  A
// suffix's rank is the position of its group's first suffix in the sorted
// order (32 bits).  Each later round h = q, 2q, 4q, ... re-sorts the suffixes
// still in groups of two or more (every suffix, in text order, while they are
// more than half) by (group rank, rank h bytes on + 1; 0 past the end) and
// splits each group in place: new rank = group rank + offset of the subgroup's
// first suffix in the group.  A suffix alone in its group keeps its final rank
// (the ranks stay consistent, as in Larsson-Sadakane doubling) and its sa entry
// is written then; it stops when no group of two or more is left.
#include "wtg_internal.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace {

// byte values present in the text: present[b] = 1
__global__ __launch_bounds__(256) void k_sa_present(const uint8_t* __restrict__ t, uint64_t n,
                                                    uint32_t* __restrict__ present) {
  __shared__ uint32_t s[256];
  s[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 16;
#pragma unroll
  for (int j = 0; j < 16; j++)
    if (base + j < n) s[t[base + j]] = 1u;
  __syncthreads();
  if (s[threadIdx.x]) present[threadIdx.x] = 1u;
}

// round-0 key of suffix i: its first q bytes, each as code[byte] (1..sigma, in
// byte order) in c bits, 0 past the text end; q*c <= 63, so q <= 63.  The
// block's 256 + 63 codes are staged in LDS.
__global__ __launch_bounds__(256) void k_sa_init(const uint8_t* __restrict__ t, uint64_t n,
                                                 const uint16_t* __restrict__ code, uint32_t c, uint32_t q,
                                                 uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
  __shared__ uint16_t s_code[256];
  __shared__ uint16_t s_win[256 + 64];
  s_code[threadIdx.x] = code[threadIdx.x];
  __syncthreads();
  const uint64_t b0 = (uint64_t)blockIdx.x * 256;
  for (uint32_t j = threadIdx.x; j < 256 + 64; j += 256) {
    const uint64_t p = b0 + j;
    s_win[j] = p < n ? s_code[t[p]] : (uint16_t)0;
  }
  __syncthreads();
  const uint64_t i = b0 + threadIdx.x;
  if (i >= n) return;
  uint64_t k = 0;
  for (uint32_t j = 0; j < q; j++) k = (k << c) | (uint64_t)s_win[threadIdx.x + j];
  key[i] = k;
  idx[i] = (uint32_t)i;
}

// after a sort of m keys, per sorted position k: fb[k] = first position of k's
// run of equal keys and fa[k] = first position of k's run of equal rank parts
// (key >> sh; not computed when whole_group), both as inclusive max-scans of
// run starts within 4096-position tiles, tile maxima to tb / ta (completed by
// the exclusive max-scan of the tile maxima)
__global__ __launch_bounds__(256) void k_sa_runs_tiles(const uint64_t* __restrict__ key, uint64_t m,
                                                       int whole_group, uint32_t sh, uint32_t* __restrict__ fa,
                                                       uint32_t* __restrict__ fb, uint32_t* __restrict__ ta,
                                                       uint32_t* __restrict__ tb) {
  // the tile's keys go through LDS (coalesced loads; thread t then takes keys
  // 16t..16t+15), padded one slot per 16 against bank conflicts; the same
  // space then holds the two scanned outputs for coalesced stores
  constexpr uint32_t PAD = 4096 + 4096 / 16;
  __shared__ uint64_t s_k[PAD];
  __shared__ uint32_t s_a[4], s_b[4];
  uint32_t* s_o = (uint32_t*)s_k;  // [0, PAD) fa, [PAD, 2 PAD) fb
  const uint64_t tile0 = (uint64_t)blockIdx.x * 4096;
#pragma unroll
  for (uint32_t j = 0; j < 16; j++) {
    const uint32_t p = j * 256 + threadIdx.x;
    if (tile0 + p < m) s_k[p + (p >> 4)] = key[tile0 + p];
  }
  __syncthreads();
  const uint32_t l0 = threadIdx.x * 16;
  const uint64_t base = tile0 + l0;
  uint32_t va[16], vb[16], ma = 0, mb = 0;
  uint64_t prev = 0;
  if (base < m) {
    if (l0 > 0) prev = s_k[(l0 - 1) + ((l0 - 1) >> 4)];
    else if (tile0 > 0) prev = key[tile0 - 1];
  }
#pragma unroll
  for (uint32_t j = 0; j < 16; j++) {
    const uint64_t p = base + j;
    uint32_t a = 0, b = 0;
    if (p < m) {
      const uint64_t kc = s_k[threadIdx.x * 17 + j];
      b = (p == 0 || kc != prev) ? (uint32_t)p : 0u;
      a = (!whole_group && (p == 0 || (kc >> sh) != (prev >> sh))) ? (uint32_t)p : 0u;
      prev = kc;
    }
    ma = a > ma ? a : ma;
    mb = b > mb ? b : mb;
    va[j] = ma;
    vb[j] = mb;
  }
  uint32_t xa = ma, xb = mb;
  const int lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t ya = (uint32_t)__shfl_up((int)xa, d);
    const uint32_t yb = (uint32_t)__shfl_up((int)xb, d);
    if (lane >= d) {
      xa = ya > xa ? ya : xa;
      xb = yb > xb ? yb : xb;
    }
  }
  if (lane == 63) {
    s_a[w] = xa;
    s_b[w] = xb;
  }
  __syncthreads();  // also: every thread is done reading s_k
  uint32_t pa = 0, pb = 0;
  for (int k = 0; k < w; k++) {
    pa = s_a[k] > pa ? s_a[k] : pa;
    pb = s_b[k] > pb ? s_b[k] : pb;
  }
  uint32_t ea = (uint32_t)__shfl_up((int)xa, 1), eb = (uint32_t)__shfl_up((int)xb, 1);
  if (lane == 0) ea = eb = 0;
  ea = ea > pa ? ea : pa;
  eb = eb > pb ? eb : pb;
#pragma unroll
  for (uint32_t j = 0; j < 16; j++) {
    s_o[threadIdx.x * 17 + j] = va[j] > ea ? va[j] : ea;
    s_o[PAD + threadIdx.x * 17 + j] = vb[j] > eb ? vb[j] : eb;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < 16; j++) {
    const uint32_t p = j * 256 + threadIdx.x;
    if (tile0 + p < m) {
      if (!whole_group) fa[tile0 + p] = s_o[p + (p >> 4)];
      fb[tile0 + p] = s_o[PAD + p + (p >> 4)];
    }
  }
  if (threadIdx.x == 0) {
    uint32_t xa2 = 0, xb2 = 0;
    for (int k = 0; k < 4; k++) {
      xa2 = s_a[k] > xa2 ? s_a[k] : xa2;
      xb2 = s_b[k] > xb2 ? s_b[k] : xb2;
    }
    if (!whole_group) ta[blockIdx.x] = xa2;
    tb[blockIdx.x] = xb2;
  }
}

// exclusive max-scan of the tile maxima in place (one block, 4096 per chunk)
__global__ __launch_bounds__(256) void k_sa_tile_prefix(uint32_t* __restrict__ tmax, uint64_t ntiles) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  for (uint64_t c0 = 0; c0 < ntiles; c0 += 4096) {
    const uint64_t base = c0 + threadIdx.x * 16;
    uint32_t v[16], m = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      v[j] = base + j < ntiles ? tmax[base + j] : 0u;
      m = v[j] > m ? v[j] : m;
    }
    uint32_t x = m;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, d);
      if (lane >= d) x = y > x ? y : x;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t pre = s_carry;
    for (int k = 0; k < w; k++) pre = s_w[k] > pre ? s_w[k] : pre;
    uint32_t ex = (uint32_t)__shfl_up((int)x, 1);
    if (lane == 0) ex = 0;
    uint32_t run = ex > pre ? ex : pre;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (base + j < ntiles) tmax[base + j] = run;  // exclusive
      run = v[j] > run ? v[j] : run;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t mx = s_carry;
      for (int k = 0; k < 4; k++) mx = s_w[k] > mx ? s_w[k] : mx;
      s_carry = mx;
    }
    __syncthreads();
  }
}

// keys (rank << sh) | (rank h bytes on + 1; 0 past the end), sh = bits of n:
// of the unresolved suffixes act[0..m) or, with act null, of every suffix in
// text order (coalesced rank reads while most suffixes are unresolved; a
// resolved suffix sorts alone in its group and keeps its rank)
__global__ void k_sa_keys(const uint32_t* __restrict__ rank, const uint32_t* __restrict__ act, uint64_t m,
                          uint64_t n, uint64_t h, uint32_t sh, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t i = act ? act[k] : (uint32_t)k;
  key[k] = ((uint64_t)rank[i] << sh) | (i + h < n ? (uint64_t)rank[i + h] + 1 : 0ull);
  val[k] = i;
}

// new rank of sorted position k: the group's old rank (0 in round 0) plus the
// offset of k's subgroup head from the group's first position (fa, fb, ta, tb
// from k_sa_runs_tiles and the tile prefix).  A suffix alone in its subgroup has
// its final rank r: sa[r] = suffix (r grows with k, so these stores are nearly
// sequential) and keep[k] = 0; else keep[k] = 1.
__global__ void k_sa_update(const uint64_t* __restrict__ key, const uint32_t* __restrict__ val, uint64_t m,
                            uint64_t n, int whole_group, uint32_t sh, const uint32_t* __restrict__ fa,
                            const uint32_t* __restrict__ ta, const uint32_t* __restrict__ fb,
                            const uint32_t* __restrict__ tb, uint32_t* __restrict__ rank, uint32_t* __restrict__ keep,
                            int64_t* __restrict__ sa, uint32_t* __restrict__ err) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint64_t t = k / 4096;
  const uint32_t first = whole_group ? 0u : (fa[k] > ta[t] ? fa[k] : ta[t]);
  const uint32_t head = fb[k] > tb[t] ? fb[k] : tb[t];
  const uint32_t g = whole_group ? 0u : (uint32_t)(key[k] >> sh);
  const uint32_t r = g + (head - first);
  const uint32_t i = val[k];
  // after round 0 the rank changes only past a group's first subgroup
  if (whole_group || r != g) rank[i] = r;
  const bool single = head == (uint32_t)k && (k + 1 == m || key[k + 1] != key[k]);
  keep[k] = single ? 0u : 1u;
  // a suffix already alone in its group has its sa entry from an earlier round
  const bool was_single = !whole_group && first == (uint32_t)k && (k + 1 == m || (key[k + 1] >> sh) != g);
  if (single && !was_single) {
    if (r < n) sa[r] = (int64_t)i;
    else atomicOr(err, WTG_FLAG_SA_RANKS);  // ranks inconsistent: not expected
  }
}

__global__ void k_sa_compact(const uint32_t* __restrict__ val, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ pos, uint64_t m, uint32_t* __restrict__ act) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m && keep[k]) act[pos[k]] = val[k];
}

// bwt[i] = t[sa[i]-1], or t[n-1] where sa[i] = 0 (fm-index/par/WT.C:38-41)
__global__ void k_bwt(const uint8_t* __restrict__ t, const int64_t* __restrict__ sa, uint64_t n,
                      uint8_t* __restrict__ bwt, uint32_t* __restrict__ err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = sa[i];
  if (p < 0 || (uint64_t)p >= n) {  // not a suffix array of this text
    atomicOr(err, WTG_FLAG_BAD_SA);
    bwt[i] = 0;
    return;
  }
  bwt[i] = p ? t[p - 1] : t[n - 1];
}

}  // namespace

extern "C" int wtg_bwt_from_sa_gpu(wtg_ctx* ctx, const uint8_t* d_text, const int64_t* d_sa, uint64_t n,
                                   uint8_t* d_bwt) {
  WTG_REQUIRE(ctx && (n == 0 || (d_text && d_sa && d_bwt)), "wtg_bwt_from_sa_gpu: null argument");
  if (n == 0) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "bwt", k_bwt, dim3((unsigned)wtg_ceil_div(n, 256)), dim3(256), 0, d_text, d_sa, n, d_bwt,
             ctx->d_flag);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

// a failed launch releases the buffers (goto out) instead of returning
#define SA_LAUNCH_CHECK()                                                  \
  do {                                                                     \
    const hipError_t _e = hipGetLastError();                               \
    if (_e != hipSuccess) {                                                \
      wtg_set_error(hipGetErrorString(_e));                                \
      rc = WTG_ERR_HIP;                                                    \
      goto out;                                                            \
    }                                                                      \
  } while (0)

extern "C" int wtg_suffix_array_gpu(wtg_ctx* ctx, const uint8_t* d_text, uint64_t n, int64_t* d_sa) {
  WTG_REQUIRE(ctx && (n == 0 || (d_text && d_sa)), "wtg_suffix_array_gpu: null argument");
  WTG_REQUIRE(n < (1ull << 31), "wtg_suffix_array_gpu: n >= 2^31");
  if (n == 0) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  // Each round sorts only the suffixes whose group is not yet a singleton
  // (ranks of resolved suffixes are final and stay consistent with the order).
  const uint64_t tiles = wtg_ceil_div(n, 4096);
  uint64_t *k0 = nullptr, *k1 = nullptr, *st = nullptr;
  uint32_t *v0 = nullptr, *v1 = nullptr, *act = nullptr, *rank = nullptr, *fa = nullptr, *fb = nullptr;
  uint32_t *ta = nullptr, *tb = nullptr, *keep = nullptr, *pos = nullptr, *ctr = nullptr, *present = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = WTG_OK;
  if ((rc = wtg_malloc(ctx, (void**)&k0, n * 8)) != WTG_OK || (rc = wtg_malloc(ctx, (void**)&k1, n * 8)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&v0, n * 4)) != WTG_OK || (rc = wtg_malloc(ctx, (void**)&v1, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&act, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&rank, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&fa, n * 4)) != WTG_OK || (rc = wtg_malloc(ctx, (void**)&fb, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&ta, tiles * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&tb, tiles * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&keep, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&pos, (n + 1) * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&st, (tiles + 1) * 8)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&ctr, 256)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&present, 256 * 4 + 256 * 2)) != WTG_OK)
    goto out;
  for (unsigned eb : {63u, 64u}) {  // temporary storage for the widest sort
    size_t b = 0;
    if (rocprim::radix_sort_pairs(nullptr, b, k0, k1, v0, v1, (uint32_t)n, 0, eb, ctx->stream) != hipSuccess) {
      wtg_set_error("wtg_suffix_array_gpu: sort size query failed");
      rc = WTG_ERR_HIP;
      goto out;
    }
    tmp_bytes = b > tmp_bytes ? b : tmp_bytes;
  }
  if ((rc = wtg_malloc(ctx, &tmp, tmp_bytes)) != WTG_OK) goto out;
  {
    uint32_t nbits = 1;
    while (nbits < 32 && (1ull << nbits) <= n) nbits++;
    // round 0 codes the sigma distinct bytes as 1..sigma in c bits (order kept),
    // so it sorts by q = floor(63 / c) bytes: 7 for 256 values, 21 for DNA
    uint16_t* d_code = (uint16_t*)(present + 256);  // codes up to 256: 16 bits
    uint32_t h_present[256];
    uint16_t h_code[256];
    if ((rc = wtg_fill_bytes_zero(ctx, present, 256 * 4)) != WTG_OK) goto out;
    WTG_LAUNCH(ctx, "sa_present", k_sa_present, dim3((unsigned)wtg_ceil_div(n, 4096)), dim3(256), 0, d_text, n,
               present);
    SA_LAUNCH_CHECK();
    if ((rc = wtg_stage_d2h(ctx, h_present, present, sizeof(h_present))) != WTG_OK) goto out;
    uint32_t sigma = 0;
    for (int b = 0; b < 256; b++) h_code[b] = h_present[b] ? (uint16_t)++sigma : 0;
    uint32_t cbits = 1;
    while ((1u << cbits) <= sigma) cbits++;
    const uint32_t q = 63 / cbits;
    if (hipMemcpyAsync(d_code, h_code, sizeof(h_code), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      wtg_set_error("wtg_suffix_array_gpu: code table copy failed");
      rc = WTG_ERR_HIP;
      goto out;
    }
    uint64_t m = n, h = 0;
    for (int round = 0; round < 40 && m > 0; round++) {
      if (round == 0) {
        WTG_LAUNCH(ctx, "sa_init", k_sa_init, dim3((unsigned)wtg_ceil_div(n, 256)), dim3(256), 0, d_text, n, d_code,
                   cbits, q, k0, v0);
      } else {
        const bool all = 2 * m > n;  // every suffix in text order
        if (all) m = n;
        WTG_LAUNCH(ctx, "sa_keys", k_sa_keys, dim3((unsigned)wtg_ceil_div(m, 256)), dim3(256), 0, rank,
                   all ? (const uint32_t*)nullptr : act, m, n, h, nbits, k0, v0);
      }
      SA_LAUNCH_CHECK();
      const unsigned end_bit = round == 0 ? q * cbits : 2u * nbits;
      const unsigned grid = (unsigned)wtg_ceil_div(m, 256);
      const uint64_t mt = wtg_ceil_div(m, 4096);
      if (rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k1, v0, v1, (uint32_t)m, 0, end_bit, ctx->stream) !=
          hipSuccess) {
        wtg_set_error("wtg_suffix_array_gpu: sort failed");
        rc = WTG_ERR_HIP;
        goto out;
      }
      const int whole = round == 0 ? 1 : 0;
      WTG_LAUNCH(ctx, "sa_runs", k_sa_runs_tiles, dim3((unsigned)mt), dim3(256), 0, k1, m, whole, nbits, fa, fb, ta,
                 tb);
      SA_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_tile_prefix", k_sa_tile_prefix, dim3(1), dim3(256), 0, tb, mt);
      SA_LAUNCH_CHECK();
      if (!whole) {
        WTG_LAUNCH(ctx, "sa_tile_prefix", k_sa_tile_prefix, dim3(1), dim3(256), 0, ta, mt);
        SA_LAUNCH_CHECK();
      }
      WTG_LAUNCH(ctx, "sa_update", k_sa_update, dim3(grid), dim3(256), 0, k1, v1, m, n, whole, nbits, fa, ta, fb, tb,
                 rank, keep, d_sa, ctx->d_flag);
      SA_LAUNCH_CHECK();
      // compact the unresolved suffixes (exclusive scan of keep)
      const uint64_t sct = wtg_ceil_div(m, 4096);
      if ((rc = wtg_fill_bytes_zero(ctx, st, (sct + 1) * 8)) != WTG_OK) goto out;
      if ((rc = wtg_fill_bytes_zero(ctx, ctr, 4)) != WTG_OK) goto out;
      WTG_LAUNCH(ctx, "sa_keep_scan", k_scan_u32<uint32_t>, dim3((unsigned)sct), dim3(256), 0, keep, pos, m, st, ctr,
                 ctx->d_flag);
      SA_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_compact", k_sa_compact, dim3(grid), dim3(256), 0, v1, keep, pos, m, act);
      SA_LAUNCH_CHECK();
      uint32_t mnext = 0;
      if ((rc = wtg_stage_d2h(ctx, &mnext, pos + m, 4)) != WTG_OK) goto out;
      m = mnext;
      h = h == 0 ? q : 2 * h;
      if (m > 0 && (round == 39 || h >= 2 * n)) {
        wtg_set_error("wtg_suffix_array_gpu: ranks did not separate");
        rc = WTG_ERR_DEVICE;
        goto out;
      }
    }
  }
  rc = wtg_finish(ctx);
out:
  if (rc != WTG_OK) wtg_abandon(ctx);
  wtg_free(ctx, k0); wtg_free(ctx, k1); wtg_free(ctx, v0); wtg_free(ctx, v1); wtg_free(ctx, act);
  wtg_free(ctx, rank); wtg_free(ctx, fa); wtg_free(ctx, fb); wtg_free(ctx, ta); wtg_free(ctx, tb);
  wtg_free(ctx, keep); wtg_free(ctx, pos); wtg_free(ctx, st); wtg_free(ctx, ctr); wtg_free(ctx, tmp);
  wtg_free(ctx, present);
  return rc;
}
#undef SA_LAUNCH_CHECK
