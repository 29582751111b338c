// wtg_sa_gpu.hip -- suffix array of a byte text on the GPU by prefix doubling
// (the reference calls divsufsort, fm-index/par/WT.C:35-36; any exact suffix
// sorter gives the same array).  Round 0 sorts all suffixes by their first 7
// bytes, each coded as byte+1 in 9 bits with 0 past the text end, so a suffix
// that ends sorts before its extensions.  A suffix's rank is the position of
// its group's first suffix in the sorted order (32 bits).  Each later round
// h = 7, 14, 28, ... re-sorts only the suffixes still in groups of two or more,
// by (group rank, rank h bytes on + 1; 0 past the end), and splits each group
// in place: new rank = group rank + offset of the subgroup's first suffix in
// the group.  Suffixes left alone in a group keep their final rank and drop out
// of later rounds (the ranks stay consistent, as in Larsson-Sadakane doubling).
// It stops when no group of two or more is left; sa[rank[i]] = i.
#include "wtg_internal.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace {

__global__ void k_sa_init(const uint8_t* __restrict__ t, uint64_t n, uint64_t* __restrict__ key,
                          uint32_t* __restrict__ idx) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = 0;
#pragma unroll
  for (int j = 0; j < 7; j++) k = (k << 9) | (i + j < n ? (uint64_t)t[i + j] + 1 : 0ull);
  key[i] = k;
  idx[i] = (uint32_t)i;
}

// inclusive max-scan of head flags (p where a run starts, else 0) within 4096-value
// tiles (tile maxima to tmax); with the exclusive scan of the tile maxima this
// gives each sorted position the first position of its run (tile maxima to tmax)
__global__ __launch_bounds__(256) void k_sa_maxscan_tiles(uint32_t* __restrict__ head, uint64_t n,
                                                          uint32_t* __restrict__ tmax) {
  __shared__ uint32_t s_w[4];
  const uint64_t base = (uint64_t)blockIdx.x * 4096 + threadIdx.x * 16;
  uint32_t v[16], m = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    v[j] = base + j < n ? head[base + j] : 0u;
    m = v[j] > m ? v[j] : m;
    v[j] = m;
  }
  // wave inclusive max of the thread maxima, then across the 4 waves
  uint32_t x = m;
  const int lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, d);
    if (lane >= d) x = y > x ? y : x;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int k = 0; k < w; k++) pre = s_w[k] > pre ? s_w[k] : pre;
  uint32_t ex = (uint32_t)__shfl_up((int)x, 1);
  if (lane == 0) ex = 0;
  ex = ex > pre ? ex : pre;
#pragma unroll
  for (int j = 0; j < 16; j++)
    if (base + j < n) head[base + j] = v[j] > ex ? v[j] : ex;
  if (threadIdx.x == 0) {
    uint32_t mx = 0;
    for (int k = 0; k < 4; k++) mx = s_w[k] > mx ? s_w[k] : mx;
    tmax[blockIdx.x] = mx;
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

// keys of the unresolved suffixes: (current rank, rank h bytes on + 1; 0 past the end)
__global__ void k_sa_keys_act(const uint32_t* __restrict__ rank, const uint32_t* __restrict__ act, uint64_t m,
                              uint64_t n, uint64_t h, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t i = act[k];
  key[k] = ((uint64_t)rank[i] << 32) | (i + h < n ? (uint64_t)rank[i + h] + 1 : 0ull);
  val[k] = i;
}

// after a sort of m keys: fa = k where the rank part (key >> 32) starts a run,
// fb = k where the whole key starts a run (0 elsewhere; max-scanned next)
__global__ void k_sa_flags(const uint64_t* __restrict__ key, uint64_t m, int whole_group, uint32_t* __restrict__ fa,
                           uint32_t* __restrict__ fb) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const bool b = k == 0 || key[k] != key[k - 1];
  const bool a = !whole_group && (k == 0 || (key[k] >> 32) != (key[k - 1] >> 32));
  fa[k] = a ? (uint32_t)k : 0u;
  fb[k] = b ? (uint32_t)k : 0u;
}

// new rank of sorted position k: the group's old rank (0 in round 0) plus the
// offset of k's subgroup head from the group's first position; keep[k] = 1 when
// k's subgroup has more than one suffix
__global__ void k_sa_update(const uint64_t* __restrict__ key, const uint32_t* __restrict__ val, uint64_t m,
                            int whole_group, const uint32_t* __restrict__ fa, const uint32_t* __restrict__ ta,
                            const uint32_t* __restrict__ fb, const uint32_t* __restrict__ tb,
                            uint32_t* __restrict__ rank, uint32_t* __restrict__ keep) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint64_t t = k / 4096;
  const uint32_t first = whole_group ? 0u : (fa[k] > ta[t] ? fa[k] : ta[t]);
  const uint32_t head = fb[k] > tb[t] ? fb[k] : tb[t];
  const uint32_t g = whole_group ? 0u : (uint32_t)(key[k] >> 32);
  rank[val[k]] = g + (head - first);
  const bool single = head == (uint32_t)k && (k + 1 == m || key[k + 1] != key[k]);
  keep[k] = single ? 0u : 1u;
}

__global__ void k_sa_compact(const uint32_t* __restrict__ val, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ pos, uint64_t m, uint32_t* __restrict__ act) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m && keep[k]) act[pos[k]] = val[k];
}

__global__ void k_sa_final(const uint32_t* __restrict__ rank, uint64_t n, int64_t* __restrict__ sa,
                           uint32_t* __restrict__ err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = rank[i];
  if (r < n) sa[r] = (int64_t)i;
  else atomicOr(err, 2u);
}

// bwt[i] = t[sa[i]-1], or t[n-1] where sa[i] = 0 (fm-index/par/WT.C:38-41)
__global__ void k_bwt(const uint8_t* __restrict__ t, const int64_t* __restrict__ sa, uint64_t n,
                      uint8_t* __restrict__ bwt, uint32_t* __restrict__ err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = sa[i];
  if (p < 0 || (uint64_t)p >= n) {  // not a suffix array of this text
    atomicOr(err, 2u);
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
  uint32_t *ta = nullptr, *tb = nullptr, *keep = nullptr, *pos = nullptr, *ctr = nullptr;
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
      (rc = wtg_malloc(ctx, (void**)&ctr, 256)) != WTG_OK)
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
    uint64_t m = n, h = 0;
    for (int round = 0; round < 40 && m > 0; round++) {
      const unsigned grid = (unsigned)wtg_ceil_div(m, 256);
      const uint64_t mt = wtg_ceil_div(m, 4096);
      if (round == 0) {
        WTG_LAUNCH(ctx, "sa_init", k_sa_init, dim3(grid), dim3(256), 0, d_text, n, k0, v0);
      } else {
        WTG_LAUNCH(ctx, "sa_keys", k_sa_keys_act, dim3(grid), dim3(256), 0, rank, act, m, n, h, k0, v0);
      }
      WTG_LAUNCH_CHECK();
      const unsigned end_bit = round == 0 ? 63u : 32u + nbits;
      if (rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k1, v0, v1, (uint32_t)m, 0, end_bit, ctx->stream) !=
          hipSuccess) {
        wtg_set_error("wtg_suffix_array_gpu: sort failed");
        rc = WTG_ERR_HIP;
        goto out;
      }
      const int whole = round == 0 ? 1 : 0;
      WTG_LAUNCH(ctx, "sa_flags", k_sa_flags, dim3(grid), dim3(256), 0, k1, m, whole, fa, fb);
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_maxscan", k_sa_maxscan_tiles, dim3((unsigned)mt), dim3(256), 0, fb, m, tb);
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_tile_prefix", k_sa_tile_prefix, dim3(1), dim3(256), 0, tb, mt);
      WTG_LAUNCH_CHECK();
      if (!whole) {
        WTG_LAUNCH(ctx, "sa_maxscan", k_sa_maxscan_tiles, dim3((unsigned)mt), dim3(256), 0, fa, m, ta);
        WTG_LAUNCH_CHECK();
        WTG_LAUNCH(ctx, "sa_tile_prefix", k_sa_tile_prefix, dim3(1), dim3(256), 0, ta, mt);
        WTG_LAUNCH_CHECK();
      }
      WTG_LAUNCH(ctx, "sa_update", k_sa_update, dim3(grid), dim3(256), 0, k1, v1, m, whole, fa, ta, fb, tb, rank,
                 keep);
      WTG_LAUNCH_CHECK();
      // compact the unresolved suffixes (exclusive scan of keep)
      const uint64_t sct = wtg_ceil_div(m, 4096);
      if ((rc = wtg_fill_bytes_zero(ctx, st, (sct + 1) * 8)) != WTG_OK) goto out;
      if ((rc = wtg_fill_bytes_zero(ctx, ctr, 4)) != WTG_OK) goto out;
      WTG_LAUNCH(ctx, "sa_keep_scan", k_scan_u32<uint32_t>, dim3((unsigned)sct), dim3(256), 0, keep, pos, m, st, ctr,
                 ctx->d_flag);
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_compact", k_sa_compact, dim3(grid), dim3(256), 0, v1, keep, pos, m, act);
      WTG_LAUNCH_CHECK();
      uint32_t mnext = 0;
      if (hipMemcpyAsync(&mnext, pos + m, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        wtg_set_error("wtg_suffix_array_gpu: count copy failed");
        rc = WTG_ERR_HIP;
        goto out;
      }
      m = mnext;
      h = h == 0 ? 7 : 2 * h;
      if (m > 0 && (round == 39 || h >= 2 * n)) {
        wtg_set_error("wtg_suffix_array_gpu: ranks did not separate");
        rc = WTG_ERR_DEVICE;
        goto out;
      }
    }
    WTG_LAUNCH(ctx, "sa_final", k_sa_final, dim3((unsigned)wtg_ceil_div(n, 256)), dim3(256), 0, rank, n, d_sa,
               ctx->d_flag);
    WTG_LAUNCH_CHECK();
  }
  rc = wtg_finish(ctx);
out:
  wtg_free(ctx, k0); wtg_free(ctx, k1); wtg_free(ctx, v0); wtg_free(ctx, v1); wtg_free(ctx, act);
  wtg_free(ctx, rank); wtg_free(ctx, fa); wtg_free(ctx, fb); wtg_free(ctx, ta); wtg_free(ctx, tb);
  wtg_free(ctx, keep); wtg_free(ctx, pos); wtg_free(ctx, st); wtg_free(ctx, ctr); wtg_free(ctx, tmp);
  return rc;
}
