// wtg_sa_gpu.hip -- suffix array of a byte text on the GPU by prefix doubling
// (the reference calls divsufsort, fm-index/par/WT.C:35-36; any exact suffix
// sorter gives the same array).  Round 0 sorts the suffixes by their first 7
// bytes, each coded as byte+1 in 9 bits with 0 past the text end, so a suffix
// that ends sorts before its extensions.  Each later round h = 7, 14, 28, ...
// sorts by (rank of the h-prefix, rank of the h-prefix h bytes on, +1; 0 past
// the end) with a device radix sort; a suffix's rank is the position of its
// group's first suffix in the sorted order, so ranks fit 32 bits and the key
// 64.  It stops once every rank is distinct.
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

// rank of sorted position p = the first position of its key group: heads
// scattered as (p if key[p] != key[p-1] else 0), then an inclusive max-scan;
// here the head values, scanned by the caller
__global__ void k_sa_heads(const uint64_t* __restrict__ key, uint64_t n, uint32_t* __restrict__ head) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  head[p] = (p == 0 || key[p] != key[p - 1]) ? (uint32_t)p : 0u;
}

// inclusive max-scan of head[] within 4096-value tiles (tile maxima to tmax)
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

// ranks by suffix position: rank[idx[p]] = max(head[p], earlier tiles' maximum);
// also counts the distinct groups (heads) into *groups
__global__ void k_sa_rank(const uint32_t* __restrict__ head, const uint32_t* __restrict__ tmax,
                          const uint32_t* __restrict__ idx, const uint64_t* __restrict__ key, uint64_t n,
                          uint32_t* __restrict__ rank, unsigned long long* __restrict__ groups) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = p < n;
  uint32_t r = 0;
  bool h = false;
  if (in) {
    const uint32_t pre = tmax[p / 4096];  // exclusive max over the earlier tiles
    r = head[p] > pre ? head[p] : pre;
    rank[idx[p]] = r;
    h = p == 0 || key[p] != key[p - 1];
  }
  const uint64_t b = wave_ballot(h);
  if (lane_id() == 0 && b) atomicAdd(groups, (unsigned long long)popc64(b));
}

__global__ void k_sa_keys(const uint32_t* __restrict__ rank, uint64_t n, uint64_t h, uint64_t* __restrict__ key,
                          uint32_t* __restrict__ idx) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = ((uint64_t)rank[i] << 32) | (i + h < n ? (uint64_t)rank[i + h] + 1 : 0ull);
  idx[i] = (uint32_t)i;
}

__global__ void k_sa_out(const uint32_t* __restrict__ idx, uint64_t n, int64_t* __restrict__ sa) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sa[i] = (int64_t)idx[i];
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
  const uint64_t tiles = wtg_ceil_div(n, 4096);
  uint64_t *k0 = nullptr, *k1 = nullptr;
  uint32_t *i0 = nullptr, *i1 = nullptr, *head = nullptr, *rank = nullptr, *tmax = nullptr;
  unsigned long long* groups = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = WTG_OK;
  const unsigned grid = (unsigned)wtg_ceil_div(n, 256);
  // sort keys of at most 63 bits (round 0) / 32 + bits(n) bits (later rounds)
  if ((rc = wtg_malloc(ctx, (void**)&k0, n * 8)) != WTG_OK || (rc = wtg_malloc(ctx, (void**)&k1, n * 8)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&i0, n * 4)) != WTG_OK || (rc = wtg_malloc(ctx, (void**)&i1, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&head, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&rank, n * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&tmax, tiles * 4)) != WTG_OK ||
      (rc = wtg_malloc(ctx, (void**)&groups, 8)) != WTG_OK)
    goto out;
  for (unsigned eb : {63u, 64u}) {  // temporary storage for the widest sort
    size_t b = 0;
    if (rocprim::radix_sort_pairs(nullptr, b, k0, k1, i0, i1, (uint32_t)n, 0, eb, ctx->stream) != hipSuccess) {
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
    uint64_t h = 0;
    for (int round = 0; round < 40; round++) {
      if (round == 0) {
        WTG_LAUNCH(ctx, "sa_init", k_sa_init, dim3(grid), dim3(256), 0, d_text, n, k0, i0);
      } else {
        WTG_LAUNCH(ctx, "sa_keys", k_sa_keys, dim3(grid), dim3(256), 0, rank, n, h, k0, i0);
      }
      WTG_LAUNCH_CHECK();
      const unsigned end_bit = round == 0 ? 63u : 32u + nbits;
      if (rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k1, i0, i1, (uint32_t)n, 0, end_bit, ctx->stream) !=
          hipSuccess) {
        wtg_set_error("wtg_suffix_array_gpu: sort failed");
        rc = WTG_ERR_HIP;
        goto out;
      }
      if ((rc = wtg_fill_u64(ctx, groups, 1, 0)) != WTG_OK) goto out;
      WTG_LAUNCH(ctx, "sa_heads", k_sa_heads, dim3(grid), dim3(256), 0, k1, n, head);
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_maxscan", k_sa_maxscan_tiles, dim3((unsigned)tiles), dim3(256), 0, head, n, tmax);
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_tile_prefix", k_sa_tile_prefix, dim3(1), dim3(256), 0, tmax, tiles);
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "sa_rank", k_sa_rank, dim3(grid), dim3(256), 0, head, tmax, i1, k1, n, rank, groups);
      WTG_LAUNCH_CHECK();
      unsigned long long g = 0;
      if (hipMemcpyAsync(&g, groups, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        wtg_set_error("wtg_suffix_array_gpu: group count copy failed");
        rc = WTG_ERR_HIP;
        goto out;
      }
      h = h == 0 ? 7 : 2 * h;
      if (g == n) break;
      if (round == 39 || h >= 2 * n) {
        wtg_set_error("wtg_suffix_array_gpu: ranks did not separate");
        rc = WTG_ERR_DEVICE;
        goto out;
      }
    }
    WTG_LAUNCH(ctx, "sa_out", k_sa_out, dim3(grid), dim3(256), 0, i1, n, d_sa);
    WTG_LAUNCH_CHECK();
  }
  rc = wtg_finish(ctx);
out:
  wtg_free(ctx, k0); wtg_free(ctx, k1); wtg_free(ctx, i0); wtg_free(ctx, i1);
  wtg_free(ctx, head); wtg_free(ctx, rank); wtg_free(ctx, tmax); wtg_free(ctx, groups); wtg_free(ctx, tmp);
  return rc;
}
