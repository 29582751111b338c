// wtg_internal.hpp -- context, error plumbing and gfx950 device helpers shared
// by the wavelet-tree / rank / select kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>
#include "wtg.h"

// ------------------------------------------------------------------ errors
void wtg_set_error(const std::string& msg);
#define WTG_CHECK_HIP(expr)                                                      \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      wtg_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));          \
      return WTG_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)
#define WTG_REQUIRE(cond, msg)                                                   \
  do {                                                                           \
    if (!(cond)) { wtg_set_error(msg); return WTG_ERR_ARG; }                     \
  } while (0)
#define WTG_TRY(expr)                                                            \
  do { int _rc = (expr); if (_rc != WTG_OK) return _rc; } while (0)

// ---------------------------------------------------------------- context
struct wtg_kstat { double ms = 0; uint64_t launches = 0; };
struct wtg_kpending { std::string name; hipEvent_t a, b; };

struct wtg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int blocking = 1;
  int level_method = 0;  // 0: multi-level direct passes, 1: level-by-level split
  // growable workspace (device); regions carved per call
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // device error flag (bounded waits set it), host-visible copy on sync
  uint32_t* d_flag = nullptr;
  // pinned host staging for small transfers (async copies need no bounce
  // buffer): [0, 256) the error flag, [WTG_STAGE_D2H, +16 KiB) device-to-host
  // reads (each followed by a stream sync before use), [WTG_STAGE_H2D, +32 KiB)
  // host-to-device tables (h2d_ev marks the last such copy)
  uint8_t* h_stage = nullptr;
  hipEvent_t h2d_ev = nullptr;
  // stream timer
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // per-kernel timing
  bool pooled = false;   // wtg_malloc / wtg_free use the stream-ordered pool
  bool ktime = false;
  std::string kfilter;   // time only launches whose name starts with this (empty: all)
  std::vector<wtg_kpending> kpending;
  std::vector<hipEvent_t> kpool;
  std::map<std::string, wtg_kstat> kstats;
};

// workspace of at least `bytes`, 256-byte aligned regions carved by the caller
int wtg_ws_reserve(wtg_ctx* ctx, size_t bytes);
constexpr size_t WTG_STAGE_AUX = 256;  // [256, 4096): reads collected by a later sync of the call
constexpr size_t WTG_STAGE_D2H = 4096, WTG_STAGE_D2H_MAX = 16384;
constexpr size_t WTG_STAGE_H2D = 32768, WTG_STAGE_H2D_MAX = 32768;
constexpr size_t WTG_STAGE_BYTES = WTG_STAGE_H2D + WTG_STAGE_H2D_MAX;
// copy device bytes to host dst through the pinned staging area (one stream sync)
int wtg_stage_d2h(wtg_ctx* ctx, void* dst, const void* d_src, size_t bytes);
// stream-ordered host-to-device copy of a small table through the staging area
// (waits for the previous such copy before reusing the buffer)
int wtg_stage_h2d(wtg_ctx* ctx, void* d_dst, const void* src, size_t bytes);
// bits of the device error flag (ctx->d_flag), decoded by wtg_ctx_sync
constexpr uint32_t WTG_FLAG_WAIT = 1u;      // a bounded look-back wait expired
constexpr uint32_t WTG_FLAG_BADSYM = 2u;    // input symbol >= 2^levels / tables inconsistent with the keys
constexpr uint32_t WTG_FLAG_SA_RANKS = 4u;  // suffix ranks inconsistent
constexpr uint32_t WTG_FLAG_BAD_SA = 8u;    // a suffix array entry outside [0, n)

// called at the end of every public entry point: sync if blocking, check flag
int wtg_finish(wtg_ctx* ctx);
// error path of an entry point that stops early: wait for the stream and clear
// the device flag, so a flag raised by this call does not fail the next one
// (the error already set is kept)
void wtg_abandon(wtg_ctx* ctx);
int wtg_set_device(wtg_ctx* ctx);

// kernel launch with optional per-kernel event timing
struct wtg_kscope {
  wtg_ctx* ctx; const char* name; hipEvent_t a = nullptr, b = nullptr;
  wtg_kscope(wtg_ctx* c, const char* n);
  ~wtg_kscope();
};
#define WTG_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                   \
  do {                                                                           \
    wtg_kscope _ks((ctx), (name));                                               \
    hipLaunchKernelGGL(kernel, (grid), (block), (shmem), (ctx)->stream,          \
                       __VA_ARGS__);                                             \
  } while (0)
#define WTG_LAUNCH_CHECK()                                                       \
  WTG_CHECK_HIP(hipGetLastError())

int wtg_fill_u64(wtg_ctx* ctx, void* d, uint64_t n_u64, uint64_t value);
// internal entry points shared between translation units
int wtg_levels_build_async(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n,
                           uint32_t levels, uint64_t* d_words, wtg_node* d_nodes,
                           uint64_t node_slots, unsigned long long* d_distinct);
// level pipeline over levels [start, levels) of keys already grouped by their
// top `start` bits (counts in d_cnt_start[2^start]); level k at bit (k-start)*n
// (d_ones_start: the first level's ones per node when the caller has them)
int wtg_levels_build_from(wtg_ctx* ctx, const void* d_keys, int key_bytes, uint64_t n, uint32_t levels,
                          uint32_t start, const uint32_t* d_cnt_start, uint64_t* d_words,
                          const uint32_t* d_ones_start = nullptr);
size_t wtg_levels_ws_bytes(uint64_t n, uint32_t levels, int key_bytes);
int wtg_rank_select_run(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir,
                        bool build, wtg_select* sel1, wtg_select* sel0);
// The same in two halves, for callers that fold the long-superblock counts'
// host read into a read of their own: begin enqueues everything up to the two
// counts (at pend->d_nl; the workspace must stay untouched until end), end
// allocates and fills the long pools.  known_ones (build only): the number of
// ones when the caller knows it, which saves a host read.
// For many vectors at once (one host read per half for all of them): give
// each its own workspace (ws_ext, wtg_rank_select_ws_bytes) and an arena for
// the select content, and pass known_ones = WTG_ONES_DEFER: begin then stops
// after the rank pass, with the number of ones at ws_ext + WTG_RS_ONES_OFF
// (uint64), and wtg_rank_select_count, given that number, enqueues the rest up
// to the long counts (uint32 per structure, bit t at ws_ext + WTG_RS_NL_OFF + 4t).
constexpr uint64_t WTG_ONES_UNKNOWN = ~0ull;
constexpr uint64_t WTG_ONES_DEFER = ~0ull - 1;
constexpr size_t WTG_RS_ONES_OFF = 64, WTG_RS_NL_OFF = 192;
// bump allocation from one device block; what it holds is freed with the
// block, never by wtg_select_free
struct wtg_arena {
  char* base;
  size_t size, used;
};
struct wtg_rs_pending {
  bool active;
  const uint64_t* d_words;
  uint64_t nbits;
  const uint32_t* d_nl;
  wtg_select* outs[2];
  const uint64_t* sbend[2];
  const unsigned long long* last[2];
  const char* fin[2];
  wtg_arena* ar;
  bool want[2];
};
int wtg_rank_select_begin(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir, bool build,
                          wtg_select* sel1, wtg_select* sel0, uint64_t known_ones, wtg_rs_pending* pend,
                          char* ws_ext = nullptr, wtg_arena* ar = nullptr);
int wtg_rank_select_count(wtg_ctx* ctx, wtg_rs_pending* pend, uint64_t ones);
int wtg_rank_select_end(wtg_ctx* ctx, const wtg_rs_pending* pend, const uint32_t nl[2]);
size_t wtg_rank_select_ws_bytes(uint64_t nbits);
// arena bytes of one select structure: before the count (by nbits), after it
// (by the number of ones), and its long pool (by the long count)
size_t wtg_select_arena_worst(uint64_t nbits);
size_t wtg_select_arena_count(uint64_t ones);
size_t wtg_select_arena_pool(uint32_t n_long);
// multi-level direct build (up to 8 levels per pass; n >= 1, L <= 24, n < 2^32)
// and whether it is on (WTG_LEVEL_SPLIT=1 selects the level-by-level split)
int wtg_mld_enabled();
int wtg_mld_build_from(wtg_ctx* ctx, const void* d_keys, int key_bytes, uint64_t n, uint32_t L, uint32_t s0,
                       const uint32_t* d_cnt0, const uint64_t* d_src_off, uint64_t* d_words);
size_t wtg_mld_ws_bytes(uint64_t n, uint32_t L, uint32_t s0);
// the top Lt <= 24 levels of an L-level tree (keys d_sym >> (L - Lt)), level l at bit l * n
int wtg_mld_build_top(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint32_t Lt,
                      uint64_t* d_words);
// levels 25..64: nodes numbered by rank within the level (wtg_deep.hip)
int wtg_deep_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint64_t* d_words,
                   unsigned long long* d_distinct);
int wtg_mld_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint64_t* d_words,
                  wtg_node* d_nodes, uint64_t node_slots, unsigned long long* d_distinct);
// resolves the per-device address of the allocation-free rank9 scratch (wtg_ctx_create)
int wtg_noalloc_init_device(int device);
int wtg_max_value(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint64_t* out);
int wtg_fill_bytes_zero(wtg_ctx* ctx, void* d, uint64_t bytes);

static inline size_t wtg_align256(size_t x) { return (x + 255) & ~size_t(255); }
static inline uint64_t wtg_ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static inline int wtg_log2up_host(uint64_t x) {
  int a = 0; if (x == 0) return 0; uint64_t b = x - 1; while (b) { b >>= 1; a++; } return a;
}
static inline int wtg_hi_host(uint64_t x) { int r = 0; if (!x) return 0; while (x >>= 1) r++; return r; }

// ------------------------------------------------------------- device side
#define WTG_WAVE 64
#define WTG_SPIN_LIMIT (1u << 26)   // bounded waits: iterations before giving up

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t popc64(uint64_t x) { return (uint32_t)__popcll(x); }

template <typename U>
__device__ __forceinline__ U wave_sum(U v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d);
  return v;
}

// Inclusive scans over the 64 lanes with DPP moves (no LDS round trip): row_shr
// 1, 2, 4, 8 scan each 16-lane row (lanes with no source in the row take the
// identity), then row_bcast 15 combines row 0's last lane into row 1 and row 2's
// into row 3, and row_bcast 31 the last lane of rows 0-1 into rows 2-3.
template <int CTRL, int RM>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t ident, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)v, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_f64(double ident, double v) {
  const uint64_t u = (uint64_t)__double_as_longlong(v), o = (uint64_t)__double_as_longlong(ident);
  const uint32_t lo = dpp_u32<CTRL, RM>((uint32_t)o, (uint32_t)u);
  const uint32_t hi = dpp_u32<CTRL, RM>((uint32_t)(o >> 32), (uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
#define WTG_DPP_SCAN_STEPS(STEP) STEP(0x111, 0xf) STEP(0x112, 0xf) STEP(0x114, 0xf) STEP(0x118, 0xf) \
  STEP(0x142, 0xa) STEP(0x143, 0xc)
__device__ __forceinline__ uint32_t wave_incl_sum_dpp(uint32_t v) {
#define WTG_STEP(C, R) v += dpp_u32<C, R>(0u, v);
  WTG_DPP_SCAN_STEPS(WTG_STEP)
#undef WTG_STEP
  return v;
}
__device__ __forceinline__ int32_t wave_incl_max_dpp(int32_t v) {
#define WTG_STEP(C, R) { const int32_t t = (int32_t)dpp_u32<C, R>(0x80000000u, (uint32_t)v); v = t > v ? t : v; }
  WTG_DPP_SCAN_STEPS(WTG_STEP)
#undef WTG_STEP
  return v;
}
// inclusive scan over the 64 lanes, Hillis-Steele with shuffles
template <typename U>
__device__ __forceinline__ U wave_incl_scan_shfl(U v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    U t = __shfl_up(v, d);
    if (l >= d) v += t;
  }
  return v;
}
// inclusive scan over the 64 lanes: 32-bit values by DPP moves, wider ones by
// shuffles
template <typename U>
__device__ __forceinline__ U wave_incl_scan(U v) {
  if constexpr (sizeof(U) == 4) {
    return (U)wave_incl_sum_dpp((uint32_t)v);
  } else {
    return wave_incl_scan_shfl(v);
  }
}
// running max of a and min of b over the lanes (no NaNs, no -0: max and min
// are exact and equal the sequential walk's); a lane with no source in a step
// reads its own value
__device__ __forceinline__ void wave_incl_maxmin_f64_dpp(double& a, double& b) {
#define WTG_STEP(C, R) { const double ta = dpp_f64<C, R>(a, a), tb = dpp_f64<C, R>(b, b); \
    a = __builtin_fmax(ta, a); b = __builtin_fmin(tb, b); }
  WTG_DPP_SCAN_STEPS(WTG_STEP)
#undef WTG_STEP
}

// position (0..63) of the k-th (0-based) set bit of x; x must have > k ones
__device__ __forceinline__ uint32_t select_in_word(uint64_t x, uint32_t k) {
  uint32_t pos = 0;
  uint32_t c = __popc((uint32_t)x);
  if (k >= c) { k -= c; pos += 32; x >>= 32; }
  c = __popc((uint32_t)x & 0xFFFFu);
  if (k >= c) { k -= c; pos += 16; x >>= 16; }
  c = __popc((uint32_t)x & 0xFFu);
  if (k >= c) { k -= c; pos += 8; x >>= 8; }
  c = __popc((uint32_t)x & 0xFu);
  if (k >= c) { k -= c; pos += 4; x >>= 4; }
  c = __popc((uint32_t)x & 0x3u);
  if (k >= c) { k -= c; pos += 2; x >>= 2; }
  c = (uint32_t)(x & 1u);
  if (k >= c) { pos += 1; }
  return pos;
}

// ---------------------------------------------------- decoupled look-back
// status word = flag<<62 | value; flag 1 = tile aggregate, 2 = inclusive prefix.
#define LB_AGG 1ull
#define LB_PRE 2ull
#define LB_VMASK ((1ull << 62) - 1)

__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t flag, uint64_t v) {
  __hip_atomic_store(p, (flag << 62) | (v & LB_VMASK), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_load(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes the aggregate of `tile` (one lane).  Tile 0 publishes its
// inclusive prefix directly.
__device__ __forceinline__ void lb_publish(uint64_t* states, uint64_t tile, uint64_t agg) {
  lb_store(&states[tile], tile == 0 ? LB_PRE : LB_AGG, agg);
}

// Called by ONE full wave (all 64 lanes) after lb_publish: walks back over the
// predecessors, publishes the inclusive prefix and returns the exclusive
// prefix (same in all lanes).  Every wait is bounded: on expiry the device
// error flag is raised and the partial sum is used, so the kernel always drains.
__device__ __forceinline__ uint64_t lb_wait(uint64_t* states, uint64_t tile, uint64_t agg,
                                            uint32_t* err) {
  const int l = lane_id();
  if (tile == 0) return 0;
  uint64_t excl = 0;
  int64_t base = (int64_t)tile - 1;
  uint32_t spins = 0;
  while (true) {
    int64_t j = base - l;
    uint64_t s = (j >= 0) ? lb_load(&states[j]) : (LB_PRE << 62);
    while (true) {
      bool notready = (s >> 62) == 0;
      if (wave_ballot(notready) == 0) break;
      if (++spins > WTG_SPIN_LIMIT) {
        if (l == 0) atomicOr(err, WTG_FLAG_WAIT);
        s = (LB_PRE << 62);  // give up: treat as prefix 0 (results invalid, flag raised)
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      if (notready) s = lb_load(&states[j]);
    }
    uint64_t pm = wave_ballot((s >> 62) == LB_PRE);
    uint32_t first = pm ? (uint32_t)__builtin_ctzll(pm) : 64u;
    uint64_t v = ((uint32_t)l <= first) ? (s & LB_VMASK) : 0ull;
    excl += wave_sum(v);
    if (pm) break;
    base -= 64;
  }
  if (l == 0) lb_store(&states[tile], LB_PRE, excl + agg);
  return excl;
}

// publish + wait in one call (one full wave)
__device__ __forceinline__ uint64_t lb_exclusive(uint64_t* states, uint64_t tile, uint64_t agg,
                                                 uint32_t* err) {
  if (lane_id() == 0) lb_publish(states, tile, agg);
  return lb_wait(states, tile, agg, err);
}

// ------------------------------------------------------ block scan helper
// exclusive scan of one u32 per thread over a 256-thread block (the wave scans
// by shuffles or, DPP, by DPP moves)
template <bool DPP = false>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* s_wave,
                                                        uint32_t& total) {
  const int w = threadIdx.x >> 6, l = lane_id();
  uint32_t incl = DPP ? wave_incl_sum_dpp(v) : wave_incl_scan_shfl(v);
  if (l == 63) s_wave[w] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t t = s_wave[k];
    if (k < w) off += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return incl - v + off;
}

// ------------------------------------------- exclusive scan of a u32 array
// out[0..n] (out[n] = total); 256 threads x 16 items per tile; OUT = uint32_t
// or uint64_t (the running sum is 64-bit either way)
template <typename OUT>
__global__ __launch_bounds__(256) void k_scan_u32(const uint32_t* __restrict__ in,
                                                  OUT* __restrict__ out, uint64_t n,
                                                  uint64_t* states, uint32_t* counter,
                                                  uint32_t* err) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_pre;
  if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t base = tile * 4096 + (uint64_t)threadIdx.x * 16;
  uint32_t v[16];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan_u32(sum, s_wave, tot);
  if (threadIdx.x < 64) {
    uint64_t pre = lb_exclusive(states, tile, tot, err);
    if (threadIdx.x == 0) s_pre = pre;
  }
  __syncthreads();
  uint64_t run = s_pre + ex;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    if (base + k < n) out[base + k] = (OUT)run;
    run += v[k];
  }
  if (base < n && base + 16 >= n) out[n] = (OUT)run;
  if (n == 0 && tile == 0 && threadIdx.x == 0) out[0] = 0;
}


// ------------------------------------------------ rank / select device side
// rank_support_v<1,1>::rank (rank/rank_support_v_par.hpp:131-142)
__device__ __forceinline__ uint64_t rank9_dev(const uint64_t* __restrict__ D,
                                              const uint64_t* __restrict__ words, uint64_t i) {
  const uint64_t b = i >> 9;
  uint64_t r = D[2 * b] + ((D[2 * b + 1] >> (63 - 9 * ((i & 511) >> 6))) & 0x1FF);
  if (i & 63) r += popc64(words[i >> 6] & ((1ull << (i & 63)) - 1));
  return r;
}
__device__ __forceinline__ int bit_dev(const uint64_t* __restrict__ words, uint64_t i) {
  return (int)((words[i >> 6] >> (i & 63)) & 1ull);
}

struct SelQ {
  const uint64_t* superblock;
  const uint32_t* mini;
  const int64_t* long_index;
  const uint64_t* long_pool;
  const uint64_t* words;
  uint64_t W;
  int bit;
};

__device__ __forceinline__ uint64_t select_plain(const SelQ& s, uint64_t i) {
  // select/select_support_mcl_par.hpp:564-611
  i -= 1;
  const uint64_t sb = i >> 12, off = i & 0xFFF;
  const int64_t li = s.long_index[sb];
  if (li >= 0) return s.long_pool[(uint64_t)li * 4096 + off];
  const uint64_t base = s.superblock[sb] + s.mini[sb * 64 + (off >> 6)];
  uint32_t r = (uint32_t)(off & 63);
  if (r == 0) return base;
  uint64_t pos = base + 1;
  uint64_t wp = pos >> 6;
  const uint32_t wo = (uint32_t)(pos & 63);
  if (wp >= s.W) return WTG_NONE;
  uint64_t x = s.words[wp];
  if (!s.bit) x = ~x;
  x >>= wo;
  uint32_t c = popc64(x);
  if (c >= r) return (wp << 6) + wo + select_in_word(x, r - 1);
  r -= c;
  for (;;) {
    ++wp;
    if (wp >= s.W) return WTG_NONE;
    x = s.words[wp];
    if (!s.bit) x = ~x;
    c = popc64(x);
    if (c >= r) return (wp << 6) + select_in_word(x, r - 1);
    r -= c;
  }
}

