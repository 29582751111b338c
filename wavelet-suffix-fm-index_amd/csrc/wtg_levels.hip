// wtg_levels.hip -- (a) level-wise wavelet-tree construction on gfx950.
//
// Reference semantics: wavelet/levelWT/WT.C:43-227 (== serialWT/WT.C:32-103):
// level l holds P_l = the stable sort of the symbols by x >> (L-l); the bit at
// global position l*n+i is (P_l[i] >> (L-1-l)) & 1.
//
// GPU pipeline for one build (all integer, HBM bound):
//   1. k_hist       histogram of symbol values (2^L bins)
//   2. k_scan_u32   C = exclusive prefix of the histogram (node starts)
//   3. k_nodetab    per level l < L-1 and node prefix p: A(p) = S(p)-Zs(p),
//                   B(p) = nz(p)+Zs(p) (S = node start, nz = zeros of the node,
//                   Zs = zeros of all earlier nodes of the level)
//   4. k_nodes      WTnode records (optional)
//   5. k_level<l>   per level: 16-byte vector loads, the level's bit words
//                   formed from wave ballots, zero counts scanned with mbcnt +
//                   LDS + a decoupled look-back, and the stable scatter
//                   dest = bit ? i - Z(i) + B(p) : Z(i) + A(p)
//                   into the next ping-pong buffer (last level: bits only).
#include "wtg_internal.hpp"

#define FLAG_LOOKBACK 1u
#define FLAG_BADSYM 2u

// ------------------------------------------------------------ histogram
template <typename T>
__global__ __launch_bounds__(256) void k_hist_lds(const T* __restrict__ s, uint64_t n,
                                                  uint32_t* __restrict__ hist, uint32_t nbins,
                                                  uint32_t* err) {
  extern __shared__ uint32_t sh[];
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (; i < n; i += stride) {
    uint64_t x = (uint64_t)s[i];
    if (x < nbins) atomicAdd(&sh[x], 1u); else bad = true;
  }
  if (bad) atomicOr(err, FLAG_BADSYM);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x)
    if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_hist_global(const T* __restrict__ s, uint64_t n,
                                                     uint32_t* __restrict__ hist, uint32_t nbins,
                                                     uint32_t* err) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (; i < n; i += stride) {
    uint64_t x = (uint64_t)s[i];
    if (x < nbins) atomicAdd(&hist[x], 1u); else bad = true;
  }
  if (bad) atomicOr(err, FLAG_BADSYM);
}

__global__ void k_count_nonzero(const uint32_t* __restrict__ h, uint64_t n,
                                unsigned long long* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (; i < n; i += stride) c += h[i] != 0;
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

// ------------------------------------------------------------ node tables
struct NodeTabArgs {
  const uint32_t* C;     // exclusive prefix of the 2^L-bin histogram, C[2^L] = n
  uint2* tab;            // level l (< L-1) at offset 2^l - 1
  uint64_t* states;      // per tile (tiles of one level are contiguous)
  uint32_t* counter;
  uint32_t* err;
  uint32_t L;
  uint32_t tile_base[25];  // first tile of level l; tile_base[L-1] = total tiles
};

__global__ __launch_bounds__(256) void k_nodetab(NodeTabArgs a) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_pre;
  if (threadIdx.x == 0) s_tile = atomicAdd(a.counter, 1u);
  __syncthreads();
  const uint32_t t = s_tile;
  uint32_t l = 0;
  while (l + 1 < a.L - 1 && a.tile_base[l + 1] <= t) l++;
  const uint32_t tile_in_level = t - a.tile_base[l];
  const uint64_t nnodes = 1ull << l;
  const uint32_t R = 1u << (a.L - l);
  const uint64_t p0 = (uint64_t)tile_in_level * 4096 + (uint64_t)threadIdx.x * 16;
  uint32_t S[16], nz[16];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint64_t p = p0 + k;
    if (p < nnodes) {
      uint32_t st = a.C[p * R], mid = a.C[p * R + R / 2];
      S[k] = st; nz[k] = mid - st;
    } else { S[k] = 0; nz[k] = 0; }
    sum += nz[k];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan_u32(sum, s_wave, tot);
  if (threadIdx.x < 64) {
    uint64_t pre = lb_exclusive(a.states + a.tile_base[l], tile_in_level, tot, a.err);
    if (threadIdx.x == 0) s_pre = pre;
  }
  __syncthreads();
  uint32_t zs = (uint32_t)(s_pre + ex);
  uint2* tab = a.tab + (nnodes - 1);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint64_t p = p0 + k;
    if (p < nnodes) tab[p] = make_uint2(S[k] - zs, nz[k] + zs);
    zs += nz[k];
  }
}

// ---------------------------------------------------------------- WTnodes
__global__ void k_nodes(const uint32_t* __restrict__ C, uint32_t L, uint64_t n,
                        wtg_node* __restrict__ nodes, uint64_t slots) {
  uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= slots) return;
  wtg_node r;
  r.bitmapPtr = 0; r.length = 0; r.parent = WTG_NONE; r.leftChild = WTG_NONE; r.rightChild = WTG_NONE;
  uint64_t heap_end = (1ull << L) - 1;
  if (v < heap_end) {
    uint32_t l = 63 - __clzll(v + 1);
    uint64_t p = v + 1 - (1ull << l);
    uint64_t R = 1ull << (L - l);
    uint64_t st = C[p * R], en = C[(p + 1) * R];
    if (en > st || v == 0) {
      r.bitmapPtr = (uint64_t)l * n + st;
      r.length = en - st;
      r.parent = v ? (v - 1) / 2 : WTG_NONE;
      if (l + 1 < L) {
        uint64_t mid = C[p * R + R / 2];
        if (mid > st) r.leftChild = 2 * v + 1;
        if (en > mid) r.rightChild = 2 * v + 2;
      }
    }
  }
  nodes[v] = r;
}

// ---------------------------------------------------------------- levels
// spread the low G bits of c so that bit i lands at bit i*V (V = 64/G)
template <int V>
__device__ __forceinline__ uint64_t spread_bits(uint64_t x) {
  if constexpr (V == 2) {
    x &= 0xFFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
  } else if constexpr (V == 4) {
    x &= 0xFFFFull;
    x = (x | (x << 24)) & 0x000000FF000000FFull;
    x = (x | (x << 12)) & 0x000F000F000F000Full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
  } else if constexpr (V == 8) {
    x &= 0xFFull;
    x = (x | (x << 28)) & 0x0000000F0000000Full;
    x = (x | (x << 14)) & 0x0003000300030003ull;
    x = (x | (x << 7)) & 0x0101010101010101ull;
  } else {  // V == 16
    x &= 0xFull;
    x = (x | (x << 30)) & 0x0000000300000003ull;
    x = (x | (x << 15)) & 0x0001000100010001ull;
  }
  return x;
}

struct LevelArgs {
  const void* src;
  void* dst;
  uint64_t* words;
  uint64_t bit_base;     // l*n
  uint32_t n;
  uint32_t bitpos;       // L-1-l
  uint32_t node_shift;   // L-l (node prefix p = x >> node_shift)
  uint32_t nnodes;       // 2^l
  const uint2* tab;      // this level's (A,B) table
  uint64_t* states;
  uint32_t* counter;
  uint32_t* err;
};

template <typename T>
struct VecOf;
template <> struct VecOf<uint8_t> { static constexpr int V = 16; };
template <> struct VecOf<uint16_t> { static constexpr int V = 8; };
template <> struct VecOf<uint32_t> { static constexpr int V = 4; };
template <> struct VecOf<uint64_t> { static constexpr int V = 2; };

#define LV_STEPS 4

template <typename T, bool SCATTER, bool VEC>
__global__ __launch_bounds__(256) void k_level(LevelArgs a) {
  constexpr int V = VecOf<T>::V;
  constexpr int G = 64 / V;              // lanes per bit word
  constexpr int TILE = 256 * LV_STEPS * V;
  constexpr int NW = TILE / 64;          // words per tile
  constexpr int NPL = (V == 16) ? 5 : (V == 8) ? 4 : (V == 4) ? 3 : 2;  // bit-planes of a count <= V
  __shared__ uint64_t s_words[NW + 1];
  __shared__ uint32_t s_wz[LV_STEPS * 4];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_pre;

  const int tid = threadIdx.x, w = tid >> 6, lane = lane_id();
  uint32_t tile;
  if (SCATTER) {
    if (tid == 0) s_tile = atomicAdd(a.counter, 1u);
    __syncthreads();
    tile = s_tile;
  } else {
    tile = blockIdx.x;
  }
  const uint64_t tile_start = (uint64_t)tile * TILE;
  const uint64_t n = a.n;
  const uint64_t valid = (n - tile_start < (uint64_t)TILE) ? n - tile_start : (uint64_t)TILE;
  const bool full = valid == (uint64_t)TILE;
  const T* src = (const T*)a.src;

  T x[LV_STEPS][V];
#pragma unroll
  for (int s = 0; s < LV_STEPS; s++) {
    const uint64_t e0 = tile_start + (uint64_t)(s * 256 + tid) * V;
    if (VEC && full) {
      const uint4 q = *(const uint4*)(src + e0);
      const T* qp = (const T*)&q;
#pragma unroll
      for (int v = 0; v < V; v++) x[s][v] = qp[v];
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) x[s][v] = (e0 + v < n) ? src[e0 + v] : (T)0;
    }
  }

  uint32_t mk[LV_STEPS], vmk[LV_STEPS], pre[LV_STEPS];
#pragma unroll
  for (int s = 0; s < LV_STEPS; s++) {
    const uint64_t e0 = tile_start + (uint64_t)(s * 256 + tid) * V;
    uint32_t m = 0, vm = 0;
#pragma unroll
    for (int v = 0; v < V; v++) {
      m |= (uint32_t)(((uint64_t)x[s][v] >> a.bitpos) & 1u) << v;
      vm |= (uint32_t)(e0 + v < n) << v;
    }
    m &= vm;
    mk[s] = m;
    vmk[s] = vm;
    // bit words of this (step, wave): lane j < V assembles word j from V ballots
    uint64_t word = 0;
    const int j = lane & (V - 1);
#pragma unroll
    for (int v = 0; v < V; v++) {
      uint64_t B = wave_ballot((m >> v) & 1u);
      word |= spread_bits<V>(B >> (j * G)) << v;
    }
    if (lane < V) s_words[(s * 4 + w) * V + lane] = word;
    // wave-exclusive prefix of per-lane zero counts via bit-plane ballots
    const uint32_t zc = (uint32_t)__popc(~m & vm);
    uint32_t p = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < NPL; b++) {
      uint64_t Bp = wave_ballot((zc >> b) & 1u);
      p += mbcnt64(Bp) << b;
      tot += popc64(Bp) << b;
    }
    pre[s] = p;
    if (lane == 0) s_wz[s * 4 + w] = tot;
  }
  __syncthreads();

  // ---- bit words out (funnel shift when the level base is not 64-aligned)
  {
    const uint64_t gb = a.bit_base + tile_start;
    const uint32_t sh = (uint32_t)(gb & 63);
    const uint64_t q0 = gb >> 6;
    const uint64_t qlast = (gb + valid - 1) >> 6;
    const uint32_t nout = (uint32_t)(qlast - q0 + 1);
    for (uint32_t t = tid; t < nout; t += 256) {
      uint64_t cur = (t < (uint32_t)NW) ? s_words[t] : 0ull;
      uint64_t val;
      if (sh == 0) val = cur;
      else {
        uint64_t prev = (t > 0) ? s_words[t - 1] : 0ull;
        val = (cur << sh) | (prev >> (64 - sh));
      }
      const uint64_t q = q0 + t;
      const bool fullword = (q * 64 >= gb) && (q * 64 + 64 <= gb + valid);
      if (fullword) a.words[q] = val;
      else if (val) atomicOr((unsigned long long*)&a.words[q], (unsigned long long)val);
    }
  }
  if (!SCATTER) return;

  // ---- tile zero total and look-back for the global zero prefix
  uint32_t woff[LV_STEPS];
  uint32_t ttot = 0;
  {
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < LV_STEPS * 4; k++) {
      const uint32_t t = s_wz[k];
#pragma unroll
      for (int s = 0; s < LV_STEPS; s++)
        if (k == s * 4 + w) woff[s] = run;
      run += t;
    }
    ttot = run;
  }
  if (w == 0) {
    uint64_t p = lb_exclusive(a.states, tile, ttot, a.err);
    if (lane == 0) s_pre = p;
  }
  __syncthreads();
  const uint64_t zprev = s_pre;

  // ---- stable scatter into the next level's order
  T* dst = (T*)a.dst;
#pragma unroll
  for (int s = 0; s < LV_STEPS; s++) {
    const uint64_t e0 = tile_start + (uint64_t)(s * 256 + tid) * V;
    const uint32_t m = mk[s], vm = vmk[s];
    const uint32_t zbase = woff[s] + pre[s];
#pragma unroll
    for (int v = 0; v < V; v++) {
      if (!((vm >> v) & 1u)) continue;
      const uint64_t i = e0 + v;
      const uint64_t Z = zprev + zbase + (uint32_t)__popc(~m & vm & ((1u << v) - 1u));
      uint64_t pnode = (uint64_t)x[s][v] >> a.node_shift;
      if (pnode >= a.nnodes) { atomicOr(a.err, FLAG_BADSYM); continue; }
      const uint2 ab = a.tab[pnode];
      const uint64_t dest = ((m >> v) & 1u) ? (i - Z + ab.y) : (Z + ab.x);
      if (dest < n) dst[dest] = x[s][v];
      else atomicOr(a.err, FLAG_BADSYM);
    }
  }
}

// ================================================================== host
namespace {
struct LevelPlan {
  uint32_t L;
  uint64_t n;
  int w;
  uint64_t tiles;       // tiles per level
  uint32_t tile_elems;
  size_t off_hist, off_C, off_tab, off_scanst, off_ntst, off_states, off_ctr, off_buf0, off_buf1, total;
  uint64_t scan_tiles;
  uint32_t nt_tile_base[25];
  uint32_t nt_tiles;
};

int tile_elems_for(int w) { return 256 * LV_STEPS * (16 / w); }

LevelPlan plan_levels(uint64_t n, uint32_t L, int w) {
  LevelPlan P{};
  P.L = L; P.n = n; P.w = w;
  P.tile_elems = tile_elems_for(w);
  P.tiles = wtg_ceil_div(n, P.tile_elems);
  const uint64_t nb = 1ull << L;
  P.scan_tiles = wtg_ceil_div(nb, 4096);
  // node-table tiles: level l < L-1 has ceil(2^l/4096) tiles
  uint32_t t = 0;
  for (uint32_t l = 0; l + 1 < L; l++) {
    P.nt_tile_base[l] = t;
    t += (uint32_t)wtg_ceil_div(1ull << l, 4096);
  }
  for (uint32_t l = (L > 1 ? L - 1 : 0); l < 25; l++) P.nt_tile_base[l] = t;
  P.nt_tiles = t;
  size_t o = 0;
  // everything from off_hist up to off_buf0 is cleared at the start of a build
  P.off_hist = o; o = wtg_align256(o + nb * 4);
  P.off_C = o; o = wtg_align256(o + (nb + 1) * 4);
  P.off_scanst = o; o = wtg_align256(o + (P.scan_tiles + 1) * 8);
  P.off_ntst = o; o = wtg_align256(o + ((size_t)t + 1) * 8);
  P.off_states = o; o = wtg_align256(o + (size_t)L * (P.tiles + 1) * 8);
  P.off_ctr = o; o = wtg_align256(o + 64 * 4);
  P.off_tab = o; o = wtg_align256(o + (nb / 2 + 1) * 8);
  P.off_buf0 = o; o = wtg_align256(o + n * w);
  P.off_buf1 = o; o = wtg_align256(o + n * w);
  P.total = o;
  return P;
}

template <typename T>
int launch_hist(wtg_ctx* ctx, const T* s, uint64_t n, uint32_t* hist, uint32_t nb) {
  if (n == 0) return WTG_OK;
  uint64_t blocks = wtg_ceil_div(n, 256 * 16);
  if (blocks > 4096) blocks = 4096;
  if (nb <= 16384) {
    WTG_LAUNCH(ctx, "hist", k_hist_lds<T>, dim3((unsigned)blocks), dim3(256), nb * 4, s, n, hist, nb,
               ctx->d_flag);
  } else {
    WTG_LAUNCH(ctx, "hist", k_hist_global<T>, dim3((unsigned)blocks), dim3(256), 0, s, n, hist, nb,
               ctx->d_flag);
  }
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}

template <typename T>
int launch_level(wtg_ctx* ctx, const LevelArgs& la, uint64_t tiles, bool scatter, bool vec) {
  dim3 g((unsigned)tiles), b(256);
  if (scatter) {
    if (vec) WTG_LAUNCH(ctx, "level_split", (k_level<T, true, true>), g, b, 0, la);
    else WTG_LAUNCH(ctx, "level_split", (k_level<T, true, false>), g, b, 0, la);
  } else {
    if (vec) WTG_LAUNCH(ctx, "level_bits", (k_level<T, false, true>), g, b, 0, la);
    else WTG_LAUNCH(ctx, "level_bits", (k_level<T, false, false>), g, b, 0, la);
  }
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}

template <typename T>
int levels_build_t(wtg_ctx* ctx, const T* d_sym, uint64_t n, uint32_t L, uint64_t* d_words,
                   wtg_node* d_nodes, uint64_t node_slots, unsigned long long* d_distinct) {
  const int w = sizeof(T);
  LevelPlan P = plan_levels(n, L, w);
  WTG_TRY(wtg_ws_reserve(ctx, P.total));
  char* ws = (char*)ctx->ws;
  uint32_t* hist = (uint32_t*)(ws + P.off_hist);
  uint32_t* C = (uint32_t*)(ws + P.off_C);
  uint2* tab = (uint2*)(ws + P.off_tab);
  uint64_t* scanst = (uint64_t*)(ws + P.off_scanst);
  uint64_t* ntst = (uint64_t*)(ws + P.off_ntst);
  uint64_t* states = (uint64_t*)(ws + P.off_states);
  uint32_t* ctr = (uint32_t*)(ws + P.off_ctr);
  void* buf[2] = {ws + P.off_buf0, ws + P.off_buf1};
  const uint64_t nb = 1ull << L;
  const uint64_t nwords = wtg_levels_words(n, L);

  // clear histogram, scan/look-back states and counters (one contiguous range)
  WTG_TRY(wtg_fill_bytes_zero(ctx, ws + P.off_hist, P.off_tab - P.off_hist));
  WTG_TRY(wtg_fill_bytes_zero(ctx, d_words, nwords * 8));

  WTG_TRY(launch_hist<T>(ctx, d_sym, n, hist, (uint32_t)nb));
  if (d_distinct) {
    WTG_TRY(wtg_fill_u64(ctx, d_distinct, 1, 0));
    WTG_LAUNCH(ctx, "distinct", k_count_nonzero, dim3((unsigned)std::min<uint64_t>(wtg_ceil_div(nb, 256), 1024)),
               dim3(256), 0, hist, nb, d_distinct);
    WTG_LAUNCH_CHECK();
  }
  WTG_LAUNCH(ctx, "scan_hist", k_scan_u32<0>, dim3((unsigned)P.scan_tiles), dim3(256), 0, hist, C, nb,
             scanst, ctr + 0, ctx->d_flag);
  WTG_LAUNCH_CHECK();
  if (L > 1) {
    NodeTabArgs na{};
    na.C = C; na.tab = tab; na.states = ntst; na.counter = ctr + 1; na.err = ctx->d_flag;
    na.L = L;
    for (int l = 0; l < 25; l++) na.tile_base[l] = P.nt_tile_base[l];
    WTG_LAUNCH(ctx, "nodetab", k_nodetab, dim3(P.nt_tiles), dim3(256), 0, na);
    WTG_LAUNCH_CHECK();
  }
  if (d_nodes && node_slots) {
    WTG_LAUNCH(ctx, "nodes", k_nodes, dim3((unsigned)wtg_ceil_div(node_slots, 256)), dim3(256), 0,
               C, L, n, d_nodes, node_slots);
    WTG_LAUNCH_CHECK();
  }
  if (n == 0) return WTG_OK;
  const bool vec = ((uintptr_t)d_sym & 15) == 0;
  const void* src = d_sym;
  for (uint32_t l = 0; l < L; l++) {
    LevelArgs la{};
    la.src = src;
    la.dst = buf[l & 1];
    la.words = d_words;
    la.bit_base = (uint64_t)l * n;
    la.n = (uint32_t)n;
    la.bitpos = L - 1 - l;
    la.node_shift = L - l;
    la.nnodes = 1u << l;
    la.tab = tab + ((1u << l) - 1);
    la.states = states + (size_t)l * (P.tiles + 1);
    la.counter = ctr + 2 + l;
    la.err = ctx->d_flag;
    const bool scatter = (l + 1 < L);
    WTG_TRY(launch_level<T>(ctx, la, P.tiles, scatter, l == 0 ? vec : true));
    src = buf[l & 1];
  }
  return WTG_OK;
}
}  // namespace

extern "C" uint64_t wtg_levels_words(uint64_t n, uint32_t levels) {
  return (n * (uint64_t)levels + 63) / 64;
}

static int levels_check(wtg_ctx* ctx, int sym_bytes, uint64_t n, uint32_t L) {
  WTG_REQUIRE(ctx, "null ctx");
  WTG_REQUIRE(sym_bytes == 1 || sym_bytes == 2 || sym_bytes == 4 || sym_bytes == 8,
              "sym_bytes must be 1, 2, 4 or 8");
  WTG_REQUIRE(L >= 1, "levels must be >= 1");
  if (L > 24) {
    wtg_set_error("levels > 24 not supported (dense per-level node tables)");
    return WTG_ERR_UNSUPPORTED;
  }
  if (n >= (1ull << 32)) {
    wtg_set_error("n >= 2^32 not supported");
    return WTG_ERR_UNSUPPORTED;
  }
  return WTG_OK;
}

int wtg_levels_build_async(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n,
                           uint32_t levels, uint64_t* d_words, wtg_node* d_nodes,
                           uint64_t node_slots, unsigned long long* d_distinct) {
  WTG_TRY(levels_check(ctx, sym_bytes, n, levels));
  WTG_REQUIRE(n == 0 || (d_sym && d_words), "null device buffer");
  WTG_TRY(wtg_set_device(ctx));
  switch (sym_bytes) {
    case 1: return levels_build_t<uint8_t>(ctx, (const uint8_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
    case 2: return levels_build_t<uint16_t>(ctx, (const uint16_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
    case 4: return levels_build_t<uint32_t>(ctx, (const uint32_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
    default: return levels_build_t<uint64_t>(ctx, (const uint64_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
  }
}

extern "C" int wtg_levels_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n,
                                uint32_t levels, uint64_t* d_words, wtg_node* d_nodes,
                                uint64_t node_slots) {
  WTG_TRY(wtg_levels_build_async(ctx, d_sym, sym_bytes, n, levels, d_words, d_nodes, node_slots,
                                 nullptr));
  return wtg_finish(ctx);
}

// ------------------------------------------------------------------ remap
template <typename T>
__global__ void k_max(const T* __restrict__ s, uint64_t n, unsigned long long* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t m = 0;
  for (; i < n; i += stride) { uint64_t x = s[i]; m = x > m ? x : m; }
  for (int d = 1; d < 64; d <<= 1) { uint64_t t = __shfl_xor(m, d); m = t > m ? t : m; }
  if (lane_id() == 0) atomicMax(out, (unsigned long long)m);
}
template <typename T>
__global__ void k_flag(const T* __restrict__ s, uint64_t n, uint32_t* __restrict__ A) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) A[s[i]] = 1u;
}
template <typename T>
__global__ void k_apply_map(T* __restrict__ s, uint64_t n, const uint32_t* __restrict__ A) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) s[i] = (T)A[s[i]];
}

template <typename T>
static int max_t(wtg_ctx* ctx, const T* s, uint64_t n, uint64_t* out) {
  if (n == 0) { *out = 0; return WTG_OK; }
  unsigned grid = (unsigned)std::min<uint64_t>(wtg_ceil_div(n, 256), 4096);
  unsigned long long* mx = (unsigned long long*)((char*)ctx->d_flag + 64);
  WTG_TRY(wtg_fill_u64(ctx, mx, 1, 0));
  WTG_LAUNCH(ctx, "max", k_max<T>, dim3(grid), dim3(256), 0, s, n, mx);
  WTG_LAUNCH_CHECK();
  unsigned long long k = 0;
  WTG_CHECK_HIP(hipMemcpyAsync(&k, mx, 8, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = k;
  return WTG_OK;
}

int wtg_max_value(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint64_t* out) {
  switch (sym_bytes) {
    case 1: return max_t<uint8_t>(ctx, (const uint8_t*)d_sym, n, out);
    case 2: return max_t<uint16_t>(ctx, (const uint16_t*)d_sym, n, out);
    case 4: return max_t<uint32_t>(ctx, (const uint32_t*)d_sym, n, out);
    default: return max_t<uint64_t>(ctx, (const uint64_t*)d_sym, n, out);
  }
}

template <typename T>
static int remap_t(wtg_ctx* ctx, T* s, uint64_t n, uint64_t* sigma) {
  if (n == 0) { *sigma = 0; return WTG_OK; }
  unsigned grid = (unsigned)std::min<uint64_t>(wtg_ceil_div(n, 256), 4096);
  uint64_t k = 0;
  WTG_TRY(max_t<T>(ctx, s, n, &k));
  k += 1;  // 1 + max
  if (k + 1 >= (1ull << 31)) {
    wtg_set_error("remap: max symbol too large for a dense flag array");
    return WTG_ERR_UNSUPPORTED;
  }
  const uint64_t na = k + 1;
  size_t o_A = 0, o_B = wtg_align256(na * 4), o_st = o_B + wtg_align256((na + 1) * 4);
  size_t o_ctr = o_st + wtg_align256((wtg_ceil_div(na, 4096) + 1) * 8);
  WTG_TRY(wtg_ws_reserve(ctx, o_ctr + 256));
  char* ws = (char*)ctx->ws;
  uint32_t* A = (uint32_t*)(ws + o_A);
  uint32_t* B = (uint32_t*)(ws + o_B);
  WTG_TRY(wtg_fill_bytes_zero(ctx, ws, o_ctr + 256));
  WTG_LAUNCH(ctx, "remap_flag", k_flag<T>, dim3(grid), dim3(256), 0, s, n, A);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "remap_scan", k_scan_u32<0>, dim3((unsigned)wtg_ceil_div(na, 4096)), dim3(256), 0, A, B,
             na, (uint64_t*)(ws + o_st), (uint32_t*)(ws + o_ctr), ctx->d_flag);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "remap_apply", k_apply_map<T>, dim3(grid), dim3(256), 0, s, n, B);
  WTG_LAUNCH_CHECK();
  uint32_t sg = 0;
  WTG_CHECK_HIP(hipMemcpyAsync(&sg, B + na, 4, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *sigma = sg;
  return WTG_OK;
}

extern "C" int wtg_remap(wtg_ctx* ctx, void* d_sym, int sym_bytes, uint64_t n, uint64_t* sigma) {
  WTG_REQUIRE(ctx && sigma && (n == 0 || d_sym), "wtg_remap: null argument");
  WTG_REQUIRE(sym_bytes == 1 || sym_bytes == 2 || sym_bytes == 4 || sym_bytes == 8, "bad sym_bytes");
  WTG_TRY(wtg_set_device(ctx));
  int rc;
  switch (sym_bytes) {
    case 1: rc = remap_t<uint8_t>(ctx, (uint8_t*)d_sym, n, sigma); break;
    case 2: rc = remap_t<uint16_t>(ctx, (uint16_t*)d_sym, n, sigma); break;
    case 4: rc = remap_t<uint32_t>(ctx, (uint32_t*)d_sym, n, sigma); break;
    default: rc = remap_t<uint64_t>(ctx, (uint64_t*)d_sym, n, sigma); break;
  }
  WTG_TRY(rc);
  return wtg_finish(ctx);
}

// ------------------------------------------------- host-buffer WT() helper
extern "C" int wtg_wt_host(wtg_ctx* ctx, const void* h_sym, int sym_bytes, uint64_t n,
                           uint64_t sigma, uint64_t* h_words, wtg_node* h_nodes) {
  WTG_REQUIRE(ctx && (n == 0 || (h_sym && h_words)), "wtg_wt_host: null argument");
  WTG_REQUIRE(sigma >= 1, "sigma must be >= 1");
  const uint32_t L = (uint32_t)std::max(1, wtg_log2up_host(sigma));
  WTG_TRY(levels_check(ctx, sym_bytes, n, L));
  WTG_TRY(wtg_set_device(ctx));
  const uint64_t nw = wtg_levels_words(n, L);
  const uint64_t slots = 2 * sigma;
  void* d_sym = nullptr; uint64_t* d_words = nullptr; wtg_node* d_nodes = nullptr;
  int rc = WTG_OK;
  WTG_TRY(wtg_malloc(ctx, &d_sym, n * sym_bytes));
  if ((rc = wtg_malloc(ctx, (void**)&d_words, nw * 8)) != WTG_OK) goto out;
  if (h_nodes && (rc = wtg_malloc(ctx, (void**)&d_nodes, slots * sizeof(wtg_node))) != WTG_OK) goto out;
  if ((rc = wtg_memcpy_h2d(ctx, d_sym, h_sym, n * sym_bytes)) != WTG_OK) goto out;
  if ((rc = wtg_levels_build(ctx, d_sym, sym_bytes, n, L, d_words, d_nodes, h_nodes ? slots : 0)) != WTG_OK) goto out;
  if ((rc = wtg_memcpy_d2h(ctx, h_words, d_words, nw * 8)) != WTG_OK) goto out;
  if (h_nodes) rc = wtg_memcpy_d2h(ctx, h_nodes, d_nodes, slots * sizeof(wtg_node));
out:
  wtg_free(ctx, d_sym);
  wtg_free(ctx, d_words);
  wtg_free(ctx, d_nodes);
  return rc;
}
