// wtg_levels.hip -- (a) level-wise wavelet-tree construction on gfx950.
//
// Reference semantics: wavelet/levelWT/WT.C:43-227 (== serialWT/WT.C:32-103):
// level l holds P_l = the stable sort of the symbols by x >> (L-l); the bit at
// global position l*n+i is (P_l[i] >> (L-1-l)) & 1.
//
// GPU pipeline for one build (all integer, HBM bound).  CNT_j[q] = number of
// symbols whose top j bits equal q, kept for every j in one heap-ordered array.
//   level 0:
//      k_level<bits>  16-byte vector loads; the level's bit words from wave
//                     ballots; the tile's zero count; total zeros -> CNT_1.
//   per level k:
//      k_scan_u32     exclusive prefix of the level's per-tile zero counts.
//      k_nodetab      node starts S(p) and, for split levels, the per-node
//                     constants A(p) = S(p)-Zs(p), B(p) = nz(p)+Zs(p) (nz =
//                     zeros of node p = CNT_{k+1}[2p], Zs = zeros of the earlier
//                     nodes of the level); for k >= 1 it first derives CNT_{k+1}
//                     from CNT_k and the next-bit counts made by level k-1.
//      k_level<split> tile zero ranks from bit-plane ballots + mbcnt + LDS;
//                     the tile's stable 0/1 partition staged in LDS and stored
//                     lane-consecutively to
//                       dest = slot < Z_t ? Z_prev + slot + A(p)
//                                         : O_prev + slot - Z_t + B(p)
//                     in the next ping-pong buffer.  The slots of one group
//                     ((k+1)-prefix) are contiguous and land on contiguous
//                     destinations, so the pass also emits, per group, level
//                     k+1's bit words (assembled in slot order in LDS from
//                     ballots; interior words stored, the two edge words OR-ed),
//                     level k+1's zeros per tile and the group's next-bit ones
//                     (level k+1's table; counters replicated across tiles for
//                     shallow levels).  Level L-1 needs no pass of its own.
//   k_nodes           WTnode records (optional).
// The per-tile zero prefix comes from the bits pass, so the split pass needs no
// cross-tile wait: a single-pass decoupled look-back measured ~4x slower here
// because its device-coherent state loads sit on the critical path.
#include "wtg_internal.hpp"

#define FLAG_LOOKBACK WTG_FLAG_WAIT
#define FLAG_BADSYM WTG_FLAG_BADSYM
#define TAB_CACHE 256     // node-table entries cached in LDS per tile
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

static __host__ __device__ __forceinline__ uint64_t heap_off(uint32_t j) { return (1ull << j) - 1; }
// Counter replicas for the next-bit counts of the 2^j groups of level j: a
// tile adds to replica (tile mod reps).  A skewed input can send nearly every
// tile's contribution to one group, so the replica count follows the tile
// count, not the mean group size: enough replicas that at most 64 tiles add
// into one counter (at most 1024, and at most 2^17 counters per level).  Up to
// C1_SUM_REPS replicas of at most 2^C1_SUM_LOG_NODES groups the node-table pass
// sums them itself (no reduce launch).
#define C1_SUM_REPS 32
#define C1_SUM_LOG_NODES 10
static uint32_t c1_reps(uint32_t j, uint64_t tiles) {
  uint32_t r = 1;
  while (r < 1024 && ((uint64_t)r << 6) < tiles && ((uint64_t)r << (j + 1)) <= (1ull << 17)) r <<= 1;
  return r;
}

template <typename T>
struct VecOf;
template <> struct VecOf<uint8_t> { static constexpr int V = 16; };
template <> struct VecOf<uint16_t> { static constexpr int V = 8; };
template <> struct VecOf<uint32_t> { static constexpr int V = 4; };
template <> struct VecOf<uint64_t> { static constexpr int V = 2; };

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
  uint32_t* cnt;         // heap CNT (levels 0..L)
  const uint32_t* c1;    // level k's next-bit counts (2^k, replicas already summed)
  uint32_t reps;
  uint32_t* start;       // heap START (levels 0..L-1)
  uint2* tab;            // level k table (A,B) or nullptr
  uint64_t* st_s;        // look-back states, scan of CNT_k
  uint64_t* st_z;        // look-back states, scan of nz
  uint32_t* counter;
  uint32_t* err;
  unsigned long long* distinct;  // optional: += #nonzero CNT_{k+1}
  uint32_t k;
  uint32_t derive;       // 1: CNT_{k+1} = [CNT_k - c1, c1]
  // optional: the exclusive scan of zin[0..zn) into zout[0..zn] (the level's
  // zeros per tile), scan tile t (4096 values) done by the block that takes
  // tile t, with its own look-back states; node tiles and scan tiles share the
  // tile counter, so every wait is on a block that started earlier
  const uint32_t* zin;
  uint32_t* zout;
  uint64_t zn;
  uint64_t* zst;
  uint32_t ztiles, ntiles;
};

__global__ __launch_bounds__(256) void k_nodetab(NodeTabArgs a) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_pre[2];
  if (threadIdx.x == 0) s_tile = atomicAdd(a.counter, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  if (a.zin && tile < a.ztiles) {  // block-uniform
    const uint64_t b = (uint64_t)tile * 4096 + threadIdx.x * 16;
    uint32_t v[16], sum = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      v[j] = b + j < a.zn ? a.zin[b + j] : 0u;
      sum += v[j];
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan_u32(sum, s_wave, tot);
    if (threadIdx.x < 64) {
      const uint64_t pre = lb_exclusive(a.zst, tile, tot, a.err);
      if (threadIdx.x == 0) s_pre[0] = pre;
    }
    __syncthreads();
    uint32_t run = (uint32_t)(s_pre[0] + ex);
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (b + j < a.zn) a.zout[b + j] = run;
      run += v[j];
    }
    if (b < a.zn && b + 16 >= a.zn) a.zout[a.zn] = run;
    __syncthreads();  // s_pre and s_wave are reused below
  }
  if (tile >= a.ntiles) return;  // a scan-only block
  const uint64_t nnodes = 1ull << a.k;
  const uint32_t* ck = a.cnt + heap_off(a.k);
  uint32_t* ck1 = a.cnt + heap_off(a.k + 1);
  const uint64_t p0 = (uint64_t)tile * 4096 + (uint64_t)threadIdx.x * 16;
  uint32_t cv[16], nz[16];
  uint32_t sum_c = 0, sum_z = 0, nzc = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const uint64_t p = p0 + j;
    uint32_t c = 0, z = 0;
    if (p < nnodes) {
      c = ck[p];
      if (a.derive) {
        uint32_t one = 0;
#pragma unroll 8
        for (uint32_t r = 0; r < a.reps; r++) one += a.c1[(uint64_t)r * nnodes + p];
        z = c - one;
        ck1[2 * p] = z;
        ck1[2 * p + 1] = one;
        nzc += (z != 0) + (one != 0);
      } else {
        z = ck1[2 * p];
        nzc += (z != 0) + (ck1[2 * p + 1] != 0);
      }
    }
    cv[j] = c; nz[j] = z;
    sum_c += c; sum_z += z;
  }
  uint32_t tot_c, tot_z;
  const uint32_t ex_c = block_excl_scan_u32(sum_c, s_wave, tot_c);
  const uint32_t ex_z = block_excl_scan_u32(sum_z, s_wave, tot_z);
  if (threadIdx.x < 64) {
    const uint64_t ps = lb_exclusive(a.st_s, tile, tot_c, a.err);
    const uint64_t pz = lb_exclusive(a.st_z, tile, tot_z, a.err);
    if (threadIdx.x == 0) { s_pre[0] = ps; s_pre[1] = pz; }
  }
  __syncthreads();
  uint32_t S = (uint32_t)(s_pre[0] + ex_c), Z = (uint32_t)(s_pre[1] + ex_z);
  uint32_t* st = a.start + heap_off(a.k);
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const uint64_t p = p0 + j;
    if (p < nnodes) {
      st[p] = S;
      if (a.tab) a.tab[p] = make_uint2(S - Z, nz[j] + Z);
    }
    S += cv[j];
    Z += nz[j];
  }
  if (a.distinct) {
    uint64_t d = wave_sum((uint64_t)nzc);
    if (lane_id() == 0 && d) atomicAdd(a.distinct, (unsigned long long)d);
  }
}

// ---------------------------------------------------------------- WTnodes
__global__ void k_nodes(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ start,
                        uint32_t L, uint64_t n, wtg_node* __restrict__ nodes, uint64_t slots) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= slots) return;
  wtg_node r;
  r.bitmapPtr = 0; r.length = 0; r.parent = WTG_NONE; r.leftChild = WTG_NONE; r.rightChild = WTG_NONE;
  if (v < heap_off(L)) {
    const uint32_t l = 63 - __clzll(v + 1);
    const uint64_t c = cnt[v];
    if (c > 0 || v == 0) {
      r.bitmapPtr = (uint64_t)l * n + start[v];
      r.length = c;
      r.parent = v ? (v - 1) / 2 : WTG_NONE;
      if (l + 1 < L) {
        if (cnt[2 * v + 1] > 0) r.leftChild = 2 * v + 1;
        if (cnt[2 * v + 2] > 0) r.rightChild = 2 * v + 2;
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

// wave-exclusive prefix and wave total of a per-lane count <= V (bit planes)
template <int NPL>
__device__ __forceinline__ uint32_t plane_scan(uint32_t c, uint32_t& tot) {
  uint32_t p = 0, t = 0;
#pragma unroll
  for (int b = 0; b < NPL; b++) {
    const uint64_t B = wave_ballot((c >> b) & 1u);
    p += mbcnt64(B) << b;
    t += popc64(B) << b;
  }
  tot = t;
  return p;
}

template <typename T>
__device__ __forceinline__ T shfl_t(T v, int src) {
  if constexpr (sizeof(T) == 8) return (T)__shfl((unsigned long long)v, src);
  else return (T)__shfl((unsigned int)v, src);
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
  uint32_t check_shift;  // symbols must satisfy x >> check_shift == 0 (64: no check)
  const uint2* tab;      // this level's (A,B) table
  uint32_t* c1;          // next-bit counts of the 2^(l+1) groups of level l+1 (replicated)
  uint32_t reps;         // replicas of c1 (power of two)
  uint32_t* zc;          // bits pass: zeros per tile (out)
  uint32_t* zc_next;     // split pass: zeros per tile of level l+1 (atomically accumulated)
  uint64_t next_base;    // split pass: (l+1)*n, first bit of level l+1
  const uint32_t* zpre;  // split pass: exclusive prefix of zc (in)
  uint32_t* err;
  uint32_t tile0;        // tile of block 0
  uint32_t store;        // split pass: 0 = the next level's keys are not needed (last split)
};

#define LV_STEPS 4
#define LV_STEPS_U8 2
#define LV_STEPS_U16 2
// vector steps per thread per tile (a tile is 256 * steps * V keys): byte keys
// take 8192-key tiles, twice the blocks of 16384 (2^24 byte keys: split 38 ->
// 30 us per level; one step per tile, 4096 keys, is slower again); 16-bit keys
// 4096-key tiles (the 4096-sequence batch: 15.9 -> 15.5 ms)
template <typename T> struct StepsOf { static constexpr int S = LV_STEPS; };
template <> struct StepsOf<uint8_t> { static constexpr int S = LV_STEPS_U8; };
template <> struct StepsOf<uint16_t> { static constexpr int S = LV_STEPS_U16; };

// wave-uniform copies of a value (readfirstlane) and of one lane's value
template <typename U>
__device__ __forceinline__ U rfl(U v) {
  if constexpr (sizeof(U) == 8) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(uint64_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (U)(((uint64_t)hi << 32) | lo);
  } else {
    return (U)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  }
}
__device__ __forceinline__ uint32_t rl32(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ __forceinline__ uint64_t rl64(uint64_t v, int l) {
  return ((uint64_t)rl32((uint32_t)(v >> 32), l) << 32) | rl32((uint32_t)v, l);
}
// Groups of a split tile, one thread per group (tiles holding many small
// groups): the same outputs as the wave-per-group path below.  s_hpre[i] /
// s_npre[i] = group starts / next-level ones in slot-order words [0, i).
#define SPLIT_GROUPS_WAVE 32
template <typename T, bool COUNT, int TILE>
__device__ __forceinline__ void split_groups_many(const LevelArgs& a, uint32_t tile, uint32_t valid, uint32_t zt,
                                                  uint64_t zprev, uint64_t oprev, bool cached, uint64_t pfirst,
                                                  const T* s_stage, const uint64_t* s_nb, const uint64_t* s_head,
                                                  const uint2* s_tab, const uint32_t* s_hpre,
                                                  const uint32_t* s_npre, uint32_t ng) {
  constexpr int NW = TILE / 64;
  uint32_t* c1 = COUNT ? a.c1 + (uint64_t)(tile & (a.reps - 1)) * (2ull * a.nnodes) : nullptr;
  auto start_of = [&](uint32_t gi) -> uint32_t {  // slot of the gi-th group start
    uint32_t lo = 0, hi = NW;                      // last word with s_hpre <= gi
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (s_hpre[mid] <= gi) lo = mid; else hi = mid;
    }
    return lo * 64 + select_in_word(s_head[lo], gi - s_hpre[lo]);
  };
  auto ones_before = [&](uint32_t x) -> uint32_t {  // next-level ones in slots [0, x)
    const uint32_t wi = x >> 6;
    if (wi >= (uint32_t)NW) return s_npre[NW];
    return s_npre[wi] + popc64(s_nb[wi] & ((1ull << (x & 63)) - 1));
  };
  bool bad = false;
  for (uint32_t gi = threadIdx.x; gi < ng; gi += blockDim.x) {
    const uint32_t sG = start_of(gi);
    const uint32_t eG = gi + 1 < ng ? start_of(gi + 1) : valid;
    const uint32_t len = eG - sG;
    const uint64_t x0 = (uint64_t)s_stage[sG];
    const uint64_t pnode = x0 >> a.node_shift;
    const uint64_t g = x0 >> a.bitpos;
    uint2 ab;
    if (cached && pnode - pfirst < TAB_CACHE) ab = s_tab[pnode - pfirst];
    else if (pnode < a.nnodes) ab = a.tab[pnode];
    else { bad = true; continue; }
    const uint64_t dG = (sG < zt) ? zprev + sG + ab.x : oprev + (sG - zt) + ab.y;
    if (dG + len > a.n || len == 0) { bad = true; continue; }
    const uint64_t b0 = (dG / TILE + 1) * TILE;  // first destination-tile boundary
    const uint32_t rel0 = (dG + len > b0) ? (uint32_t)(b0 - dG) : len;
    const uint64_t GB0 = a.next_base + dG;
    const uint64_t k0 = GB0 >> 6, k1 = (GB0 + len - 1) >> 6;
    const int32_t off0 = (int32_t)(k0 * 64 - GB0);  // in (-64, 0]
    for (uint64_t k = k0; k <= k1; k++) {
      const int32_t off = off0 + (int32_t)((k - k0) * 64);
      const int32_t P = (int32_t)sG + off;
      uint64_t val;
      if (P >= 0) {
        const uint32_t w0 = (uint32_t)P >> 6, sh = (uint32_t)P & 63;
        val = s_nb[w0] >> sh;
        if (sh) val |= s_nb[w0 + 1] << (64 - sh);
      } else {
        val = s_nb[0] << (uint32_t)(-P);
      }
      const int32_t lo = off < 0 ? -off : 0;
      const int32_t hi = ((int32_t)len - off < 64) ? (int32_t)len - off : 64;
      const uint64_t msk = (hi >= 64 ? ~0ull : ((1ull << hi) - 1)) & (~0ull << lo);
      val &= msk;
      if (lo == 0 && hi == 64) a.words[k] = val;
      else if (val) atomicOr((unsigned long long*)&a.words[k], (unsigned long long)val);
    }
    const uint32_t os = ones_before(sG);
    const uint32_t ones = ones_before(eG) - os;
    const uint32_t o_lo = ones_before(sG + rel0) - os;
    const uint32_t o_hi = ones - o_lo;
    if (COUNT && ones && g < 2ull * a.nnodes) atomicAdd(&c1[g], ones);
    const uint64_t D0 = dG / TILE;
    const uint32_t z_lo = rel0 - o_lo;
    if (z_lo) atomicAdd(&a.zc_next[D0], z_lo);
    if (len > rel0) {
      const uint32_t z_hi = (len - rel0) - o_hi;
      if (z_hi) atomicAdd(&a.zc_next[D0 + 1], z_hi);
    }
  }
  if (bad) atomicOr(a.err, FLAG_BADSYM);
}

// Group phase of a split tile: per group of the next level (contiguous slots
// [sG, eG) -> contiguous destinations [dG, dG+len)), one wave per group, one
// output word per lane: level l+1's bit words (interior words stored, the two
// edge words OR-ed), the group's next-bit ones (level l+1's table) and its
// zeros per destination tile of level l+1.  Group starts and the ones counts
// come from per-lane prefixes of the slot-order head / next-bit words read
// back as wave-uniform values, so the per-group bookkeeping runs on the
// scalar unit.
template <typename T, bool COUNT, int TILE>
__device__ __forceinline__ void split_groups(const LevelArgs& a, uint32_t tile, uint32_t valid, uint32_t zt,
                                             uint64_t zprev, uint64_t oprev, bool cached, uint64_t pfirst,
                                             const T* s_stage, const uint64_t* s_nb, const uint64_t* s_head,
                                             const uint2* s_tab) {
  constexpr int NW = TILE / 64;
  constexpr int HPL = (NW + 63) / 64;  // head / next-bit words per lane
  const int w = threadIdx.x >> 6, lane = lane_id();
  uint32_t* c1 = COUNT ? a.c1 + (uint64_t)(tile & (a.reps - 1)) * (2ull * a.nnodes) : nullptr;
  uint64_t hwl[HPL], nwl[HPL];
  uint32_t hc = 0, oc = 0;
#pragma unroll
  for (int j = 0; j < HPL; j++) {
    const int idx = lane * HPL + j;
    hwl[j] = (idx < NW) ? s_head[idx] : 0ull;
    nwl[j] = (idx < NW) ? s_nb[idx] : 0ull;
    hc += popc64(hwl[j]);
    oc += popc64(nwl[j]);
  }
  const uint32_t hincl = wave_incl_scan(hc), hex = hincl - hc;
  const uint32_t oincl = wave_incl_scan(oc), oex = oincl - oc;
  const uint32_t ng = rl32(hincl, 63);
  const uint32_t otot = rl32(oincl, 63);
  if (ng > SPLIT_GROUPS_WAVE) {
    // many small groups (sparse or skewed inputs): one thread per group, with
    // the head / next-bit word prefixes shared through LDS (ng is the same in
    // every wave: all read the same slot-order arrays)
    __shared__ uint32_t s_hpre[NW + 1], s_npre[NW + 1];
    if (w == 0) {
      uint32_t hh = hex, oo = oex;
#pragma unroll
      for (int j = 0; j < HPL; j++) {
        const int idx = lane * HPL + j;
        if (idx < NW) {
          s_hpre[idx] = hh;
          s_npre[idx] = oo;
        }
        hh += popc64(hwl[j]);
        oo += popc64(nwl[j]);
      }
      if (lane == 0) {
        s_hpre[NW] = ng;
        s_npre[NW] = otot;
      }
    }
    __syncthreads();
    split_groups_many<T, COUNT, TILE>(a, tile, valid, zt, zprev, oprev, cached, pfirst, s_stage, s_nb, s_head, s_tab,
                                      s_hpre, s_npre, ng);
    return;
  }
  // next-level ones among slots [0, x) (x uniform, <= valid)
  auto ones_before = [&](uint32_t x) -> uint32_t {
    const uint32_t wi = x >> 6;
    if (wi >= (uint32_t)NW) return otot;
    const int ln = (int)(wi / HPL);
    const uint32_t jj = wi % HPL;
    uint32_t r = rl32(oex, ln);
#pragma unroll
    for (int j = 0; j < HPL; j++) {
      const uint64_t wv = rl64(nwl[j], ln);
      if ((uint32_t)j < jj) r += popc64(wv);
      else if ((uint32_t)j == jj) r += popc64(wv & ((1ull << (x & 63)) - 1));
    }
    return r;
  };
  // slot of the gi-th group start (gi uniform, < ng)
  auto group_start = [&](uint32_t gi) -> uint32_t {
    const uint64_t own = wave_ballot(hex <= gi && gi < hincl);
    const int ol = (int)__builtin_ctzll(own);
    uint32_t r = gi - rl32(hex, ol);
    uint32_t slot = 0;
#pragma unroll
    for (int j = 0; j < HPL; j++) {
      const uint64_t hv = rl64(hwl[j], ol);
      const uint32_t c = popc64(hv);
      if (r < c) { slot = (uint32_t)(ol * HPL + j) * 64 + select_in_word(hv, r); break; }
      r -= c;
    }
    return slot;
  };
  for (uint32_t gi = (uint32_t)w; gi < ng; gi += 4) {
    const uint32_t sG = group_start(gi);
    const uint32_t eG = (gi + 1 < ng) ? group_start(gi + 1) : valid;
    const uint32_t len = eG - sG;
    const uint64_t x0 = (uint64_t)rfl<T>(s_stage[sG]);
    const uint64_t pnode = x0 >> a.node_shift;
    const uint64_t g = x0 >> a.bitpos;
    uint2 ab;
    if (cached && pnode - pfirst < TAB_CACHE) {
      const uint2 t = s_tab[pnode - pfirst];
      ab = make_uint2(rfl<uint32_t>(t.x), rfl<uint32_t>(t.y));
    } else if (pnode < a.nnodes) {
      ab = a.tab[pnode];
    } else {
      if (lane == 0) atomicOr(a.err, FLAG_BADSYM);
      continue;
    }
    const uint64_t dG = (sG < zt) ? zprev + sG + ab.x : oprev + (sG - zt) + ab.y;
    if (dG + len > a.n || len == 0) { if (lane == 0) atomicOr(a.err, FLAG_BADSYM); continue; }
    const uint64_t b0 = (dG / TILE + 1) * TILE;  // first destination-tile boundary
    const uint32_t rel0 = (dG + len > b0) ? (uint32_t)(b0 - dG) : len;
    const uint64_t GB0 = a.next_base + dG;
    const uint64_t k0 = GB0 >> 6, k1 = (GB0 + len - 1) >> 6;
    const int32_t off0 = (int32_t)(k0 * 64 - GB0);  // in (-64, 0]
    const uint32_t nwo = (uint32_t)(k1 - k0 + 1);
    for (uint32_t t = (uint32_t)lane; t < nwo; t += 64) {
      const int32_t off = off0 + (int32_t)(t * 64);  // group bit at word bit 0
      const int32_t P = (int32_t)sG + off;           // its slot
      uint64_t val;
      if (P >= 0) {
        const uint32_t w0 = (uint32_t)P >> 6, sh = (uint32_t)P & 63;
        val = s_nb[w0] >> sh;
        if (sh) val |= s_nb[w0 + 1] << (64 - sh);
      } else {
        val = s_nb[0] << (uint32_t)(-P);
      }
      const int32_t lo = off < 0 ? -off : 0;
      const int32_t hi = ((int32_t)len - off < 64) ? (int32_t)len - off : 64;
      const uint64_t msk = (hi >= 64 ? ~0ull : ((1ull << hi) - 1)) & (~0ull << lo);
      val &= msk;
      const uint64_t k = k0 + t;
      if (lo == 0 && hi == 64) a.words[k] = val;
      else if (val) atomicOr((unsigned long long*)&a.words[k], (unsigned long long)val);
    }
    if (lane == 0) {
      const uint32_t os = ones_before(sG);
      const uint32_t ones = ones_before(eG) - os;
      const uint32_t o_lo = ones_before(sG + rel0) - os;
      const uint32_t o_hi = ones - o_lo;
      if (COUNT && ones && g < 2ull * a.nnodes) atomicAdd(&c1[g], ones);
      const uint64_t D0 = dG / TILE;
      const uint32_t z_lo = rel0 - o_lo;
      if (z_lo) atomicAdd(&a.zc_next[D0], z_lo);
      if (len > rel0) {
        const uint32_t z_hi = (len - rel0) - o_hi;
        if (z_hi) atomicAdd(&a.zc_next[D0 + 1], z_hi);
      }
    }
  }
}
// MODE 0 ("bits"): the level's bit words + the tile's zero count.
// MODE 1 ("split"): the stable scatter of the tile into the next level, with
//                   the tile's zero prefix taken from the scanned counts.
template <typename T, int MODE, bool VEC, bool COUNT>
__global__ __launch_bounds__(256) void k_level(LevelArgs a) {
  constexpr bool SCATTER = MODE == 1;
  constexpr int V = VecOf<T>::V;
  constexpr int G = 64 / V;              // lanes per bit word
  constexpr int TILE = 256 * StepsOf<T>::S * V;
  constexpr int NW = TILE / 64;          // words per tile
  constexpr int ROUNDS = TILE / 256;     // store rounds (one element per lane)
  constexpr int NPL = (V == 16) ? 5 : (V == 8) ? 4 : (V == 4) ? 3 : 2;  // bit planes of a count <= V
  constexpr int NCH = StepsOf<T>::S * 4;      // (step, wave) chunks
  __shared__ uint64_t s_words[SCATTER ? 1 : NW + 1];
  __shared__ uint32_t s_wz[NCH];
  __shared__ uint64_t s_nb[SCATTER ? NW + 1 : 1];
  __shared__ uint64_t s_head[SCATTER ? NW : 1];
  __shared__ T s_edge[2][SCATTER ? NCH : 1];
  __shared__ T s_stage[SCATTER ? TILE : 1];
  __shared__ uint2 s_tab[SCATTER ? TAB_CACHE : 1];

  const int tid = threadIdx.x, w = tid >> 6, lane = lane_id();
  const uint32_t tile = a.tile0 + blockIdx.x;
  const uint64_t tile_start = (uint64_t)tile * TILE;
  const uint64_t n = a.n;
  const uint64_t valid = (n - tile_start < (uint64_t)TILE) ? n - tile_start : (uint64_t)TILE;
  const bool full = valid == (uint64_t)TILE;
  const T* src = (const T*)a.src;
  uint64_t zprev = 0;
  if (SCATTER) zprev = a.zpre[tile];

  T x[StepsOf<T>::S][V];
#pragma unroll
  for (int s = 0; s < StepsOf<T>::S; s++) {
    const uint64_t e0 = tile_start + (uint64_t)(s * 256 + tid) * V;
    if (VEC && full) {
      const v4u32 q = __builtin_nontemporal_load((const v4u32*)(src + e0));
      const T* qp = (const T*)&q;
#pragma unroll
      for (int v = 0; v < V; v++) x[s][v] = qp[v];
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) x[s][v] = (e0 + v < n) ? src[e0 + v] : (T)0;
    }
  }

  uint32_t mk[StepsOf<T>::S], pre[StepsOf<T>::S];
  bool bad = false;
#pragma unroll
  for (int s = 0; s < StepsOf<T>::S; s++) {
    const uint64_t e0 = tile_start + (uint64_t)(s * 256 + tid) * V;
    uint32_t m = 0, vm = 0;
#pragma unroll
    for (int v = 0; v < V; v++) {
      m |= (uint32_t)(((uint64_t)x[s][v] >> a.bitpos) & 1u) << v;
      vm |= (uint32_t)(e0 + v < n) << v;
      if (!SCATTER && a.check_shift < 64) bad |= (((uint64_t)x[s][v]) >> a.check_shift) != 0;
    }
    m &= vm;
    // invalid elements are marked as ones: they never count as zeros and sit
    // after every valid element of the tile, so they are never staged
    mk[s] = m | ~vm;
    if (!SCATTER) {
      // bit words of this (step, wave): lane j < V assembles word j from V ballots
      uint64_t word = 0;
      const int j = lane & (V - 1);
#pragma unroll
      for (int v = 0; v < V; v++) {
        const uint64_t B = wave_ballot((m >> v) & 1u);
        word |= spread_bits<V>(B >> (j * G)) << v;
      }
      if (lane < V) s_words[(s * 4 + w) * V + lane] = word;
    }
    uint32_t tot;
    pre[s] = plane_scan<NPL>((uint32_t)__popc(~m & vm), tot);
    if (lane == 0) s_wz[s * 4 + w] = tot;
    if (SCATTER) {
      if (lane == 0) s_edge[0][s * 4 + w] = x[s][0];
      if (lane == 63) s_edge[1][s * 4 + w] = x[s][V - 1];
    }
  }
  if (bad) atomicOr(a.err, FLAG_BADSYM);
  __syncthreads();

  uint32_t woff[StepsOf<T>::S];
  uint32_t zt = 0;
#pragma unroll
  for (int k = 0; k < NCH; k++) {
#pragma unroll
    for (int s = 0; s < StepsOf<T>::S; s++)
      if (k == s * 4 + w) woff[s] = zt;
    zt += s_wz[k];
  }

  if constexpr (!SCATTER) {
    if (tid == 0) a.zc[tile] = zt;
    // bit words out (funnel shift when the level base is not 64-aligned)
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
  } else {
    // ---- node-table entries of the tile's node range into LDS
    const uint64_t pfirst = (uint64_t)s_edge[0][0] >> a.node_shift;
    const uint64_t plast = (uint64_t)s_edge[1][NCH - 1] >> a.node_shift;
    const bool cached = full && plast >= pfirst && plast - pfirst < TAB_CACHE && plast < a.nnodes;
    if (cached && (uint64_t)tid <= plast - pfirst) s_tab[tid] = a.tab[pfirst + tid];
    // ---- stable 0/1 partition of the tile into LDS
#pragma unroll
    for (int s = 0; s < StepsOf<T>::S; s++) {
      const uint32_t m = mk[s];
      const uint32_t zb = woff[s] + pre[s];
      const uint32_t j0 = (uint32_t)(s * 256 + tid) * V;
#pragma unroll
      for (int v = 0; v < V; v++) {
        const uint32_t zl = zb + (uint32_t)__popc(~m & ((1u << v) - 1u));
        const uint32_t j = j0 + v;
        if (j < valid) {
          const uint32_t slot = ((m >> v) & 1u) ? zt + (j - zl) : zl;
          s_stage[slot] = x[s][v];
        }
      }
    }
    __syncthreads();
    const uint64_t oprev = tile_start - zprev;
    T* dst = (T*)a.dst;
    const uint32_t nb = a.bitpos - 1;   // the next level's bit
    // ---- slot-order bit arrays of the next level's bit (s_nb) and of group
    // starts (s_head, group = (l+1)-prefix); LDS only, so the barrier below
    // does not wait on any global store
    for (int r = 0; r < ROUNDS; r++) {
      const uint32_t q = (uint32_t)(r * 256 + tid);
      const bool in = q < valid;
      const T xv = in ? s_stage[q] : (T)0;
      const uint64_t g = (uint64_t)xv >> a.bitpos;
      T pv = shfl_t<T>(xv, (lane + 63) & 63);
      if (lane == 0 && q > 0 && in) pv = s_stage[q - 1];
      const bool head = in && (q == 0 || ((uint64_t)pv >> a.bitpos) != g);
      const uint64_t H = wave_ballot(head);
      const uint64_t C = wave_ballot(in && ((((uint64_t)xv) >> nb) & 1u));
      if (lane == 0) { s_nb[q >> 6] = C; s_head[q >> 6] = H; }
    }
    if (tid == 0) s_nb[NW] = 0;
    __syncthreads();
    // ---- lane-consecutive stores of the staged tile (unrolled by 16: the
    // rounds' LDS reads and stores in flight together)
#pragma unroll 16
    for (int r = 0; r < ROUNDS; r++) {
      const uint32_t q = (uint32_t)(r * 256 + tid);
      if (q < valid && a.store) {
        const T xv = s_stage[q];
        const uint64_t pnode = (uint64_t)xv >> a.node_shift;
        uint2 ab;
        bool ok = true;
        if (cached && pnode - pfirst < TAB_CACHE) ab = s_tab[pnode - pfirst];
        else if (pnode < a.nnodes) ab = a.tab[pnode];
        else ok = false;
        const uint64_t dest = (q < zt) ? zprev + q + ab.x : oprev + (q - zt) + ab.y;
        if (ok && dest < n) __builtin_nontemporal_store(xv, dst + dest);
        else atomicOr(a.err, FLAG_BADSYM);
      }
    }
    split_groups<T, COUNT, TILE>(a, tile, (uint32_t)valid, zt, zprev, oprev, cached, pfirst, s_stage, s_nb,
                                 s_head, s_tab);
  }
}

// Raw buffer store of one key at element index i of a buffer described as
// `nbytes` long: the hardware drops stores at or past the end (no branch).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t key_rsrc(void* base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)nbytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ void key_store(__amdgpu_buffer_rsrc_t r, uint32_t i, T v) {
  constexpr int NT = 2;  // non-temporal
  if constexpr (sizeof(T) == 1) __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, r, (int)i, 0, NT);
  else if constexpr (sizeof(T) == 2) __builtin_amdgcn_raw_buffer_store_b16((uint16_t)v, r, (int)(i * 2), 0, NT);
  else if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, r, (int)(i * 4), 0, NT);
  else {
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 d = {(uint32_t)(uint64_t)v, (uint32_t)((uint64_t)v >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(d, r, (int)(i * 8), 0, NT);
  }
}

// Split pass over FULL tiles (every tile but a partial last one), in 32-bit
// index arithmetic (n < 2^32) and without per-element validity tests: the
// same stable scatter, next-level bit words and counts as k_level<T, 1>.
// Needs node_shift < 8*sizeof(T) for 32- and 64-bit keys (the host sends level
// 0 of 32-level uint32 keys to k_level); narrower keys shift as int.  Chunk zero prefixes by a 16-lane DPP scan; slot-order
// heads and next-level bits from two LDS reads per element; stores from the
// node table cached in LDS (a tile's nodes are consecutive).
template <typename T, bool VEC, bool COUNT>
__global__ __launch_bounds__(256) void k_split(LevelArgs a) {
  constexpr int V = VecOf<T>::V;
  constexpr int TILE = 256 * StepsOf<T>::S * V;
  constexpr int NW = TILE / 64;
  constexpr int ROUNDS = TILE / 256;
  constexpr int NPL = (V == 16) ? 5 : (V == 8) ? 4 : (V == 4) ? 3 : 2;
  constexpr int NCH = StepsOf<T>::S * 4;  // (step, wave) chunks, chunk s*4+w
  static_assert(NCH <= 16, "chunk scan covers one DPP row");
  __shared__ uint32_t s_wz[NCH];
  __shared__ uint64_t s_nb[NW + 1];
  __shared__ uint64_t s_head[NW];
  __shared__ T s_stage[TILE];
  __shared__ uint2 s_tab[TAB_CACHE];
  __shared__ T s_ends[2];

  const int tid = threadIdx.x, w = tid >> 6, lane = lane_id();
  const uint32_t tile = blockIdx.x;
  const uint32_t ts = tile * (uint32_t)TILE;
  const T* src = (const T*)a.src + ts;
  const uint32_t zprev = a.zpre[tile];

  T x[StepsOf<T>::S][V];
#pragma unroll
  for (int s = 0; s < StepsOf<T>::S; s++) {
    const uint32_t e0 = (uint32_t)(s * 256 + tid) * V;
    if (VEC) {
      const v4u32 q = __builtin_nontemporal_load((const v4u32*)(src + e0));
      const T* qp = (const T*)&q;
#pragma unroll
      for (int v = 0; v < V; v++) x[s][v] = qp[v];
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) x[s][v] = src[e0 + v];
    }
  }
  // this level's bits, the lane's zero rank inside its chunk, chunk zeros
  uint32_t mk[StepsOf<T>::S], pre[StepsOf<T>::S];
#pragma unroll
  for (int s = 0; s < StepsOf<T>::S; s++) {
    uint32_t m = 0;
#pragma unroll
    for (int v = 0; v < V; v++) m |= (uint32_t)((x[s][v] >> a.bitpos) & 1u) << v;
    mk[s] = m;
    uint32_t tot;
    pre[s] = plane_scan<NPL>((uint32_t)V - (uint32_t)__popc(m), tot);
    if (lane == 0) s_wz[s * 4 + w] = tot;
  }
  if (tid == 0) s_ends[0] = x[0][0];
  if (tid == 255) s_ends[1] = x[StepsOf<T>::S - 1][V - 1];
  __syncthreads();

  // chunk prefixes: lane c < NCH holds chunk c's zeros; inclusive scan in row 0
  const uint32_t cz = lane < NCH ? s_wz[lane] : 0u;
  uint32_t ci = cz;
  ci += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)ci, 0x111, 0xF, 0xF, false);  // row_shr:1
  ci += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)ci, 0x112, 0xF, 0xF, false);  // row_shr:2
  ci += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)ci, 0x114, 0xF, 0xF, false);  // row_shr:4
  ci += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)ci, 0x118, 0xF, 0xF, false);  // row_shr:8
  const uint32_t cx = ci - cz;
  const uint32_t zt = (uint32_t)__builtin_amdgcn_readlane((int)ci, NCH - 1);
  const int wu = __builtin_amdgcn_readfirstlane(w);
  uint32_t woff[StepsOf<T>::S];
#pragma unroll
  for (int s = 0; s < StepsOf<T>::S; s++) woff[s] = (uint32_t)__builtin_amdgcn_readlane((int)cx, s * 4 + wu);
  if (tid == 0) a.zc[tile] = zt;  // unused by the split; keeps zc like k_level's

  // node-table entries of the tile's node range into LDS
  const uint64_t pfirst = (uint64_t)s_ends[0] >> a.node_shift;
  const uint64_t plast = (uint64_t)s_ends[1] >> a.node_shift;
  const bool cached = plast >= pfirst && plast - pfirst < TAB_CACHE && plast < a.nnodes;
  if (cached && (uint64_t)tid <= plast - pfirst) s_tab[tid] = a.tab[pfirst + tid];

  // stable 0/1 partition of the tile into LDS
#pragma unroll
  for (int s = 0; s < StepsOf<T>::S; s++) {
    const uint32_t m = mk[s];
    const uint32_t zb = woff[s] + pre[s];                      // slot of the lane's first zero
    const uint32_t ob = zt + (uint32_t)(s * 256 + tid) * V - zb;  // slot of its first one
#pragma unroll
    for (int v = 0; v < V; v++) {
      const uint32_t ov = (uint32_t)__popc(m & ((1u << v) - 1u));  // ones before v
      const uint32_t slot = ((m >> v) & 1u) ? ob + ov : zb + (uint32_t)v - ov;
      s_stage[slot] = x[s][v];
    }
  }
  __syncthreads();

  // slot-order bit arrays: next level's bit (s_nb) and group starts (s_head,
  // group = (l+1)-prefix)
  const uint32_t nb = a.bitpos - 1;
  const T hlim = (T)((T)1 << a.bitpos);
#pragma unroll 4
  for (int r = 0; r < ROUNDS; r++) {
    const uint32_t q = (uint32_t)(r * 256 + tid);
    const T xv = s_stage[q];
    const T pv = s_stage[q ? q - 1 : 0];
    const bool head = q == 0 || (T)(xv ^ pv) >= hlim;
    const uint64_t H = wave_ballot(head);
    const uint64_t C = wave_ballot((xv >> nb) & 1u);
    if (lane == 0) {
      s_nb[q >> 6] = C;
      s_head[q >> 6] = H;
    }
  }
  if (tid == 0) s_nb[NW] = 0;
  __syncthreads();

  // lane-consecutive stores of the staged tile
  const uint32_t oprev = ts - zprev;
  T* dst = (T*)a.dst;
  bool bad = false;
  if (!a.store) {
    // last split: only the next level's bits and counts are needed
  } else if (cached && (uint64_t)a.n * sizeof(T) < (1ull << 31)) {
    // destinations past n (only from invalid keys, which the bits pass flags)
    // are dropped by the buffer range check
    const __amdgpu_buffer_rsrc_t rs = key_rsrc(dst, a.n * (uint32_t)sizeof(T));
    const uint32_t pf = (uint32_t)pfirst;
    const uint32_t b0 = zprev + (uint32_t)tid, b1 = oprev - zt + (uint32_t)tid;
#pragma unroll 16
    for (int r = 0; r < ROUNDS; r++) {
      const uint32_t q = (uint32_t)(r * 256 + tid);
      const T xv = s_stage[q];
      uint32_t pi = (uint32_t)(xv >> a.node_shift) - pf;
      pi = pi < TAB_CACHE - 1 ? pi : TAB_CACHE - 1;
      const uint2 ab = s_tab[pi];
      const uint32_t dest = q < zt ? b0 + (uint32_t)(r * 256) + ab.x : b1 + (uint32_t)(r * 256) + ab.y;
      key_store<T>(rs, dest, xv);
    }
  } else if (cached) {
    const uint32_t pf = (uint32_t)pfirst;
    const uint32_t b0 = zprev + (uint32_t)tid, b1 = oprev - zt + (uint32_t)tid;
#pragma unroll 16
    for (int r = 0; r < ROUNDS; r++) {
      const uint32_t q = (uint32_t)(r * 256 + tid);
      const T xv = s_stage[q];
      uint32_t pi = (uint32_t)(xv >> a.node_shift) - pf;
      pi = pi < TAB_CACHE - 1 ? pi : TAB_CACHE - 1;
      const uint2 ab = s_tab[pi];
      const uint32_t dest = q < zt ? b0 + (uint32_t)(r * 256) + ab.x : b1 + (uint32_t)(r * 256) + ab.y;
      if (dest < a.n) __builtin_nontemporal_store(xv, dst + dest);
      else bad = true;
    }
  } else {
#pragma unroll 4
    for (int r = 0; r < ROUNDS; r++) {
      const uint32_t q = (uint32_t)(r * 256 + tid);
      const T xv = s_stage[q];
      const uint64_t pnode = (uint64_t)xv >> a.node_shift;
      if (pnode < a.nnodes) {
        const uint2 ab = a.tab[pnode];
        const uint32_t dest = q < zt ? zprev + q + ab.x : oprev + (q - zt) + ab.y;
        if (dest < a.n) __builtin_nontemporal_store(xv, dst + dest);
        else bad = true;
      } else {
        bad = true;
      }
    }
  }
  if (bad) atomicOr(a.err, FLAG_BADSYM);
  split_groups<T, COUNT, TILE>(a, tile, (uint32_t)TILE, zt, zprev, oprev, cached, pfirst, s_stage, s_nb, s_head,
                               s_tab);
}

// adds replicas 1.. of a level's next-bit counters into replica 0:
// c1[g] += sum_{r>=1} c1[r*ng + g]; thread per group (coalesced rows), block
// row y covers replicas [1 + y*C1_RCH, 1 + (y+1)*C1_RCH)
#define C1_RCH 32
__global__ __launch_bounds__(256) void k_c1_reduce(uint32_t* __restrict__ c1, uint32_t reps, uint64_t ngroups) {
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ngroups) return;
  const uint32_t r0 = 1 + blockIdx.y * C1_RCH;
  const uint32_t r1 = r0 + C1_RCH < reps ? r0 + C1_RCH : reps;
  uint32_t acc = 0;
  for (uint32_t r = r0; r < r1; r++) acc += c1[(uint64_t)r * ngroups + g];
  if (acc) atomicAdd(&c1[g], acc);
}

__global__ void k_cnt_from_zeros(uint32_t* __restrict__ cnt, const uint32_t* zeros, uint64_t n) {
  if (threadIdx.x == 0) {
    cnt[0] = (uint32_t)n;
    cnt[1] = *zeros;
    cnt[2] = (uint32_t)(n - *zeros);
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
  uint64_t scan_tiles;  // tiles of the scan over the per-tile zero counts
  size_t off_cnt, off_c1, off_ntst, off_scst, off_ctr, off_start, off_tab, off_zc, off_zpre,
      off_buf0, off_buf1, total;
  uint64_t c1off[26];      // c1 region of level j (next-bit counts of its 2^j groups)
  uint32_t reps[26];       // replicas of level j's counters
  uint64_t nt_tiles[26];   // node-table tiles of level k
  uint64_t nt_base[26];    // state offset of level k (two arrays: s and z)
};

int tile_elems_for(int w) { return 256 * (w == 1 ? LV_STEPS_U8 : w == 2 ? LV_STEPS_U16 : LV_STEPS) * (16 / w); }

LevelPlan plan_levels(uint64_t n, uint32_t L, int w) {
  LevelPlan P{};
  P.L = L; P.n = n; P.w = w;
  P.tile_elems = tile_elems_for(w);
  P.tiles = wtg_ceil_div(n, P.tile_elems);
  P.scan_tiles = wtg_ceil_div(P.tiles, 4096);
  uint64_t t = 0, c = 0;
  for (uint32_t k = 0; k <= L; k++) {
    P.c1off[k] = c;
    P.reps[k] = c1_reps(k, P.tiles);
    c += (uint64_t)P.reps[k] << k;
    if (k < L) {
      P.nt_tiles[k] = wtg_ceil_div(1ull << k, 4096);
      P.nt_base[k] = t;
      t += 2 * (P.nt_tiles[k] + 1);
    }
  }
  size_t o = 0;
  // [off_cnt, off_start) is cleared at the start of a build
  P.off_cnt = o; o = wtg_align256(o + ((2ull << L) + 2) * 4);
  P.off_c1 = o; o = wtg_align256(o + (c + 1) * 4);
  P.off_ntst = o; o = wtg_align256(o + (t + 1) * 8);
  P.off_scst = o; o = wtg_align256(o + (size_t)L * (P.scan_tiles + 1) * 8);
  P.off_ctr = o; o = wtg_align256(o + 128 * 4);
  P.off_zc = o; o = wtg_align256(o + (size_t)L * (P.tiles + 1) * 4);
  P.off_start = o; o = wtg_align256(o + ((1ull << L) + 1) * 4);
  P.off_tab = o; o = wtg_align256(o + ((1ull << (L - 1)) + 1) * 8);
  P.off_zpre = o; o = wtg_align256(o + (P.tiles + 1) * 4);
  P.off_buf0 = o; o = wtg_align256(o + n * w);
  P.off_buf1 = o; o = wtg_align256(o + n * w);
  P.total = o;
  return P;
}

template <typename T, int MODE, bool VEC>
void pick_level(bool count, void (*&k)(LevelArgs)) {
  if (count && MODE == 1) k = k_level<T, MODE, VEC, MODE == 1>;
  else k = k_level<T, MODE, VEC, false>;
}

template <typename T, bool VEC>
void pick_split(bool count, void (*&k)(LevelArgs)) {
  if (count) k = k_split<T, VEC, true>;
  else k = k_split<T, VEC, false>;
}

template <typename T>
int launch_level(wtg_ctx* ctx, const LevelArgs& la, uint64_t tiles, int mode, bool vec, bool count) {
  constexpr uint64_t TILE = 256ull * StepsOf<T>::S * VecOf<T>::V;
  void (*k)(LevelArgs) = nullptr;
  if (mode == 1) {
    // full tiles: k_split; a partial last tile (and, for 32- and 64-bit keys, a
    // node shift as wide as the key): k_level.  8- and 16-bit keys are promoted
    // to int before the shift, so their full-width shift is defined.
    const uint64_t full = (sizeof(T) < 4 || la.node_shift < 8 * sizeof(T)) ? (uint64_t)la.n / TILE : 0;
    if (full) {
      if (vec) pick_split<T, true>(count, k); else pick_split<T, false>(count, k);
      // the last split (no key stores) is timed under its own name
      WTG_LAUNCH(ctx, la.store ? "level_split" : "level_split_last", k, dim3((unsigned)full), dim3(256), 0, la);
      WTG_LAUNCH_CHECK();
    }
    if (full == tiles) return WTG_OK;
    if (vec) pick_level<T, 1, true>(count, k); else pick_level<T, 1, false>(count, k);
    LevelArgs lt = la;
    lt.tile0 = (uint32_t)full;
    WTG_LAUNCH(ctx, "level_split_tail", k, dim3((unsigned)(tiles - full)), dim3(256), 0, lt);
  } else {
    if (vec) pick_level<T, 0, true>(false, k); else pick_level<T, 0, false>(false, k);
    WTG_LAUNCH(ctx, "level_bits", k, dim3((unsigned)tiles), dim3(256), 0, la);
  }
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}

// ones of the level-`start` bit per level-`start` node (the c1 a split would
// have counted); wave-aggregated when a wave's 16*64 keys share one node
template <typename T>
__global__ __launch_bounds__(256) void k_prefix_ones(const T* __restrict__ keys, uint64_t n, uint32_t bitpos,
                                                     uint32_t node_shift, uint32_t* __restrict__ c1) {
  const uint64_t base = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 16;
  uint64_t first = WTG_NONE, last = 0;
  uint32_t ones = 0;
  uint64_t cur = WTG_NONE;
  bool mixed = false;
  for (int k = 0; k < 16; k++) {
    if (base + k >= n) break;
    const uint64_t x = (uint64_t)keys[base + k];
    const uint64_t node = x >> node_shift;
    if (cur != WTG_NONE && node != cur) mixed = true;
    if (first == WTG_NONE) first = node;
    last = node;
    cur = node;
    ones += (uint32_t)((x >> bitpos) & 1u);
  }
  const uint64_t f0 = __shfl((unsigned long long)first, 0);
  const bool same = !mixed && (first == WTG_NONE || (first == f0 && last == f0));
  if (__all(same)) {
    const uint32_t t = wave_sum(ones);
    if (lane_id() == 0 && t && f0 != WTG_NONE) atomicAdd(&c1[f0], t);
    return;
  }
  for (int k = 0; k < 16; k++) {
    if (base + k >= n) break;
    const uint64_t x = (uint64_t)keys[base + k];
    if ((x >> bitpos) & 1u) atomicAdd(&c1[x >> node_shift], 1u);
  }
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Level pipeline over levels [first, L) of keys < 2^L.  first > 0: the keys
// are already grouped by their top `first` bits (node of level `first`), whose
// counts are in d_cnt_start[2^first]; bits of level k go to bit (k-first)*n.
template <typename T>
int levels_build_t(wtg_ctx* ctx, const T* d_sym, uint64_t n, uint32_t L, uint64_t* d_words,
                   wtg_node* d_nodes, uint64_t node_slots, unsigned long long* d_distinct,
                   uint32_t first = 0, const uint32_t* d_cnt_start = nullptr,
                   const uint32_t* d_ones_start = nullptr) {
  const int w = sizeof(T);
  LevelPlan P = plan_levels(n, L, w);
  WTG_TRY(wtg_ws_reserve(ctx, P.total));
  char* ws = (char*)ctx->ws;
  uint32_t* cnt = (uint32_t*)(ws + P.off_cnt);
  uint32_t* c1 = (uint32_t*)(ws + P.off_c1);
  uint32_t* start = (uint32_t*)(ws + P.off_start);
  uint2* tab = (uint2*)(ws + P.off_tab);
  uint64_t* ntst = (uint64_t*)(ws + P.off_ntst);
  uint64_t* scst = (uint64_t*)(ws + P.off_scst);
  uint32_t* ctr = (uint32_t*)(ws + P.off_ctr);
  uint32_t* zc = (uint32_t*)(ws + P.off_zc);
  uint32_t* zpre = (uint32_t*)(ws + P.off_zpre);
  void* buf[2] = {ws + P.off_buf0, ws + P.off_buf1};
  const uint64_t nwords = wtg_levels_words(n, L - first);
  const bool want_nodes = d_nodes && node_slots && first == 0;

  // clear counts, scan states and counters (one contiguous range)
  WTG_TRY(wtg_fill_bytes_zero(ctx, ws + P.off_cnt, P.off_start - P.off_cnt));
  WTG_TRY(wtg_fill_bytes_zero(ctx, d_words, nwords * 8));
  if (d_distinct) WTG_TRY(wtg_fill_u64(ctx, d_distinct, 1, 0));

  const bool vec = ((uintptr_t)d_sym & 15) == 0;
  const void* src = d_sym;
  for (uint32_t k = first; k < L; k++) {
    const bool split = k + 1 < L;
    uint32_t* zck = zc + (size_t)k * (P.tiles + 1);
    LevelArgs la{};
    la.src = src;
    la.dst = buf[k & 1];
    la.store = k + 2 < L ? 1u : 0u;  // the last split's keys are never read
    la.words = d_words;
    la.bit_base = (uint64_t)(k - first) * n;
    la.n = (uint32_t)n;
    la.bitpos = L - 1 - k;
    la.node_shift = L - k;
    la.nnodes = 1u << k;
    la.check_shift = (k == first && L < 64) ? L : 64;
    la.tab = tab + heap_off(k);
    la.c1 = c1 + P.c1off[k + 1];
    la.reps = P.reps[k + 1];
    la.zc = zck;
    la.zc_next = split ? zc + (size_t)(k + 1) * (P.tiles + 1) : nullptr;
    la.next_base = (uint64_t)(k + 1 - first) * n;
    la.zpre = zpre;
    la.err = ctx->d_flag;
    const bool kvec = k == first ? vec : true;
    if (n > 0) {
      // first level: bits pass (its bit words + zeros per tile); every later
      // level's bits and tile zeros were produced by the previous split
      if (k == first) WTG_TRY(launch_level<T>(ctx, la, P.tiles, 0, kvec, false));
      // (after a split the scan is folded into the node-table pass)
      if (k == first) {
        WTG_LAUNCH(ctx, "scan_tiles", k_scan_u32<uint32_t>, dim3((unsigned)P.scan_tiles), dim3(256), 0, zck, zpre,
                   P.tiles, scst + (size_t)k * (P.scan_tiles + 1), ctr + k, ctx->d_flag);
        WTG_LAUNCH_CHECK();
      }
      if (k == 0) {
        // CNT_0 = n, CNT_1 = (zeros, ones) of the top bit
        WTG_LAUNCH(ctx, "cnt_init", k_cnt_from_zeros, dim3(1), dim3(64), 0, cnt, zpre + P.tiles, n);
        WTG_LAUNCH_CHECK();
      } else if (k == first) {
        // CNT_start from the caller, and the level's ones per node
        const uint64_t nn = 1ull << first;
        WTG_LAUNCH(ctx, "cnt_start", k_copy_u32, dim3((unsigned)wtg_ceil_div(nn, 256)), dim3(256), 0,
                   d_cnt_start, cnt + heap_off(first), nn);
        WTG_LAUNCH_CHECK();
        if (d_ones_start) {
          WTG_LAUNCH(ctx, "ones_start", k_copy_u32, dim3((unsigned)wtg_ceil_div(nn, 256)), dim3(256), 0,
                     d_ones_start, c1 + P.c1off[k], nn);
        } else {
          WTG_LAUNCH(ctx, "prefix_ones", k_prefix_ones<T>, dim3((unsigned)wtg_ceil_div(n, 4096)), dim3(256), 0,
                     (const T*)src, n, L - 1 - k, L - k, c1 + P.c1off[k]);
        }
        WTG_LAUNCH_CHECK();
      }
    }
    if (k == 0 && first == 0 && d_distinct && L == 1) {
      WTG_LAUNCH(ctx, "distinct", k_count_nonzero, dim3(1), dim3(256), 0, cnt + heap_off(1), (uint64_t)2,
                 d_distinct);
      WTG_LAUNCH_CHECK();
    }
    // node table of level k (node starts always: cheap, needed by WTnodes)
    // (the node-table pass reads a node's replicas strided: only for few nodes)
    const bool reduce = k >= 1 && k > first && n > 0 && P.reps[k] > 1 &&
                        (P.reps[k] > C1_SUM_REPS || k > C1_SUM_LOG_NODES);
    if (reduce) {
      WTG_LAUNCH(ctx, "c1_reduce", k_c1_reduce,
                 dim3((unsigned)wtg_ceil_div(1ull << k, 256), (unsigned)wtg_ceil_div(P.reps[k] - 1, C1_RCH)), dim3(256),
                 0, c1 + P.c1off[k], P.reps[k], (uint64_t)(1ull << k));
      WTG_LAUNCH_CHECK();
    }
    {
      NodeTabArgs na{};
      na.cnt = cnt;
      na.c1 = c1 + P.c1off[k];
      // replicas hold counts only when the previous split wrote them (k > first)
      na.reps = (reduce || k <= first) ? 1u : P.reps[k];
      na.start = start;
      na.tab = split ? tab + heap_off(k) : nullptr;
      na.st_s = ntst + P.nt_base[k];
      na.st_z = ntst + P.nt_base[k] + P.nt_tiles[k] + 1;
      na.counter = ctr + 64 + k;
      na.err = ctx->d_flag;
      na.k = k;
      na.derive = k >= 1 ? 1u : 0u;
      na.distinct = (d_distinct && k + 1 == L && L > 1) ? d_distinct : nullptr;
      na.ntiles = (uint32_t)P.nt_tiles[k];
      if (split && n > 0 && k > first) {
        na.zin = zck;
        na.zout = zpre;
        na.zn = P.tiles;
        na.zst = scst + (size_t)k * (P.scan_tiles + 1);
        na.ztiles = (uint32_t)P.scan_tiles;
      }
      if (split || want_nodes || na.distinct) {
        WTG_LAUNCH(ctx, "nodetab", k_nodetab, dim3((unsigned)std::max<uint64_t>(P.nt_tiles[k], na.ztiles)),
                   dim3(256), 0, na);
        WTG_LAUNCH_CHECK();
      }
    }
    if (n == 0 || !split) continue;
    // split pass: stable scatter into the next level's order, plus the next
    // level's bit words, its zeros per tile and the next-bit counts that feed
    // level k+1's table (and, at the last split, the distinct-symbol count)
    const bool count = (k + 2 <= L - 1) || (d_distinct && k + 2 == L);
    WTG_TRY(launch_level<T>(ctx, la, P.tiles, 1, kvec, count));
    src = buf[k & 1];
  }
  if (want_nodes) {
    WTG_LAUNCH(ctx, "nodes", k_nodes, dim3((unsigned)wtg_ceil_div(node_slots, 256)), dim3(256), 0,
               cnt, start, L, n, d_nodes, node_slots);
    WTG_LAUNCH_CHECK();
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
  if (L > 64) {
    wtg_set_error("levels > 64 not supported");
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
  if (levels > 24) {  // node ids by rank in the level instead of dense prefix tables
    if (d_nodes && node_slots) {
      wtg_set_error("levels > 24: WTnode records are limited to 24 levels (pass no node array)");
      return WTG_ERR_UNSUPPORTED;
    }
    return n ? wtg_deep_build(ctx, d_sym, sym_bytes, n, levels, d_words, d_distinct) : WTG_OK;
  }
  if (n > 0 && ctx->level_method == 0 && wtg_mld_enabled())
    return wtg_mld_build(ctx, d_sym, sym_bytes, n, levels, d_words, d_nodes, node_slots, d_distinct);
  switch (sym_bytes) {
    case 1: return levels_build_t<uint8_t>(ctx, (const uint8_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
    case 2: return levels_build_t<uint16_t>(ctx, (const uint16_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
    case 4: return levels_build_t<uint32_t>(ctx, (const uint32_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
    default: return levels_build_t<uint64_t>(ctx, (const uint64_t*)d_sym, n, levels, d_words, d_nodes, node_slots, d_distinct);
  }
}

int wtg_levels_build_from(wtg_ctx* ctx, const void* d_keys, int key_bytes, uint64_t n, uint32_t levels,
                          uint32_t start, const uint32_t* d_cnt_start, uint64_t* d_words,
                          const uint32_t* d_ones_start) {
  WTG_TRY(levels_check(ctx, key_bytes, n, levels));
  WTG_REQUIRE(start < levels && (start == 0 || d_cnt_start), "levels_build_from: bad start level");
  WTG_REQUIRE(n == 0 || (d_keys && d_words), "null device buffer");
  switch (key_bytes) {
    case 1: return levels_build_t<uint8_t>(ctx, (const uint8_t*)d_keys, n, levels, d_words, nullptr, 0, nullptr, start, d_cnt_start, d_ones_start);
    case 2: return levels_build_t<uint16_t>(ctx, (const uint16_t*)d_keys, n, levels, d_words, nullptr, 0, nullptr, start, d_cnt_start, d_ones_start);
    case 4: return levels_build_t<uint32_t>(ctx, (const uint32_t*)d_keys, n, levels, d_words, nullptr, 0, nullptr, start, d_cnt_start, d_ones_start);
    default: return levels_build_t<uint64_t>(ctx, (const uint64_t*)d_keys, n, levels, d_words, nullptr, 0, nullptr, start, d_cnt_start, d_ones_start);
  }
}

size_t wtg_levels_ws_bytes(uint64_t n, uint32_t levels, int key_bytes) {
  return plan_levels(n, levels, key_bytes).total;
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
  WTG_LAUNCH(ctx, "remap_scan", k_scan_u32<uint32_t>, dim3((unsigned)wtg_ceil_div(na, 4096)), dim3(256), 0, A, B,
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
