// wtg_mld.hip -- multi-level direct construction of the level-wise wavelet
// tree: up to 8 levels per pass over the keys.
//
// Reference semantics (wavelet/levelWT/WT.C:43-227 == serialWT/WT.C:32-103):
// level l holds P_l = the stable sort of the symbols by their top l bits
// (x >> (L-l)); the bit at global position l*n + i is bit L-1-l of P_l[i].
//
// Levels are built in groups [s, s+g), g <= 8.  The keys of a group are in
// P_s order (the input order for s = 0), so the level-s nodes are contiguous
// segments.  For a key x at position i of P_s, with d = the g bits after its
// top s bits (x >> (L-s-g)) & (2^g-1), and for each level l = s + j:
//     pos_l(i) = NodeStart_l(x >> (L-l)) + #{i' < i in P_s : same top l bits}
// because P_l is the stable sort of P_s by the top l bits.  NodeStart_l(p) =
// #keys with smaller top l bits = Cum[p << (s+g-l)], Cum being the exclusive
// prefix of the histogram of the top s+g bits.  The count of equal-prefix keys
// before i splits into the tiles of i's segment before i's tile (the "carry",
// kept per d, a range of d for coarser levels) and the rank inside the tile.
//
// Tiles are segment-aligned (each covers <= MLD_T keys of one level-s node),
// so every key of a tile shares the top s bits and a tile's ranks are those of
// its own local subtree over the next g levels.  Per group:
//   k_mld_count   per tile, the key count of each leaf d (wave ballots find
//                 the lanes with equal d, the group leader adds its size);
//   k_mld_scan1/2 column-wise exclusive prefix of those counts over the tiles
//                 (carry(t, d) = P[t][d] - P[first tile of t's segment][d]);
//   k_mld_hist    the histogram of the top s+g bits from that prefix, then
//                 k_scan_u32 -> Cum (the node starts of all g levels);
//   k_mld_emit    per tile, a level pipeline in LDS: at level s+j the tile's
//                 keys are stably sorted by their top j bits of d, so the
//                 wave ballots of the level bit over 64-key chunks are the
//                 tile's level bits in node order; a stable split of each node
//                 by that bit (chunk zero prefix + per-node table) gives the
//                 next order.  The staged bits go out as runs, one per node
//                 and level (interior words stored, edge words OR-ed), and the
//                 keys, now in leaf order, to their P_{s+g} positions (the
//                 next group's input);
//   k_mld_ntiles  the next group's tiles per level-(s+g) node.
// Key traffic per group: read twice (count, emit), written once (scatter).
#include "wtg_internal.hpp"
#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

constexpr uint32_t MLD_T = 2048;            // keys per tile
constexpr uint32_t MLD_GMAX = 8;            // levels per group
constexpr uint32_t MLD_BW = MLD_T / 32 + 2; // staged bit words per level (+2 read-ahead pad)
constexpr uint32_t MLD_CPW = MLD_T / 64 / 4; // 64-key chunks per wave
static_assert(MLD_T / 64 <= 64, "the chunk prefix is one wave scan");
constexpr uint32_t MLD_SR = 64;             // tiles per block of the carry scan

struct MldArgs {
  const void* keys;          // P_s order; group 0: the input symbols (key type K)
  uint64_t n;
  uint32_t L, s, g;
  uint32_t ntiles0;          // s == 0: ceil(n / MLD_T)
  const uint32_t* tstart;    // s > 0: 2^s + 1 entries, first tile of each level-s node
  const uint32_t* tseg;      // s > 0: the level-s node of each tile
  const uint32_t* cum_s;     // s > 0: 2^s + 1 entries, keys before each level-s node
  const uint32_t* tcnt;      // [tile][2^g] leaf counts (written by k_mld_count)
  uint32_t* tcnt_w;
  const uint32_t* ploc;      // [tile][2^g] block-local exclusive prefix over tiles
  const uint32_t* bpre;      // [block][2^g] exclusive prefix of the block sums
  const uint32_t* cum;       // 2^(s+g) + 1 entries
  uint64_t* words;
  uint32_t lvl0;             // the first level of the build: level l's bits go to (l - lvl0) * n
  uint32_t check;            // 1: flag input keys with bits at or above chk_shift
  uint32_t chk_shift;
  const uint64_t* src_off;   // null, or (first group) where each level-s node's keys start in `keys`
  uint32_t kshift;           // first group: the keys are the input values >> kshift (the top L
                             // levels of a deeper tree); 0 otherwise
  uint32_t* kout;            // P_{s+g} keys as uint32 (emit with SCATTER)
  uint32_t* err;
};

__device__ __forceinline__ uint32_t mld_ntiles(const MldArgs& a) {
  return a.s == 0 ? a.ntiles0 : a.tstart[1u << a.s];
}

// tile t -> (segment, first key in P_s, keys, first tile of the segment)
__device__ __forceinline__ bool mld_tile(const MldArgs& a, uint32_t t, uint32_t& seg, uint64_t& base,
                                         uint32_t& len, uint32_t& first) {
  if (a.s == 0) {
    const uint64_t off = (uint64_t)t * MLD_T;
    if (t >= a.ntiles0 || off >= a.n) return false;
    seg = 0;
    base = off;
    len = (uint32_t)(a.n - off < MLD_T ? a.n - off : MLD_T);
    first = 0;
    return true;
  }
  const uint32_t nseg = 1u << a.s;
  if (t >= a.tstart[nseg]) return false;
  const uint32_t lo = a.tseg[t];  // the segment of tile t (k_mld_tseg)
  seg = lo;
  first = a.tstart[lo];
  const uint64_t off = (uint64_t)(t - first) * MLD_T;
  const uint64_t cnt = a.cum_s[lo + 1] - a.cum_s[lo];
  if (off >= cnt) return false;  // (cannot happen for consistent tables)
  base = a.cum_s[lo] + off;
  len = (uint32_t)(cnt - off < MLD_T ? cnt - off : MLD_T);
  return true;
}

// where tile keys start in a.keys: their P_s position, or, when the level-s
// nodes' keys lie apart (src_off), that node's source offset + the tile's
// offset inside the node
__device__ __forceinline__ uint64_t mld_src(const MldArgs& a, uint32_t seg, uint64_t base) {
  return a.src_off ? a.src_off[seg] + (base - a.cum_s[seg]) : base;
}

template <typename K>
__device__ __forceinline__ uint64_t mld_key(const void* keys, uint64_t i) {
  return (uint64_t)reinterpret_cast<const K*>(keys)[i];
}

// a 16-byte word as a native vector (loads through any address space)
typedef uint32_t mld_u4 __attribute__((ext_vector_type(4)));

// element e of a 16-byte word of K-sized keys (little endian)
template <typename K>
__device__ __forceinline__ uint64_t word_elem(const mld_u4& q, uint32_t e) {
  const uint32_t c[4] = {q.x, q.y, q.z, q.w};
  if constexpr (sizeof(K) == 8) {
    return (uint64_t)c[2 * e] | ((uint64_t)c[2 * e + 1] << 32);
  } else {
    constexpr uint32_t B = 8 * sizeof(K), PER = 32 / B;
    const uint32_t dw = c[e / PER];
    return B == 32 ? dw : (dw >> (B * (e % PER))) & ((1u << (B & 31)) - 1u);
  }
}

// the g-bit leaf of element e: bits [tsh, tsh + g) of the key (gw = bits of
// that field inside the key's width, for keys narrower than a dword)
template <typename K>
__device__ __forceinline__ uint32_t word_leaf(const mld_u4& q, uint32_t e, uint32_t tsh, uint32_t gw, uint32_t gmask) {
  if constexpr (sizeof(K) == 8) {
    return (uint32_t)(word_elem<K>(q, e) >> tsh) & gmask;
  } else {
    constexpr uint32_t B = 8 * sizeof(K), PER = 32 / B;
    const uint32_t c[4] = {q.x, q.y, q.z, q.w};
    return __builtin_amdgcn_ubfe(c[e / PER], B * (e % PER) + tsh, gw);
  }
}

// leaf counts of the tiles, one wave per MLD_CNT_TPW<K> consecutive tiles (no
// block barrier; byte and 16-bit keys take several tiles per wave, so each
// wave reads 8 KiB).  A tile's keys are read as the aligned 16-byte words
// that cover them, the loads of all of the wave's tiles in flight at once;
// counting needs no key order, so every element of every word is counted (LDS
// atomics; a slot whose 64 elements share one leaf adds once) and the
// elements of the two edge words that lie outside the tile are taken back
// afterwards.  The aligned words never leave the pages of the tile's own bytes.
// Trees of at most SG leaves per tile (SG > 0) count by ballots instead: per
// slot and leaf one ballot of the lanes holding it, the counts in scalar
// registers (a few leaves would otherwise put every lane's atomic on one of a
// few LDS words).
// tiles per wave of the count kernel: one for every key width, so small
// trees keep enough waves in flight (2^24 byte keys: 23 vs 31 us with four
// tiles per wave; the batched 16-bit keys: 239 vs 236 Gsym/s with two)
template <typename K>
constexpr uint32_t mld_cnt_tpw() { return 1u; }

__device__ __forceinline__ void wave_sync_lds_mld() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// a device address as a global-memory pointer (global, not flat, loads)
typedef const __attribute__((address_space(1))) mld_u4* mld_gword;

template <typename K, uint32_t SG>
__global__ __launch_bounds__(256) void k_mld_count(MldArgs a) {
  constexpr uint32_t KW = sizeof(K), E = 16 / KW, TPW = mld_cnt_tpw<K>();
  constexpr uint32_t NW = (MLD_T * KW / 16 + 1 + 63) / 64;  // words per lane (a tile covers <= T*KW/16 + 1)
  __shared__ uint32_t s_cnt_all[4][TPW][1u << MLD_GMAX];
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t G = 1u << a.g, lane = (uint32_t)lane_id();
  uintptr_t b0[TPW], b1[TPW], w0[TPW];
  uint64_t pbase[TPW];
  uint32_t nw[TPW], tt[TPW], plen[TPW];
  bool ok[TPW];
  mld_u4 v[TPW][NW];
#pragma unroll
  for (uint32_t u = 0; u < TPW; u++) {
    uint32_t seg, len, first;
    uint64_t base;
    tt[u] = (blockIdx.x * 4 + w) * TPW + u;
    ok[u] = mld_tile(a, tt[u], seg, base, len, first);  // wave-uniform
    b0[u] = b1[u] = w0[u] = 0;
    nw[u] = 0;
    pbase[u] = base;
    plen[u] = len;
    if (ok[u]) {
      b0[u] = (uintptr_t)a.keys + mld_src(a, seg, base) * KW;
      b1[u] = b0[u] + (uintptr_t)len * KW;
      w0[u] = b0[u] & ~(uintptr_t)15;
      nw[u] = (uint32_t)((b1[u] - w0[u] + 15) >> 4);
    }
    const mld_gword wp = (mld_gword)w0[u];
#pragma unroll
    for (uint32_t m = 0; m < NW; m++) {
      const uint32_t k = m * 64 + lane;
      v[u][m] = k < nw[u] ? wp[k] : mld_u4{0u, 0u, 0u, 0u};
    }
  }
  if (!ok[0]) return;
  // clear the group's level words: a tile clears, at each of the group's
  // levels, the words whose first bit falls on one of its positions (the
  // positions of a level-s node keep their range at every later level), so
  // every word is cleared once, by the group of its first bit's level, before
  // that group's emit ORs into it
#pragma unroll
  for (uint32_t u = 0; u < TPW; u++) {
    if (!ok[u]) break;
    for (uint32_t j = 0; j < a.g; j++) {
      const uint64_t lb = (uint64_t)(a.s + j - a.lvl0) * a.n;
      const uint64_t wa = (lb + pbase[u] + 63) >> 6, wb = (lb + pbase[u] + plen[u] + 63) >> 6;
      for (uint64_t k = wa + lane; k < wb; k += 64) a.words[k] = 0ull;
    }
  }
  const uint32_t tsh = a.L - a.s - a.g + a.kshift;  // the leaf's first bit in the key
  constexpr uint32_t B = 8 * KW;
  const uint32_t gw = KW == 8 ? a.g : (tsh >= B ? 0u : (B - tsh < a.g ? B - tsh : a.g));
  const uint32_t gmask = G - 1;
  // the bad-key test over whole dwords: bits at or above chk_shift of a key
  uint32_t cm_lo = 0, cm_hi = 0;
  if (a.check) {
    const uint32_t c = a.chk_shift;
    if constexpr (KW == 8) {
      cm_lo = c >= 32 ? 0u : ~((1u << c) - 1u);
      cm_hi = c >= 32 ? ~((1u << (c - 32)) - 1u) : ~0u;
    } else {
      const uint32_t km = B == 32 ? ~0u : (1u << (B & 31)) - 1u;
      const uint32_t hm = c >= B ? 0u : km & ~((1u << c) - 1u);
      cm_lo = cm_hi = hm * (B == 8 ? 0x01010101u : B == 16 ? 0x00010001u : 1u);
    }
  }
#pragma unroll
  for (uint32_t u = 0; u < TPW; u++)
    for (uint32_t i = lane; i < G; i += 64) s_cnt_all[w][u][i] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t acc_lo = 0, acc_hi = 0;  // OR of the interior words (edge words are tested per key below)
  bool bad = false;
#pragma unroll
  for (uint32_t u = 0; u < TPW; u++) {
    if (!ok[u]) break;  // wave-uniform (later tiles are past the end too)
    uint32_t* s_cnt = s_cnt_all[w][u];
    uint32_t sc[SG > 0 ? SG : 1];  // SG > 0: the counts per leaf (wave-uniform)
#pragma unroll
    for (uint32_t d = 0; d < (SG > 0 ? SG : 1); d++) sc[d] = 0;
#pragma unroll
    for (uint32_t m = 0; m < NW; m++) {
      if (m * 64 >= nw[u]) break;  // wave-uniform
      const uint32_t k = m * 64 + lane;
      const bool act = k < nw[u];
      const uint32_t nact = popc64(wave_ballot(act));
      if (act && k > 0 && k + 1 < nw[u]) {
        acc_lo |= v[u][m].x | v[u][m].z;
        acc_hi |= v[u][m].y | v[u][m].w;
      }
#pragma unroll
      for (uint32_t e = 0; e < E; e++) {
        const uint32_t d = word_leaf<K>(v[u][m], e, tsh, gw, gmask);
        if constexpr (SG > 0) {
#pragma unroll
          for (uint32_t q = 0; q < SG; q++) sc[q] += popc64(wave_ballot(act && d == q));
        } else {
          const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);  // lane 0 is active
          if (wave_ballot(act && d != d0) == 0) {
            if (lane == 0) atomicAdd(&s_cnt[d0], nact);
          } else if (act) {
            atomicAdd(&s_cnt[d], 1u);
          }
        }
      }
    }
    if constexpr (SG > 0) {
#pragma unroll
      for (uint32_t q = 0; q < SG; q++)
        if (lane == q && q < G) s_cnt[q] = sc[q];
      wave_sync_lds_mld();
    }
    // edge words: elements outside [b0, b1) were counted above; the others
    // get the per-key bad-key test
    if (lane < 2 && (lane == 0 || nw[u] > 1)) {
      const uint32_t k = lane == 0 ? 0u : nw[u] - 1;
      const mld_u4 q = ((mld_gword)w0[u])[k];
#pragma unroll
      for (uint32_t e = 0; e < E; e++) {
        const uintptr_t ea = w0[u] + 16 * (uintptr_t)k + e * KW;
        if (ea < b0[u] || ea >= b1[u]) {
          atomicSub(&s_cnt[word_leaf<K>(q, e, tsh, gw, gmask)], 1u);
        } else if (a.check && (word_elem<K>(q, e) >> a.chk_shift) != 0) {
          bad = true;
        }
      }
    }
  }
  if (bad || ((acc_lo & cm_lo) | (acc_hi & cm_hi)) != 0) atomicOr(a.err, WTG_FLAG_BADSYM);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (uint32_t u = 0; u < TPW; u++) {
    if (!ok[u]) break;
    for (uint32_t i = lane; i < G; i += 64) a.tcnt_w[(uint64_t)tt[u] * G + i] = s_cnt_all[w][u][i];
  }
}

// column-wise exclusive prefix of the leaf counts over the tiles, blocks of
// MLD_SR tiles: ploc = prefix inside the block, bsum = the block's totals.
// Thread (q, c) takes column c of MLD_SR/Q consecutive rows (Q = 256/G), all
// loaded before the scan.  Rows >= ntiles count 0 (row ntiles then holds the
// grand total).
__global__ __launch_bounds__(256) void k_mld_scan1(const uint32_t* __restrict__ tcnt, uint32_t* __restrict__ ploc,
                                                   uint32_t* __restrict__ bsum, const uint32_t* ntiles_dev,
                                                   uint32_t ntiles_host, uint32_t G, uint32_t nrows) {
  __shared__ uint32_t s_tot[256];
  const uint32_t nt = ntiles_dev ? *ntiles_dev : ntiles_host;
  const uint32_t Q = 256 / G < MLD_SR ? 256 / G : MLD_SR, c = threadIdx.x % G, q = threadIdx.x / G;
  const uint32_t per = q < Q ? MLD_SR / Q : 0u;  // rows of this thread (threads past Q*G idle)
  const uint64_t r0 = (uint64_t)blockIdx.x * MLD_SR + (uint64_t)q * per;
  uint32_t v[MLD_SR];
#pragma unroll
  for (uint32_t r = 0; r < MLD_SR; r++) {
    const uint64_t row = r0 + r;
    v[r] = (r < per && row < nt) ? tcnt[row * G + c] : 0u;
  }
  uint32_t acc = 0;
#pragma unroll
  for (uint32_t r = 0; r < MLD_SR; r++) {
    const uint32_t x = v[r];
    v[r] = acc;
    acc += x;
  }
  s_tot[threadIdx.x] = acc;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (uint32_t k = 0; k < Q; k++) {
    const uint32_t t = s_tot[k * G + c];
    if (k < q) pre += t;
    tot += t;
  }
#pragma unroll
  for (uint32_t r = 0; r < MLD_SR; r++) {
    const uint64_t row = r0 + r;
    if (r < per && row < nrows) ploc[row * G + c] = v[r] + pre;
  }
  if (q == 0) bsum[(uint64_t)blockIdx.x * G + c] = tot;
}

// exclusive prefix of the block sums per column: block c scans column c
__global__ __launch_bounds__(256) void k_mld_scan2(const uint32_t* __restrict__ bsum, uint32_t* __restrict__ bpre,
                                                   uint32_t nblocks, uint32_t G) {
  __shared__ uint32_t s_wave[4];
  const uint32_t c = blockIdx.x;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nblocks; b0 += 1024) {
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t b = b0 + threadIdx.x * 4 + k;
      v[k] = b < nblocks ? bsum[(uint64_t)b * G + c] : 0u;
      sum += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan_u32(sum, s_wave, tot) + carry;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t b = b0 + threadIdx.x * 4 + k;
      if (b < nblocks) bpre[(uint64_t)b * G + c] = ex;
      ex += v[k];
    }
    carry += tot;
  }
}

__device__ __forceinline__ uint32_t mld_prefix(const MldArgs& a, uint32_t row, uint32_t c) {
  const uint32_t G = 1u << a.g;
  return a.ploc[(uint64_t)row * G + c] + a.bpre[(uint64_t)(row / MLD_SR) * G + c];
}

// histogram of the top s+g bits: bin (seg << g) | d = keys of segment seg with
// leaf d = P[last tile + 1][d] - P[first tile][d]
__global__ __launch_bounds__(256) void k_mld_hist(MldArgs a, uint32_t* __restrict__ hist) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nbins = 1ull << (a.s + a.g);
  if (i >= nbins) return;
  const uint32_t seg = (uint32_t)(i >> a.g), d = (uint32_t)(i & ((1u << a.g) - 1));
  const uint32_t t0 = a.s == 0 ? 0u : a.tstart[seg];
  const uint32_t t1 = a.s == 0 ? a.ntiles0 : a.tstart[seg + 1];
  hist[i] = mld_prefix(a, t1, d) - mld_prefix(a, t0, d);
}

// tiles of each level-(s+g) node for the next group
__global__ __launch_bounds__(256) void k_mld_ntiles(const uint32_t* __restrict__ cum, uint64_t nseg,
                                                    uint32_t* __restrict__ nt) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  nt[i] = (cum[i + 1] - cum[i] + MLD_T - 1) / MLD_T;
}

// tile -> segment map, one thread per tile: the last segment k with
// tstart[k] <= t (empty segments share their start with the next one)
__global__ __launch_bounds__(256) void k_mld_tseg(const uint32_t* __restrict__ tstart, uint64_t nseg,
                                                  uint32_t* __restrict__ tseg, uint64_t maxtiles) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= maxtiles || t >= tstart[nseg]) return;
  uint64_t lo = 0, hi = nseg;  // tstart[lo] <= t < tstart[hi]
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (tstart[mid] <= t) lo = mid; else hi = mid;
  }
  tseg[t] = (uint32_t)lo;
}

// inclusive prefix sum over the 64 lanes with DPP moves (no LDS round trip):
// row_shr 1, 2, 4, 8 scan each 16-lane row (lanes shifted in from outside the
// row read 0), then row_bcast 15 adds row 0's total to row 1 and row 2's to
// row 3, and row_bcast 31 adds the total of rows 0-1 to rows 2-3
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return v;
}

// 64 bits of an LDS bit array from bit position p (p + 64 < bits + 64 pad)
__device__ __forceinline__ uint64_t lds_bits64(const uint32_t* row, uint32_t p) {
  const uint32_t k = p >> 5, o = p & 31;
  const uint64_t lo = (uint64_t)row[k] | ((uint64_t)row[k + 1] << 32);
  const uint64_t hi = row[k + 2];
  return o ? (lo >> o) | (hi << (64 - o)) : lo;
}

template <typename K, bool SCATTER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_mld_emit(MldArgs a) {
  // block scans by DPP moves in the bits-only pass, by shuffles in the pass
  // with key output (measured: the DPP form made the key-output pass 9%
  // slower and the bits-only pass faster: C2 5.68 -> 5.59 ms, C4 9.22 -> 8.96 ms)
  constexpr bool MLD_SCAN_DPP = !SCATTER;
  // LDS of the level pipeline: keys (ping-pong), tile prefix and staged bits
  // (under 20 KiB at 2048-key tiles)
  __shared__ __align__(16) uint32_t s_k[2][MLD_T];   // the tile's keys in its level order
  static_assert(MLD_T <= 65535, "tile positions fit 16 bits");
  __shared__ uint16_t s_tpre[(1u << MLD_GMAX) + 1];   // exclusive prefix over d of the tile's counts
  __shared__ __align__(8) uint32_t s_bits[MLD_GMAX][MLD_BW];  // level bits in the tile's level order
  static_assert(MLD_BW % 2 == 0, "64-bit rows");
  __shared__ uint32_t s_nxy[(1u << MLD_GMAX) - 1];    // node tables of the levels (see below)
  __shared__ uint32_t s_lvl0[MLD_GMAX];
  __shared__ uint32_t s_wave[4];
  __shared__ uint16_t s_ex[4][MLD_T / 64];  // per wave: zeros before each chunk of the level
  uint32_t seg, len, first;
  uint64_t base;
  const uint32_t t = blockIdx.x;
  if (!mld_tile(a, t, seg, base, len, first)) return;
  const uint64_t src = mld_src(a, seg, base);
  const uint32_t g = a.g, G = 1u << g, lane = (uint32_t)lane_id();
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t sh = a.L - a.s - g;
  // carry and tile counts per leaf, their exclusive prefixes over d; thread d
  // keeps its leaf's node start and carry prefix in registers until the
  // write-out (they then go to the free key buffer)
  uint32_t xk[MLD_CPW];  // the wave's chunks c = w + 4m, one key per lane
  uint32_t cum_d = 0, cpre_d, ctot;
  {
    const uint32_t d = threadIdx.x;
    uint32_t cv = 0, tv = 0;
    if (d < G) {
      cv = mld_prefix(a, t, d) - mld_prefix(a, first, d);
      tv = a.tcnt[(uint64_t)t * G + d];
      cum_d = a.cum[((uint64_t)seg << g) + d];
    }
    // the tile's keys go straight to registers (level s needs no LDS copy);
    // all loads are in flight together with the tables above (unconditional:
    // positions past the tile load the tile's first key, replaced below)
    K kr[MLD_CPW];
#pragma unroll
    for (uint32_t m = 0; m < MLD_CPW; m++) {
      const uint32_t p = (w + 4 * m) * 64 + lane;
      kr[m] = reinterpret_cast<const K*>(a.keys)[src + (p < len ? p : 0u)];
    }
    cpre_d = block_excl_scan_u32<MLD_SCAN_DPP>(cv, s_wave, ctot);
    uint32_t tot;
    const uint32_t te = block_excl_scan_u32<MLD_SCAN_DPP>(tv, s_wave, tot);
    if (d < G) s_tpre[d] = (uint16_t)te;
    if (d == 0) s_tpre[G] = (uint16_t)tot;
    // positions past the tile hold the pad key ~0: a one at every level, in
    // the last node, so each split leaves it where it is (behind the real keys)
#pragma unroll
    for (uint32_t m = 0; m < MLD_CPW; m++) {
      const uint32_t p = (w + 4 * m) * 64 + lane;
      xk[m] = p < len ? (uint32_t)((uint64_t)kr[m] >> a.kshift) : ~0u;
    }
  }
  __syncthreads();
  // node tables of all levels, from the tile's leaf counts alone.  Entry
  // r = 2^j - 1 + q is node q of level s+j: start S, ZN keys with bit 0 (its
  // first child), ZS keys with bit 0 in the nodes before it.  The split sends
  // the node's k-th zero to S + k and its k-th one to S + ZN + k: for a key at
  // p with zc zeros before it in the tile that is X + zc (zero) or Y + p - zc
  // (one), with X = S - ZS and Y = ZS + ZN, packed X | Y << 16
  {
    const uint32_t r = threadIdx.x;
    const uint32_t j = 31 - __clz(r + 1);
    uint32_t S = 0, zn = 0;
    if (r < G - 1) {
      const uint32_t q = r + 1 - (1u << j), sgj = g - j;
      S = s_tpre[q << sgj];
      zn = s_tpre[(2 * q + 1) << (sgj - 1)] - S;
    }
    uint32_t tot;
    const uint32_t zx = block_excl_scan_u32<MLD_SCAN_DPP>(zn, s_wave, tot);  // over all entries; each level rebased below
    if (r < G - 1 && r + 1 == (1u << j)) s_lvl0[j] = zx;
    __syncthreads();
    if (r < G - 1) {
      const uint32_t zs = zx - s_lvl0[j];
      s_nxy[r] = (S - zs) | ((zs + zn) << 16);
    }
  }
  __syncthreads();
  // the tile's own level pipeline: at level s+j the keys are stably sorted by
  // their top j bits of d (the tile's part of each level-(s+j) node is
  // contiguous); the ballots of the level bit in that order are the level's
  // bits, and a stable split of every node by that bit gives the next order
  const uint32_t nch = (len + 63) >> 6;
  for (uint32_t j = 0; j < g; j++) {
    const uint32_t* A = s_k[j & 1];
    uint32_t* B = s_k[(j + 1) & 1];
    const uint32_t bp = sh + g - 1 - j;  // the key bit of level s+j
    // the wave's chunks, all loads first; the keys stay in registers until
    // the split below
    if (j > 0) {
#pragma unroll
      for (uint32_t m = 0; m < MLD_CPW; m++) xk[m] = A[(w + 4 * m) * 64 + lane];
    }
#pragma unroll
    for (uint32_t m = 0; m < MLD_CPW; m++) {
      const uint64_t ones = wave_ballot((xk[m] >> bp) & 1u);
      if (lane == 0) reinterpret_cast<uint64_t*>(s_bits[j])[w + 4 * m] = ones;
    }
    if (j + 1 == g && !SCATTER) break;  // the last level's order is not needed
    __syncthreads();
    // every wave: exclusive prefix of the chunk zeros (lane = chunk), into the
    // wave's own row so that each chunk reads its base with one broadcast
    // load; zeros before tile position p = that prefix + zeros of p's chunk
    // below p
    const uint32_t v = lane < nch ? 64u - popc64(reinterpret_cast<const uint64_t*>(s_bits[j])[lane]) : 0u;
    const uint32_t incl = wave_incl_scan_dpp(v);
    if (lane < MLD_T / 64) s_ex[w][lane] = (uint16_t)(incl - v);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // stable split of each node: its zeros first, then its ones.  Every
    // chunk of the tile takes part (chunks past the keys hold pad keys,
    // which stay in place), so the node lookups of all chunks are issued
    // before the first is used
    const uint32_t* nt = s_nxy + (1u << j) - 1;
    uint32_t xy[MLD_CPW];
#pragma unroll
    for (uint32_t m = 0; m < MLD_CPW; m++) xy[m] = nt[__builtin_amdgcn_ubfe(xk[m], bp + 1, j)];
#pragma unroll
    for (uint32_t m = 0; m < MLD_CPW; m++) {
      const uint32_t c = w + 4 * m;
      const uint32_t p = c * 64 + lane;
      const uint32_t x = xk[m];
      const bool zero = ((x >> bp) & 1u) == 0;
      const uint64_t zm = wave_ballot(zero);
      const uint32_t cz = s_ex[w][c];
      const uint32_t zc = __builtin_amdgcn_mbcnt_hi((uint32_t)(zm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)zm, cz));
      const uint32_t d0 = (xy[m] & 0xFFFFu) + zc, d1 = (xy[m] >> 16) + p - zc;
      B[zero ? d0 : d1] = x;
    }
    __syncthreads();
  }
  __syncthreads();
  // write the level bits: for level s+j the tile's keys of local node q form
  // one run at NodeStart + carry.  Thread r < G-1 describes run r = (j, q);
  // then the destination words of all runs are dealt out over the block.
  const uint64_t n = a.n;
  const uint32_t nruns = G - 1;  // (j, q) pairs with j < g
  // run tables, carry prefix and leaf bases in the key buffer that no longer
  // holds keys
  char* rt = reinterpret_cast<char*>(s_k[(g + 1) & 1]);
  uint64_t* s_rgs = reinterpret_cast<uint64_t*>(rt);               // per run: global start,
  uint32_t* s_rls = reinterpret_cast<uint32_t*>(rt + 2048);        // start in the staged bits,
  uint32_t* s_rcnt = reinterpret_cast<uint32_t*>(rt + 3072);       // length,
  uint32_t* s_rw = reinterpret_cast<uint32_t*>(rt + 4096);         // first destination word item
  uint32_t* s_cpre = reinterpret_cast<uint32_t*>(rt + 5248);       // carry prefix over d (G + 1)
  uint32_t* s_lb = reinterpret_cast<uint32_t*>(rt + 6400);         // leaf d's key at tile position i goes to s_lb[d] + i
  static_assert(6400 + 4 * (1u << MLD_GMAX) <= 4 * MLD_T, "tables fit one key buffer");
  uint32_t* s_cum = s_lb;  // node starts first, leaf bases after the run tables are built
  if (threadIdx.x < G) {
    s_cpre[threadIdx.x] = cpre_d;
    s_cum[threadIdx.x] = cum_d;
  }
  if (threadIdx.x == 0) s_cpre[G] = ctot;
  __syncthreads();
  {
    const uint32_t r = threadIdx.x;
    uint32_t nw = 0;
    if (r < nruns) {
      const uint32_t j = 31 - __clz(r + 1), q = r + 1 - (1u << j);
      const uint32_t lo = q << (g - j), hi = (q + 1) << (g - j);
      const uint32_t cnt = s_tpre[hi] - s_tpre[lo];
      const uint64_t lb = (uint64_t)(a.s + j - a.lvl0) * n;  // the level's first bit
      const uint64_t gs = lb + s_cum[lo] + (s_cpre[hi] - s_cpre[lo]);
      if (gs < lb || gs + cnt > lb + n) {
        atomicOr(a.err, WTG_FLAG_BADSYM);  // tables inconsistent with the keys: write nothing
      } else if (cnt) {
        nw = (uint32_t)(((gs + cnt - 1) >> 6) - (gs >> 6) + 1);
      }
      s_rgs[r] = gs;
      s_rls[r] = s_tpre[lo];
      s_rcnt[r] = cnt;
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan_u32<MLD_SCAN_DPP>(nw, s_wave, tot);  // (its barriers order the s_cum reads above)
    if (r < nruns) s_rw[r] = ex;
    if (r == 0) s_rw[nruns] = tot;
    if (SCATTER && r < G) s_lb[r] = cum_d + (s_cpre[r + 1] - s_cpre[r]) - s_tpre[r];
    __syncthreads();
  }
  const uint32_t nwords = s_rw[nruns];
  for (uint32_t it = threadIdx.x; it < nwords; it += 256) {
    uint32_t lo = 0, hi = nruns;  // the last run whose first word item is <= it
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_rw[mid] <= it) lo = mid; else hi = mid;
    }
    const uint32_t r = lo, j = 31 - __clz(r + 1);
    const uint64_t gs = s_rgs[r], ge = gs + s_rcnt[r];
    const uint64_t wi = (gs >> 6) + (it - s_rw[r]);
    const uint64_t wb = wi << 6;
    const uint32_t b0 = gs > wb ? (uint32_t)(gs - wb) : 0u;
    const uint32_t b1 = ge < wb + 64 ? (uint32_t)(ge - wb) : 64u;
    uint64_t v = lds_bits64(s_bits[j], (uint32_t)(s_rls[r] + (wb + b0 - gs))) << b0;
    if (b1 < 64) v &= (1ull << b1) - 1;
    if (b0 == 0 && b1 == 64) a.words[wi] = v;
    else if (v) atomicOr((unsigned long long*)&a.words[wi], (unsigned long long)v);
  }
  if (SCATTER) {
    // keys in leaf order to their P_{s+g} positions
#pragma unroll
    for (uint32_t m = 0; m < MLD_CPW; m++) {
      const uint32_t i = (w + 4 * m) * 64 + lane;
      if (i < len) {
        const uint32_t x = s_k[g & 1][i];
        const uint32_t dest = s_lb[(x >> sh) & (G - 1)] + i;
        if (dest < n) a.kout[dest] = x;
        else atomicOr(a.err, WTG_FLAG_BADSYM);
      }
    }
  }
}

// WTnode records from the histogram prefix of all L bits (node (l, p) holds
// the keys with top l bits p: start Cum[p << (L-l)], count up to the next)
__global__ void k_mld_nodes(const uint32_t* __restrict__ cum, uint32_t L, uint64_t n, wtg_node* __restrict__ nodes,
                            uint64_t slots) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= slots) return;
  wtg_node r;
  r.bitmapPtr = 0; r.length = 0; r.parent = WTG_NONE; r.leftChild = WTG_NONE; r.rightChild = WTG_NONE;
  if (v < (1ull << L) - 1) {
    const uint32_t l = 63 - __clzll(v + 1);
    const uint64_t p = v + 1 - (1ull << l);
    const uint32_t sh = L - l;
    const uint64_t st = cum[p << sh], c = cum[(p + 1) << sh] - st;
    if (c > 0 || v == 0) {
      r.bitmapPtr = (uint64_t)l * n + st;
      r.length = c;
      r.parent = v ? (v - 1) / 2 : WTG_NONE;
      if (l + 1 < L) {
        const uint64_t m = cum[(2 * p + 1) << (sh - 1)];
        if (m > st) r.leftChild = 2 * v + 1;
        if (cum[(p + 1) << sh] > m) r.rightChild = 2 * v + 2;
      }
    }
  }
  nodes[v] = r;
}

__global__ void k_mld_nonzero(const uint32_t* __restrict__ h, uint64_t n, unsigned long long* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (; i < n; i += stride) c += h[i] != 0;
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

struct MldPlan {
  uint32_t ngroups, gs[4], gg[4];
  uint64_t maxtiles, maxblocks, maxbins;
  size_t off_keys[2], off_tcnt, off_ploc, off_bsum, off_bpre, off_hist, off_cum[2], off_tst, off_nt, off_tseg,
      off_scst, off_ctr, total;
};

MldPlan mld_plan(uint64_t n, uint32_t L, uint32_t s0) {
  MldPlan P{};
  uint32_t s = s0;
  while (s < L) {
    const uint32_t g = L - s < MLD_GMAX ? L - s : MLD_GMAX;
    P.gs[P.ngroups] = s;
    P.gg[P.ngroups] = g;
    P.ngroups++;
    s += g;
  }
  // a group over level-s nodes has at most ceil(n/T) + min(n, 2^s) tiles
  P.maxtiles = wtg_ceil_div(n, MLD_T) + std::min<uint64_t>(n, 1ull << P.gs[P.ngroups - 1]) + 1;
  P.maxblocks = wtg_ceil_div(P.maxtiles + 1, MLD_SR) + 1;
  P.maxbins = 1ull << L;
  size_t o = 0;
  for (int b = 0; b < 2; b++) {
    P.off_keys[b] = o;
    if (P.ngroups > 1) o = wtg_align256(o + n * 4);
  }
  P.off_tcnt = o; o = wtg_align256(o + (P.maxtiles + 1) * 256 * 4);
  P.off_ploc = o; o = wtg_align256(o + (P.maxtiles + 1) * 256 * 4);
  P.off_bsum = o; o = wtg_align256(o + P.maxblocks * 256 * 4);
  P.off_bpre = o; o = wtg_align256(o + P.maxblocks * 256 * 4);
  P.off_hist = o; o = wtg_align256(o + P.maxbins * 4);
  for (int b = 0; b < 2; b++) { P.off_cum[b] = o; o = wtg_align256(o + (P.maxbins + 1) * 4); }
  P.off_tst = o; o = wtg_align256(o + (P.maxbins + 1) * 4);
  P.off_nt = o; o = wtg_align256(o + P.maxbins * 4);
  P.off_tseg = o; o = wtg_align256(o + P.maxtiles * 4);
  P.off_scst = o; o = wtg_align256(o + (wtg_ceil_div(P.maxbins + 1, 4096) + 2) * 8);
  P.off_ctr = o; o = wtg_align256(o + 256);  // right after the scan states: one fill clears both
  P.total = o;
  return P;
}

// Levels [s0, L) of the keys d_sym (in P_{s0} order: sorted by their top s0
// bits, d_cnt0 = the 2^s0 node sizes; s0 = 0: any order, d_cnt0 unused);
// level l's bits go to bit (l - s0) * n of d_words.  kshift > 0 (s0 = 0): the
// keys are d_sym >> kshift, i.e. these are the top L levels of an
// (L + kshift)-level tree.
template <typename K>
int mld_build_t(wtg_ctx* ctx, const K* d_sym, uint64_t n, uint32_t L, uint32_t s0, const uint32_t* d_cnt0,
                const uint64_t* d_src_off, uint64_t* d_words, wtg_node* d_nodes, uint64_t node_slots,
                unsigned long long* d_distinct, uint32_t kshift = 0) {
  const MldPlan P = mld_plan(n, L, s0);
  WTG_TRY(wtg_ws_reserve(ctx, P.total));
  char* ws = (char*)ctx->ws;
  uint32_t* tcnt = (uint32_t*)(ws + P.off_tcnt);
  uint32_t* ploc = (uint32_t*)(ws + P.off_ploc);
  uint32_t* bsum = (uint32_t*)(ws + P.off_bsum);
  uint32_t* bpre = (uint32_t*)(ws + P.off_bpre);
  uint32_t* hist = (uint32_t*)(ws + P.off_hist);
  uint32_t* cum[2] = {(uint32_t*)(ws + P.off_cum[0]), (uint32_t*)(ws + P.off_cum[1])};
  uint32_t* tst = (uint32_t*)(ws + P.off_tst);
  uint32_t* nt = (uint32_t*)(ws + P.off_nt);
  uint32_t* tseg = (uint32_t*)(ws + P.off_tseg);
  uint64_t* scst = (uint64_t*)(ws + P.off_scst);
  uint32_t* ctr = (uint32_t*)(ws + P.off_ctr);
  uint32_t* kbuf[2] = {(uint32_t*)(ws + P.off_keys[0]), (uint32_t*)(ws + P.off_keys[1])};
  // (the level words are cleared by the leaf-count passes)
  if (d_distinct) WTG_TRY(wtg_fill_u64(ctx, d_distinct, 1, 0));
  const uint32_t ntiles0 = (uint32_t)wtg_ceil_div(n, MLD_T);
  const void* src = d_sym;
  const uint32_t* cum_prev = nullptr;
  if (s0 > 0) {
    // the level-s0 nodes: their starts (cum[1]; group 0 writes cum[0]), tiles
    // per node, first tiles and the tile -> node map
    const uint64_t nb0 = 1ull << s0, st0 = wtg_ceil_div(nb0, 4096);
    WTG_TRY(wtg_fill_bytes_zero(ctx, scst, (char*)ctr + 256 - (char*)scst));
    WTG_LAUNCH(ctx, "mld_cum", k_scan_u32<uint32_t>, dim3((unsigned)st0), dim3(256), 0, d_cnt0, cum[1], nb0, scst,
               ctr, ctx->d_flag);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "mld_ntiles", k_mld_ntiles, dim3((unsigned)wtg_ceil_div(nb0, 256)), dim3(256), 0, cum[1], nb0,
               nt);
    WTG_LAUNCH_CHECK();
    WTG_TRY(wtg_fill_bytes_zero(ctx, scst, (char*)ctr + 256 - (char*)scst));
    WTG_LAUNCH(ctx, "mld_tiles", k_scan_u32<uint32_t>, dim3((unsigned)st0), dim3(256), 0, nt, tst, nb0, scst, ctr,
               ctx->d_flag);
    WTG_LAUNCH_CHECK();
    const uint64_t nt0 = wtg_ceil_div(n, MLD_T) + std::min<uint64_t>(n, nb0);
    WTG_LAUNCH(ctx, "mld_tseg", k_mld_tseg, dim3((unsigned)wtg_ceil_div(nt0, 256)), dim3(256), 0, tst, nb0, tseg,
               (uint64_t)P.maxtiles);
    WTG_LAUNCH_CHECK();
    cum_prev = cum[1];
  }
  for (uint32_t gi = 0; gi < P.ngroups; gi++) {
    const uint32_t s = P.gs[gi], g = P.gg[gi], G = 1u << g;
    const bool last = gi + 1 == P.ngroups;
    MldArgs a{};
    a.keys = src;
    a.n = n;
    a.L = L;
    a.s = s;
    a.g = g;
    a.ntiles0 = ntiles0;
    a.tstart = s ? tst : nullptr;
    a.tseg = s ? tseg : nullptr;
    a.cum_s = s ? cum_prev : nullptr;
    a.tcnt = tcnt;
    a.tcnt_w = tcnt;
    a.ploc = ploc;
    a.bpre = bpre;
    a.cum = cum[gi & 1];
    a.words = d_words;
    a.lvl0 = s0;
    // the input: no bits at or above L - s0 below the top s0 (none above bit 63)
    a.check = gi == 0 && L + kshift < 64 ? 1u : 0u;
    a.chk_shift = d_src_off ? L - s0 : L + kshift;
    a.src_off = gi == 0 ? d_src_off : nullptr;
    a.kshift = gi == 0 ? kshift : 0u;
    a.kout = last ? nullptr : kbuf[gi & 1];
    a.err = ctx->d_flag;
    // grid: an upper bound of the tiles (blocks past the real count exit)
    const uint64_t grid = s == 0 ? ntiles0 : wtg_ceil_div(n, MLD_T) + std::min<uint64_t>(n, 1ull << s);
    const uint32_t nrows = (uint32_t)grid + 1;
    const uint32_t nblocks = (uint32_t)wtg_ceil_div(nrows, MLD_SR);
    if (gi == 0) {
      const dim3 cg((unsigned)wtg_ceil_div(grid, 4 * mld_cnt_tpw<K>()));
      if (G <= 4) WTG_LAUNCH(ctx, "mld_count", (k_mld_count<K, 4>), cg, dim3(256), 0, a);
      else if (G <= 8) WTG_LAUNCH(ctx, "mld_count", (k_mld_count<K, 8>), cg, dim3(256), 0, a);
      else WTG_LAUNCH(ctx, "mld_count", (k_mld_count<K, 0>), cg, dim3(256), 0, a);
    } else {
      const dim3 cg((unsigned)wtg_ceil_div(grid, 4));
      if (G <= 4) WTG_LAUNCH(ctx, "mld_count", (k_mld_count<uint32_t, 4>), cg, dim3(256), 0, a);
      else if (G <= 8) WTG_LAUNCH(ctx, "mld_count", (k_mld_count<uint32_t, 8>), cg, dim3(256), 0, a);
      else WTG_LAUNCH(ctx, "mld_count", (k_mld_count<uint32_t, 0>), cg, dim3(256), 0, a);
    }
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "mld_scan", k_mld_scan1, dim3(nblocks), dim3(256), 0, tcnt, ploc, bsum,
               s ? (const uint32_t*)(tst + (1u << s)) : nullptr, ntiles0, G, nrows);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "mld_scan", k_mld_scan2, dim3(G), dim3(256), 0, bsum, bpre, nblocks, G);
    WTG_LAUNCH_CHECK();
    const uint64_t nbins = 1ull << (s + g);
    WTG_LAUNCH(ctx, "mld_hist", k_mld_hist, dim3((unsigned)wtg_ceil_div(nbins, 256)), dim3(256), 0, a, hist);
    WTG_LAUNCH_CHECK();
    const uint64_t stiles = wtg_ceil_div(nbins, 4096);
    WTG_TRY(wtg_fill_bytes_zero(ctx, scst, (char*)ctr + 256 - (char*)scst));
    WTG_LAUNCH(ctx, "mld_cum", k_scan_u32<uint32_t>, dim3((unsigned)stiles), dim3(256), 0, hist, cum[gi & 1], nbins,
               scst, ctr, ctx->d_flag);
    WTG_LAUNCH_CHECK();
    if (last && d_distinct) {
      WTG_LAUNCH(ctx, "mld_distinct", k_mld_nonzero, dim3((unsigned)std::min<uint64_t>(wtg_ceil_div(nbins, 256), 1024)),
                 dim3(256), 0, hist, nbins, d_distinct);
      WTG_LAUNCH_CHECK();
    }
    // the level bits of the group (and the keys in P_{s+g} order)
    if (gi == 0) {
      if (last) WTG_LAUNCH(ctx, "mld_emit", (k_mld_emit<K, false>), dim3((unsigned)grid), dim3(256), 0, a);
      else WTG_LAUNCH(ctx, "mld_emit", (k_mld_emit<K, true>), dim3((unsigned)grid), dim3(256), 0, a);
    } else {
      if (last) WTG_LAUNCH(ctx, "mld_emit", (k_mld_emit<uint32_t, false>), dim3((unsigned)grid), dim3(256), 0, a);
      else WTG_LAUNCH(ctx, "mld_emit", (k_mld_emit<uint32_t, true>), dim3((unsigned)grid), dim3(256), 0, a);
    }
    WTG_LAUNCH_CHECK();
    if (!last) {
      // the next group's tiles per node of level s+g
      WTG_LAUNCH(ctx, "mld_ntiles", k_mld_ntiles, dim3((unsigned)wtg_ceil_div(nbins, 256)), dim3(256), 0,
                 cum[gi & 1], nbins, nt);
      WTG_LAUNCH_CHECK();
      WTG_TRY(wtg_fill_bytes_zero(ctx, scst, (char*)ctr + 256 - (char*)scst));
      WTG_LAUNCH(ctx, "mld_tiles", k_scan_u32<uint32_t>, dim3((unsigned)stiles), dim3(256), 0, nt, tst, nbins, scst,
                 ctr, ctx->d_flag);
      WTG_LAUNCH_CHECK();
      const uint64_t ntiles_next = wtg_ceil_div(n, MLD_T) + std::min<uint64_t>(n, nbins);
      WTG_LAUNCH(ctx, "mld_tseg", k_mld_tseg, dim3((unsigned)wtg_ceil_div(ntiles_next, 256)), dim3(256), 0, tst,
                 nbins, tseg, (uint64_t)P.maxtiles);
      WTG_LAUNCH_CHECK();
      src = kbuf[gi & 1];
      cum_prev = cum[gi & 1];
    }
  }
  if (d_nodes && node_slots && s0 == 0) {
    WTG_LAUNCH(ctx, "nodes", k_mld_nodes, dim3((unsigned)wtg_ceil_div(node_slots, 256)), dim3(256), 0,
               cum[(P.ngroups - 1) & 1], L, n, d_nodes, node_slots);
    WTG_LAUNCH_CHECK();
  }
  return WTG_OK;
}

}  // namespace

// 1: build with the multi-level direct passes; 0: the level-by-level split pipeline
int wtg_mld_enabled() {
  static int on = -1;
  if (on < 0) {
    const char* e = std::getenv("WTG_LEVEL_SPLIT");
    on = (e && e[0] == '1') ? 0 : 1;
  }
  return on;
}

static int mld_dispatch(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint32_t s0,
                        const uint32_t* d_cnt0, const uint64_t* d_src_off, uint64_t* d_words, wtg_node* d_nodes,
                        uint64_t node_slots, unsigned long long* d_distinct, uint32_t kshift = 0) {
  switch (sym_bytes) {
    case 1: return mld_build_t<uint8_t>(ctx, (const uint8_t*)d_sym, n, L, s0, d_cnt0, d_src_off, d_words, d_nodes,
                                        node_slots, d_distinct, kshift);
    case 2: return mld_build_t<uint16_t>(ctx, (const uint16_t*)d_sym, n, L, s0, d_cnt0, d_src_off, d_words, d_nodes,
                                         node_slots, d_distinct, kshift);
    case 4: return mld_build_t<uint32_t>(ctx, (const uint32_t*)d_sym, n, L, s0, d_cnt0, d_src_off, d_words, d_nodes,
                                         node_slots, d_distinct, kshift);
    default: return mld_build_t<uint64_t>(ctx, (const uint64_t*)d_sym, n, L, s0, d_cnt0, d_src_off, d_words,
                                          d_nodes, node_slots, d_distinct, kshift);
  }
}

// n >= 1, 1 <= L <= 24, n < 2^32
int wtg_mld_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint64_t* d_words,
                  wtg_node* d_nodes, uint64_t node_slots, unsigned long long* d_distinct) {
  return mld_dispatch(ctx, d_sym, sym_bytes, n, L, 0, nullptr, nullptr, d_words, d_nodes, node_slots, d_distinct);
}

// levels [s0, L) of keys already sorted by their top s0 bits (d_cnt0: the 2^s0
// node sizes), level l at bit (l - s0) * n of d_words; 0 <= s0 < L <= 24.
// d_src_off (nullable, s0 > 0): node v's keys are d_keys[d_src_off[v] ..] and
// hold only their low L - s0 bits (the node gives the top s0)
int wtg_mld_build_from(wtg_ctx* ctx, const void* d_keys, int key_bytes, uint64_t n, uint32_t L, uint32_t s0,
                       const uint32_t* d_cnt0, const uint64_t* d_src_off, uint64_t* d_words) {
  return mld_dispatch(ctx, d_keys, key_bytes, n, L, s0, d_cnt0, s0 ? d_src_off : nullptr, d_words, nullptr, 0,
                      nullptr);
}

size_t wtg_mld_ws_bytes(uint64_t n, uint32_t L, uint32_t s0) { return mld_plan(n, L, s0).total; }

// the top Lt levels (Lt <= 24) of an L-level tree over d_sym (L > Lt): the
// multi-level passes over the keys' top Lt bits; level l at bit l * n
int wtg_mld_build_top(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n, uint32_t L, uint32_t Lt,
                      uint64_t* d_words) {
  return mld_dispatch(ctx, d_sym, sym_bytes, n, Lt, 0, nullptr, nullptr, d_words, nullptr, 0, nullptr, L - Lt);
}
