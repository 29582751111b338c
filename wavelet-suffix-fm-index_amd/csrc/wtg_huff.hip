// wtg_huff.hip -- sdsl::wt_huff<> over a byte text on the GPU, its batched
// queries, and FM-index backward-search count over the wt_huff of a BWT.
//
// Construction (wt_huff/wt_pc_par.hpp:321-361; the reference fills the bits by
// recursive stable partitions, build_recursive :193-257).  Here the text is read
// three times and the bits written once, with no per-level data movement:
//   1. k_huff_hist   byte counts -> host builds the Huffman shape (a few us)
//   2. k_huff_count  per (inner node v, segment s) count of text positions whose
//                    code passes through v.  Segment = one wave's S consecutive
//                    bytes.  Up to 32 inner nodes each lane adds a per-symbol
//                    word of byte counters (one byte per node on the symbol's
//                    path) for each of its bytes; larger trees take a segment
//                    byte histogram and add it along each symbol's path.
//                    Table laid out [v][s] with inner nodes in BFS order,
//                    so ONE flat exclusive scan of it gives each (v, s) run's
//                    global bit position: bv_pos[v] is the sum of the inner
//                    nodes before v, the prefix within row v the bits of v before
//                    segment s.
//   3. k_huff_emit   up to 32 inner nodes (k_huff_emit_lane): each lane gathers
//                    the branch bits of its 16 bytes node by node in registers
//                    and a wave scan per node places them in the node's run of a
//                    per-wave LDS staging area.  Larger trees (k_huff_emit): the
//                    wave walks its segment in 64-byte chunks, level by level;
//                    per (chunk, level) the lanes are sorted stably by node id
//                    with one ballot per id bit, a ds_permute puts each node's
//                    bits into one contiguous slice of a ballot word, and one
//                    lane per node ORs its slice onto the node's run.  At the
//                    segment end every run is written out as whole words (edge
//                    words OR-ed).
// Then rank9 + select1 + select0 over the bits, and the per-node rank bases.
#include "wtg_internal.hpp"
#include <cstring>
#include <queue>
#include <vector>

namespace {
constexpr int HF_THREADS = 256;          // 4 waves per block
constexpr int HF_PIECE = 1024;           // bytes per wave per piece (16 per lane)
constexpr uint32_t HF_LEAF = 0x100u;     // child value >= 0x100: leaf, symbol = value - 0x100

// Device copy of the shape.  Inner nodes are indexed 0..n_inner-1 in BFS
// order (the order of their bv_pos).
struct HuffTab {
  uint64_t code[256];        // code bits of each byte, bit l = branch at depth l
  uint64_t bvpos[256];       // per inner node
  uint64_t rbase[256];       // rank1 at bvpos (after the rank directory is built)
  uint64_t C[257];           // C[c] = #bytes < c
  uint16_t child[256][2];    // per inner node: inner index or HF_LEAF + symbol
  uint8_t len[256];          // code length, 0 = absent
  uint8_t lvl_first[64];     // first inner node id at depth l (ids are BFS, so a level is a range)
  uint8_t lvl_bits[64];      // bits for an id offset within level l plus one "none" value
  uint32_t n_inner, max_depth, sigma, single;  // single: the symbol when sigma == 1
};
}  // namespace

struct wtg_huff {
  wtg_huff_shape shape{};
  HuffTab tab{};
  HuffTab* d_tab = nullptr;
  uint64_t* words = nullptr;   // ceil(bv_size/64) words + one zero padding word
  uint64_t* dir = nullptr;
  wtg_select s1{}, s0{};
  uint64_t* d_hist = nullptr;  // 256 counts
  // query copy of the bits: 64-byte blocks of (rank1 at the block start, 7 words),
  // block b covering bits [448b, 448b + 448), so a rank reads one 64-byte sector
  uint64_t* fmr = nullptr;
};

// ------------------------------------------------------------------ kernels
namespace {

// 16 text bytes at p (16-byte vector load when the text is 16-byte aligned)
__device__ __forceinline__ uint4 load16(const uint8_t* __restrict__ text, uint64_t n, uint64_t p, bool aligned) {
  uint4 v;
  if (aligned && p + 16 <= n) {
    v = *reinterpret_cast<const uint4*>(text + p);
  } else {
    uint8_t b[16];
#pragma unroll
    for (int k = 0; k < 16; k++) b[k] = p + k < n ? text[p + k] : 0;
    memcpy(&v, b, 16);
  }
  return v;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// byte histogram: LDS atomics into HF_HIST_REPS replicas per wave (replica =
// lane & (HF_HIST_REPS-1)), rows padded to 257 words so the replicas of one
// byte sit in different banks (a low-entropy text such as DNA otherwise puts
// every lane of an atomic on one of four banks).  All pieces are loaded before
// counting.  Measured on 2^27 DNA bytes, every byte by an atomic: 4 replicas
// 89 us, 8 or 16 replicas 95 us, 4 or 16 unpadded 124 us; taking a wave's
// distinct values one at a time (ballot per value, one lane adds the count)
// 261 us.  With the sampled values counted in registers (below) and only the
// other bytes by atomics: 73 us (108 us in the same run without the sample).
constexpr int HF_HIST_PIECES = 8;  // pieces of 1 KiB per wave
constexpr int HF_HIST_REPS = 4;
constexpr int HF_HIST_ROW = 257;
constexpr int HF_HIST_SET = 8;  // sampled values a wave counts 4 bytes at a time
constexpr uint64_t HF_HIST_BLOCKS = 1024;  // grid cap (4 per CU)
__global__ __launch_bounds__(HF_THREADS) void k_huff_hist(const uint8_t* __restrict__ text, uint64_t n,
                                                          unsigned long long* __restrict__ hist, bool aligned) {
  __shared__ uint32_t s_h[HF_THREADS / 64][HF_HIST_REPS * HF_HIST_ROW];
  uint32_t* flat = &s_h[0][0];
  for (int k = threadIdx.x; k < (HF_THREADS / 64) * HF_HIST_REPS * HF_HIST_ROW; k += HF_THREADS) flat[k] = 0;
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t* h = s_h[w] + (lane & (HF_HIST_REPS - 1)) * HF_HIST_ROW;
  // wave-strided over 8 KiB chunks (the grid is capped at HF_HIST_BLOCKS so
  // the per-block LDS merge and global adds are amortised)
  const uint64_t chunk = (uint64_t)HF_HIST_PIECES * HF_PIECE;
  for (uint64_t c = (uint64_t)blockIdx.x * (HF_THREADS / 64) + w; c * chunk < n;
       c += (uint64_t)gridDim.x * (HF_THREADS / 64)) {
  const uint64_t base0 = c * chunk;
  uint4 v[HF_HIST_PIECES];
#pragma unroll
  for (int pc = 0; pc < HF_HIST_PIECES; pc++) {
    const uint64_t base = base0 + (uint64_t)pc * HF_PIECE + (uint64_t)lane * 16;
    v[pc] = base < n ? load16(text, n, base, aligned) : make_uint4(0, 0, 0, 0);
  }
  // The wave's alphabet sample: the distinct values of byte 0 of each lane's
  // first piece, at most HF_HIST_SET of them (wave-uniform).  Those values are
  // counted 4 bytes at a time in registers (exact zero-byte mask of
  // x ^ value*0x01010101, popcount); the other bytes (and partial pieces) go to
  // the LDS atomics, one byte per lane per step.  More than HF_HIST_SET sampled
  // values: every byte goes to the atomics.
  const bool smp_ok = base0 + (uint64_t)lane * 16 < n;
  const uint32_t smp = v[0].x & 0xFFu;
  uint32_t setv[HF_HIST_SET];
  uint64_t left = wave_ballot(smp_ok);
#pragma unroll
  for (int q = 0; q < HF_HIST_SET; q++) {
    setv[q] = 0x100u;  // no value
    if (left) {
      const uint32_t lv = (uint32_t)__builtin_amdgcn_readlane((int)smp, (int)__builtin_ctzll(left));
      setv[q] = lv;
      left &= ~wave_ballot(smp_ok && smp == lv);
    }
  }
  if (left) setv[0] = 0x100u, setv[1] = 0x100u, setv[2] = 0x100u, setv[3] = 0x100u, setv[4] = 0x100u,
            setv[5] = 0x100u, setv[6] = 0x100u, setv[7] = 0x100u;
  static_assert(HF_HIST_SET == 8, "setv reset above lists 8 entries");
  uint32_t cnt[HF_HIST_SET];
#pragma unroll
  for (int q = 0; q < HF_HIST_SET; q++) cnt[q] = 0;
#pragma unroll
  for (int pc = 0; pc < HF_HIST_PIECES; pc++) {
    const uint64_t base = base0 + (uint64_t)pc * HF_PIECE + (uint64_t)lane * 16;
    const uint32_t valid = base >= n ? 0u : (n - base >= 16 ? 16u : (uint32_t)(n - base));
    const uint32_t x0 = v[pc].x, x1 = v[pc].y, x2 = v[pc].z, x3 = v[pc].w;
    // um: bit 7 of each byte still to count by an atomic
    uint32_t u0 = 0, u1 = 0, u2 = 0, u3 = 0;
    if (valid == 16) {
      uint32_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
#pragma unroll
      for (int q = 0; q < HF_HIST_SET; q++) {
        if (setv[q] < 256u) {
          const uint32_t pat = setv[q] * 0x01010101u;
          uint32_t y, t, z, c = 0;
          y = x0 ^ pat; t = ((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y; z = ~(t | 0x7F7F7F7Fu); m0 |= z; c += __builtin_popcount(z);
          y = x1 ^ pat; t = ((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y; z = ~(t | 0x7F7F7F7Fu); m1 |= z; c += __builtin_popcount(z);
          y = x2 ^ pat; t = ((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y; z = ~(t | 0x7F7F7F7Fu); m2 |= z; c += __builtin_popcount(z);
          y = x3 ^ pat; t = ((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y; z = ~(t | 0x7F7F7F7Fu); m3 |= z; c += __builtin_popcount(z);
          cnt[q] += c;
        }
      }
      u0 = ~m0 & 0x80808080u;
      u1 = ~m1 & 0x80808080u;
      u2 = ~m2 & 0x80808080u;
      u3 = ~m3 & 0x80808080u;
    } else {
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const uint32_t bit = (uint32_t)k < valid ? 0x80u << (8 * (k & 3)) : 0u;
        if (k < 4) u0 |= bit; else if (k < 8) u1 |= bit; else if (k < 12) u2 |= bit; else u3 |= bit;
      }
    }
    // at most 16 steps: each lane clears one pending byte per step
    while (wave_ballot((u0 | u1 | u2 | u3) != 0)) {
      if (u0 | u1 | u2 | u3) {
        const int e = u0 ? 0 : u1 ? 1 : u2 ? 2 : 3;
        const uint32_t um = e == 0 ? u0 : e == 1 ? u1 : e == 2 ? u2 : u3;
        const uint32_t xw = e == 0 ? x0 : e == 1 ? x1 : e == 2 ? x2 : x3;
        const uint32_t bp = (uint32_t)__builtin_ctz(um);  // 7, 15, 23 or 31
        atomicAdd(&h[(xw >> (bp - 7)) & 0xFFu], 1u);
        const uint32_t clr = ~(1u << bp);
        if (e == 0) u0 &= clr; else if (e == 1) u1 &= clr; else if (e == 2) u2 &= clr; else u3 &= clr;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < HF_HIST_SET; q++) {
    if (setv[q] < 256u) {
      const uint32_t tot = wave_sum(cnt[q]);
      if (lane == 0 && tot) atomicAdd(&s_h[w][setv[q]], tot);
    }
  }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < 256; c += HF_THREADS) {
    uint32_t t = 0;
    for (int r = 0; r < (HF_THREADS / 64) * HF_HIST_REPS; r++) t += flat[r * HF_HIST_ROW + c];
    if (t) atomicAdd(&hist[c], (unsigned long long)t);
  }
}

struct HuffLds {
  uint64_t code[256];
  uint16_t child[256][2];
  uint8_t len[256];
  uint8_t lvl_first[64], lvl_bits[64];
};

__device__ __forceinline__ void load_tab(HuffLds& L, const HuffTab* __restrict__ T) {
  for (int k = threadIdx.x; k < 256; k += blockDim.x) {
    L.code[k] = T->code[k];
    L.len[k] = T->len[k];
    L.child[k][0] = T->child[k][0];
    L.child[k][1] = T->child[k][1];
    if (k < 64) {
      L.lvl_first[k] = T->lvl_first[k];
      L.lvl_bits[k] = T->lvl_bits[k];
    }
  }
}

struct SegArgs {
  const uint8_t* text;
  uint64_t n;
  uint32_t S;            // bytes per segment (multiple of HF_PIECE)
  uint64_t nseg;
  uint32_t n_inner;
  uint32_t stage_words;  // per wave
  const HuffTab* tab;
  uint32_t* cnt;         // [n_inner][nseg]
  const uint64_t* pos;   // [n_inner][nseg] (+1): global bit of each run
  uint64_t* words;       // tree bits (zeroed before emit)
  bool aligned;          // text is 16-byte aligned
};

__device__ __forceinline__ void load_piece(uint8_t* s_txt, const uint8_t* __restrict__ text, uint64_t n,
                                           uint64_t base, int lane, bool aligned) {
  *reinterpret_cast<uint4*>(s_txt + lane * 16) = load16(text, n, base + (uint64_t)lane * 16, aligned);
}

// pass 2 for trees of at most 8*K inner nodes: per-lane byte counters.  s_exp[k][c]
// holds a 1 in byte (v & 7) of word k = v >> 3 for every inner node v on the
// path of symbol c, so adding s_exp[.][c] for each text byte counts, per lane,
// the positions passing each node; a lane sees S/64 <= 64 bytes of a segment,
// so a byte never overflows.  At the segment end the even and odd bytes are
// split into 16-bit fields and summed over the wave.
template <int K>
__global__ __launch_bounds__(HF_THREADS) void k_huff_count_swar(SegArgs a) {
  __shared__ uint64_t s_exp[K][256];
  __shared__ uint32_t s_cnt[HF_THREADS / 64][8 * K];
  for (int c = threadIdx.x; c < 256; c += HF_THREADS) {
    uint64_t e[K];
#pragma unroll
    for (int k = 0; k < K; k++) e[k] = 0;
    const uint32_t len = a.tab->len[c];
    const uint64_t code = a.tab->code[c];
    uint32_t node = 0;
    for (uint32_t l = 0; l < len; l++) {
#pragma unroll
      for (int k = 0; k < K; k++)
        if ((node >> 3) == (uint32_t)k) e[k] += 1ull << (8 * (node & 7));
      node = a.tab->child[node][(code >> l) & 1];
    }
#pragma unroll
    for (int k = 0; k < K; k++) s_exp[k][c] = e[k];
  }
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t* cnt = s_cnt[w];
  const uint64_t waves = (uint64_t)gridDim.x * (HF_THREADS / 64);
  constexpr uint64_t EVEN = 0x00FF00FF00FF00FFull;
  for (uint64_t s = (uint64_t)blockIdx.x * (HF_THREADS / 64) + w; s < a.nseg; s += waves) {
    uint64_t acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0;
    for (uint32_t off = 0; off < a.S; off += HF_PIECE) {
      const uint64_t base = s * a.S + off + (uint64_t)lane * 16;
      if (base - (uint64_t)lane * 16 >= a.n) break;
      const uint4 wv = load16(a.text, a.n, base, a.aligned);
      const uint32_t valid = base >= a.n ? 0u : (a.n - base >= 16 ? 16u : (uint32_t)(a.n - base));
#pragma unroll
      for (int b = 0; b < 16; b++) {
        const uint32_t word = b < 4 ? wv.x : b < 8 ? wv.y : b < 12 ? wv.z : wv.w;
        const uint32_t c = (word >> (8 * (b & 3))) & 0xFFu;
        if ((uint32_t)b < valid) {
#pragma unroll
          for (int k = 0; k < K; k++) acc[k] += s_exp[k][c];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint64_t ev = acc[k] & EVEN, od = (acc[k] >> 8) & EVEN;
      // fields: ev lo = nodes 0,2; ev hi = 4,6; od lo = 1,3; od hi = 5,7 (of word k)
      const uint32_t p0 = (uint32_t)__shfl(wave_incl_scan((uint32_t)ev), 63);
      const uint32_t p1 = (uint32_t)__shfl(wave_incl_scan((uint32_t)(ev >> 32)), 63);
      const uint32_t p2 = (uint32_t)__shfl(wave_incl_scan((uint32_t)od), 63);
      const uint32_t p3 = (uint32_t)__shfl(wave_incl_scan((uint32_t)(od >> 32)), 63);
      if (lane == 0) {
        cnt[8 * k + 0] = p0 & 0xFFFFu; cnt[8 * k + 2] = p0 >> 16;
        cnt[8 * k + 4] = p1 & 0xFFFFu; cnt[8 * k + 6] = p1 >> 16;
        cnt[8 * k + 1] = p2 & 0xFFFFu; cnt[8 * k + 3] = p2 >> 16;
        cnt[8 * k + 5] = p3 & 0xFFFFu; cnt[8 * k + 7] = p3 >> 16;
      }
    }
    wave_sync_lds();
    if ((uint32_t)lane < a.n_inner) a.cnt[(uint64_t)lane * a.nseg + s] = cnt[lane];
    wave_sync_lds();
  }
}

// pass 2 for larger trees: a byte histogram of the segment (LDS atomics into 4
// padded replicas), then each present symbol adds its count to every inner node
// on its path.
__global__ __launch_bounds__(HF_THREADS) void k_huff_count_hist(SegArgs a) {
  __shared__ uint32_t s_h[HF_THREADS / 64][4 * HF_HIST_ROW];
  __shared__ uint32_t s_cnt[HF_THREADS / 64][256];
  __shared__ uint64_t s_code[256];
  __shared__ uint16_t s_child[256][2];
  __shared__ uint8_t s_len[256];
  for (int k = threadIdx.x; k < 256; k += HF_THREADS) {
    s_code[k] = a.tab->code[k];
    s_len[k] = a.tab->len[k];
    s_child[k][0] = a.tab->child[k][0];
    s_child[k][1] = a.tab->child[k][1];
  }
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t* hw = s_h[w];
  uint32_t* h = hw + (lane & 3) * HF_HIST_ROW;
  uint32_t* cnt = s_cnt[w];
  const uint64_t waves = (uint64_t)gridDim.x * (HF_THREADS / 64);
  for (uint64_t s = (uint64_t)blockIdx.x * (HF_THREADS / 64) + w; s < a.nseg; s += waves) {
    for (int k = lane; k < 4 * HF_HIST_ROW; k += 64) hw[k] = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) cnt[64 * q + lane] = 0;
    wave_sync_lds();
    for (uint32_t off = 0; off < a.S; off += HF_PIECE) {
      const uint64_t base = s * a.S + off + (uint64_t)lane * 16;
      if (base - (uint64_t)lane * 16 >= a.n) break;
      const uint4 wv = load16(a.text, a.n, base, a.aligned);
      const uint32_t valid = base >= a.n ? 0u : (a.n - base >= 16 ? 16u : (uint32_t)(a.n - base));
#pragma unroll
      for (int b = 0; b < 16; b++) {
        const uint32_t word = b < 4 ? wv.x : b < 8 ? wv.y : b < 12 ? wv.z : wv.w;
        if ((uint32_t)b < valid) atomicAdd(&h[(word >> (8 * (b & 3))) & 0xFFu], 1u);
      }
    }
    wave_sync_lds();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t c = (uint32_t)(64 * q + lane);
      const uint32_t t = hw[c] + hw[HF_HIST_ROW + c] + hw[2 * HF_HIST_ROW + c] + hw[3 * HF_HIST_ROW + c];
      const uint32_t len = t ? (uint32_t)s_len[c] : 0u;
      const uint64_t code = s_code[c];
      uint32_t node = 0;
      for (uint32_t l = 0; l < len; l++) {
        atomicAdd(&cnt[node], t);
        node = s_child[node][(code >> l) & 1];
      }
    }
    wave_sync_lds();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t v = (uint32_t)(64 * q + lane);
      if (v < a.n_inner) a.cnt[(uint64_t)v * a.nseg + s] = cnt[v];
    }
    wave_sync_lds();
  }
}

// 64 bits of the staging area starting at local bit d (d may be negative: zeros)
__device__ __forceinline__ uint64_t stage_bits(const uint64_t* stg, int64_t d) {
  if (d < 0) return d <= -64 ? 0ull : stg[0] << (uint32_t)(-d);
  const uint64_t i = (uint64_t)d >> 6;
  const uint32_t sh = (uint32_t)(d & 63);
  return sh ? (stg[i] >> sh) | (stg[i + 1] << (64 - sh)) : stg[i];
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t l) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), (int)l) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)l);
}

// writes every run of segment s from the wave's staging area as whole words
// (edge words OR-ed); len(v), gstart(v), lstart(v) give node v's run length,
// global bit and staging bit as wave-uniform values
template <typename FL, typename FG, typename FS>
__device__ __forceinline__ void write_runs(const SegArgs& a, uint64_t s, const uint64_t* stg, int lane, FL&& len_of,
                                           FG&& gstart, FS&& lstart) {
  for (uint32_t v = 0; v < a.n_inner; v++) {
    const uint32_t len = len_of(v);
    if (!len) continue;
    const uint64_t g0 = gstart(v);
    const int64_t lsv = (int64_t)lstart(v);
    const uint64_t w0 = g0 >> 6, w1 = (g0 + len - 1) >> 6;
    for (uint64_t kw = w0 + lane; kw <= w1; kw += 64) {
      const uint64_t lo = kw * 64 > g0 ? kw * 64 : g0;
      const uint64_t hi = kw * 64 + 64 < g0 + len ? kw * 64 + 64 : g0 + len;
      const uint32_t nb = (uint32_t)(hi - lo), at = (uint32_t)(lo - kw * 64);
      const uint64_t mask = nb == 64 ? ~0ull : ((1ull << nb) - 1) << at;
      const uint64_t val = stage_bits(stg, lsv + (int64_t)(kw * 64) - (int64_t)g0) & mask;
      if (nb == 64) a.words[kw] = val;
      else if (val) atomicOr(reinterpret_cast<unsigned long long*>(&a.words[kw]), (unsigned long long)val);
    }
  }
}

// pass 3: tree bits
__global__ __launch_bounds__(HF_THREADS) void k_huff_emit(SegArgs a) {
  __shared__ HuffLds L;
  __shared__ uint32_t s_ls[HF_THREADS / 64][256];   // per wave: node run start in the staging area
  __shared__ uint32_t s_run[HF_THREADS / 64][256];  // per wave: bits appended to the node's run
  extern __shared__ uint64_t s_dyn[];
  load_tab(L, a.tab);
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint64_t* stg = s_dyn + (size_t)w * (a.stage_words + HF_PIECE / 8);
  uint8_t* s_txt = reinterpret_cast<uint8_t*>(stg + a.stage_words);
  uint32_t* ls = s_ls[w];
  uint32_t* rn = s_run[w];
  const uint64_t waves = (uint64_t)gridDim.x * (HF_THREADS / 64);
  for (uint64_t s = (uint64_t)blockIdx.x * (HF_THREADS / 64) + w; s < a.nseg; s += waves) {
    wave_sync_lds();
    // local run starts: exclusive scan of this segment's node counts
    uint64_t gpos[4];  // each run's global bit position, loaded up front (lane j: node j+64q)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t v = (uint32_t)(64 * q + lane);
      gpos[q] = v < a.n_inner ? a.pos[(uint64_t)v * a.nseg + s] : 0ull;
    }
    uint32_t carry = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t v = (uint32_t)(64 * q + lane);
      const uint32_t c = v < a.n_inner ? a.cnt[(uint64_t)v * a.nseg + s] : 0u;
      const uint32_t inc = wave_incl_scan(c);
      ls[v] = carry + inc - c;
      rn[v] = 0;
      carry += (uint32_t)__shfl(inc, 63);
    }
    const uint32_t used = (carry + 63) / 64 + 2;
    for (uint32_t k = lane; k < used; k += 64) stg[k] = 0;
    for (uint32_t off = 0; off < a.S; off += HF_PIECE) {
      const uint64_t pbase = s * a.S + off;
      if (pbase >= a.n) break;
      wave_sync_lds();
      load_piece(s_txt, a.text, a.n, pbase, lane, a.aligned);
      wave_sync_lds();
#pragma unroll 1
      for (int k = 0; k < HF_PIECE / 64; k++) {
        const uint64_t p = pbase + (uint64_t)k * 64 + lane;
        const uint32_t c = s_txt[k * 64 + lane];
        const uint32_t len = p < a.n ? L.len[c] : 0u;
        const uint64_t code = L.code[c];
        uint32_t node = 0;
        for (uint32_t l = 0;; l++) {
          const bool act = l < len;
          if (!wave_ballot(act)) break;
          const uint32_t bit = act ? (uint32_t)(code >> l) & 1u : 0u;
          // stable sort of the lanes by node id (finished lanes last): eq ends
          // as the lanes with my node, lt counts the lanes with a smaller one.
          // Ids are taken relative to the level's first node, in its bit width.
          const uint32_t kbits = L.lvl_bits[l];
          const uint32_t vid = act ? node - L.lvl_first[l] : (1u << kbits) - 1;
          uint64_t eq = ~0ull;
          uint32_t lt = 0;
          for (int kb = (int)kbits - 1; kb >= 0; kb--) {
            const uint32_t mb = (vid >> kb) & 1u;
            const uint64_t B = wave_ballot(mb != 0);
            lt += mb ? popc64(eq & ~B) : 0u;
            eq &= mb ? B : ~B;
          }
          const uint32_t rank = mbcnt64(eq);
          // lane lt + rank receives my bit: each node's bits become one
          // contiguous slice of W, in text order
          const int got = __builtin_amdgcn_ds_permute((int)((lt + rank) * 4), (int)bit);
          const uint64_t W = wave_ballot(got != 0);
          if (act && rank == 0) {  // one lane per node appends the node's slice
            const uint32_t g = popc64(eq);
            const uint32_t at = ls[node] + rn[node];
            rn[node] += g;
            const uint64_t slice = (W >> lt) & (g == 64 ? ~0ull : ((1ull << g) - 1));
            const uint32_t wi = at >> 6, sh = at & 63;
            atomicOr(reinterpret_cast<unsigned long long*>(&stg[wi]), (unsigned long long)(slice << sh));
            if (sh && sh + g > 64)
              atomicOr(reinterpret_cast<unsigned long long*>(&stg[wi + 1]), (unsigned long long)(slice >> (64 - sh)));
          }
          if (act) node = L.child[node][bit];  // leaves drop out at l+1 == len
        }
      }
    }
    wave_sync_lds();
    write_runs(a, s, stg, lane, [&](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)rn[v]); },
               [&](uint32_t v) {
                 const uint32_t qv = v >> 6;
                 const uint64_t gq = qv == 0 ? gpos[0] : qv == 1 ? gpos[1] : qv == 2 ? gpos[2] : gpos[3];
                 return readlane64(gq, v & 63);
               },
               [&](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)ls[v]); });
  }
}

// pass 3 for trees of at most 32 inner nodes and 1024-byte segments: lane i
// takes bytes 16i..16i+15 of the segment and, node by node, gathers the branch
// bits of its bytes that pass the node into a register (at most 16 bits); a wave
// scan of the per-node bit counts (two nodes per scan, 16-bit fields) places
// each lane's bits in the node's run of the LDS staging area (one or two LDS
// ORs per node and lane).  The runs are then written out in one pass over all
// the words they touch (lane k: the k-th word, its node found by a search over
// the per-node word prefix).
// With at most HF_SPREAD inner nodes the pass bits are kept spread, node j's
// in bits 5j..5j+4 of the low word: one add per byte then advances the count
// of every node at once (16 bytes per lane, so a field never carries).
constexpr uint32_t HF_SPREAD = 6;
__global__ __launch_bounds__(HF_THREADS) void k_huff_emit_lane(SegArgs a) {
  // per symbol: bit 32+j = branch 1 at inner node j; low word: bit j = passes
  // inner node j (or 1 << 5j, the spread form for small trees)
  __shared__ uint64_t s_pb[256];
  extern __shared__ uint64_t s_dyn[];
  const bool spread = a.n_inner <= HF_SPREAD;
  for (int c = threadIdx.x; c < 256; c += HF_THREADS) {
    uint64_t e = 0;
    const uint32_t len = a.tab->len[c];
    const uint64_t code = a.tab->code[c];
    uint32_t node = 0;
    for (uint32_t l = 0; l < len; l++) {
      const uint32_t b = (uint32_t)(code >> l) & 1u;
      e |= (spread ? 1ull << (5 * node) : 1ull << node) | ((uint64_t)b << (32 + node));
      node = a.tab->child[node][b];
    }
    s_pb[c] = e;
  }
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint64_t* stg = s_dyn + (size_t)w * a.stage_words;
  const uint64_t waves = (uint64_t)gridDim.x * (HF_THREADS / 64);
  const uint32_t ni = a.n_inner;
  // lane v < n_inner: node v's run in the segment (length, global start); lane
  // i: text bytes 16i..16i+15.  The next segment's are loaded before the
  // current one is processed.
  struct SegIn {
    uint32_t cv, valid;
    uint64_t gp;
    uint4 wv;
  };
  const bool has = (uint32_t)lane < ni;
  auto load_seg = [&](uint64_t s) {
    SegIn r;
    r.cv = has ? a.cnt[(uint64_t)lane * a.nseg + s] : 0u;
    r.gp = has ? a.pos[(uint64_t)lane * a.nseg + s] : 0ull;
    const uint64_t base = s * HF_PIECE + (uint64_t)lane * 16;
    r.valid = base >= a.n ? 0u : (a.n - base >= 16 ? 16u : (uint32_t)(a.n - base));
    r.wv = r.valid ? load16(a.text, a.n, base, a.aligned) : make_uint4(0, 0, 0, 0);
    return r;
  };
  uint64_t s = (uint64_t)blockIdx.x * (HF_THREADS / 64) + w;
  SegIn cur{};
  if (s < a.nseg) cur = load_seg(s);
  for (; s < a.nseg; s += waves) {
    SegIn nxt{};
    if (s + waves < a.nseg) nxt = load_seg(s + waves);
    const uint32_t cv = cur.cv;
    const uint64_t gp = cur.gp;
    const uint32_t inc = wave_incl_scan(cv);
    const uint32_t lsv = inc - cv;
    const uint32_t total = (uint32_t)__shfl((int)inc, 63);
    // the 16 bytes' table entries (unconditional reads, masked: a conditional
    // read becomes a branch per byte)
    uint32_t pm[16], bm[16];
#pragma unroll
    for (int b = 0; b < 16; b++) {
      const uint32_t word = b < 4 ? cur.wv.x : b < 8 ? cur.wv.y : b < 12 ? cur.wv.z : cur.wv.w;
      const uint64_t e = s_pb[(word >> (8 * (b & 3))) & 0xFFu];
      const uint32_t vm = (uint32_t)b < cur.valid ? ~0u : 0u;
      pm[b] = (uint32_t)e & vm;
      bm[b] = (uint32_t)(e >> 32) & vm;
    }
    const uint32_t used = (total + 63) / 64 + 2;
    for (uint32_t k = lane; k < used; k += 64) stg[k] = 0;
    wave_sync_lds();
    // nodes j and j+1: their gathered bits (n0, n1 of them) into their runs
    auto place = [&](uint32_t j, uint32_t buf0, uint32_t n0, uint32_t buf1, uint32_t n1) {
      const uint32_t pk = n0 | (n1 << 16);
      const uint32_t ex2 = wave_incl_scan(pk) - pk;
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const uint32_t cntj = t ? n1 : n0;
        const uint32_t bj = t ? buf1 : buf0;
        const uint32_t at = (uint32_t)__builtin_amdgcn_readlane((int)lsv, (int)(j + t)) + ((ex2 >> (16 * t)) & 0xFFFFu);
        if (cntj) {
          const uint32_t wi = at >> 6, sh = at & 63;
          atomicOr(reinterpret_cast<unsigned long long*>(&stg[wi]), (unsigned long long)bj << sh);
          if (sh + cntj > 64)
            atomicOr(reinterpret_cast<unsigned long long*>(&stg[wi + 1]), (unsigned long long)bj >> (64 - sh));
        }
      }
    };
    if (spread) {
      // every node at once: field j of cnt counts the earlier bytes that pass
      // node j, which is where this byte's branch bit goes in node j's bits
      uint32_t buf[HF_SPREAD] = {}, cnt = 0;
#pragma unroll
      for (int b = 0; b < 16; b++) {
#pragma unroll
        for (uint32_t j = 0; j < HF_SPREAD; j++) buf[j] |= ((bm[b] >> j) & 1u) << ((cnt >> (5 * j)) & 31u);
        cnt += pm[b];
      }
#pragma unroll
      for (uint32_t j = 0; j < HF_SPREAD; j += 2)
        if (j < ni) place(j, buf[j], (cnt >> (5 * j)) & 31u, buf[j + 1], (cnt >> (5 * j + 5)) & 31u);
    } else {
      for (uint32_t j = 0; j < ni; j += 2) {
        uint32_t buf0 = 0, buf1 = 0, n0 = 0, n1 = 0;
#pragma unroll
        for (int b = 0; b < 16; b++) {
          buf0 |= ((bm[b] >> j) & 1u) << n0;
          n0 += (pm[b] >> j) & 1u;
          buf1 |= ((bm[b] >> (j + 1)) & 1u) << n1;
          n1 += (pm[b] >> (j + 1)) & 1u;
        }
        place(j, buf0, n0, buf1, n1);
      }
    }
    wave_sync_lds();
    // words touched by node v's run, their prefix over the nodes
    const uint32_t wc = cv ? (uint32_t)(((gp + cv - 1) >> 6) - (gp >> 6) + 1) : 0u;
    const uint32_t winc = wave_incl_scan(wc);
    const uint32_t wst = winc - wc;
    const uint32_t tw = (uint32_t)__shfl((int)winc, 63);
    for (uint32_t k0 = 0; k0 < tw; k0 += 64) {  // uniform loop: the shuffles run on the full wave
      const uint32_t k = k0 + lane;
      uint32_t v = 0;
      for (uint32_t q = 1; q < ni; q++)
        if ((uint32_t)__builtin_amdgcn_readlane((int)wst, (int)q) <= k) v = q;
      const uint32_t len = (uint32_t)__shfl((int)cv, (int)v);
      const uint64_t g0 = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(gp >> 32), (int)v) << 32) |
                          (uint32_t)__shfl((int)(uint32_t)gp, (int)v);
      const int64_t lsx = (int64_t)(uint32_t)__shfl((int)lsv, (int)v);
      const uint32_t wsv = (uint32_t)__shfl((int)wst, (int)v);
      if (k >= tw) continue;
      const uint64_t kw = (g0 >> 6) + (k - wsv);
      const uint64_t lo = kw * 64 > g0 ? kw * 64 : g0;
      const uint64_t hi = kw * 64 + 64 < g0 + len ? kw * 64 + 64 : g0 + len;
      const uint32_t nb = (uint32_t)(hi - lo), at = (uint32_t)(lo - kw * 64);
      const uint64_t mask = nb == 64 ? ~0ull : ((1ull << nb) - 1) << at;
      const uint64_t val = stage_bits(stg, lsx + (int64_t)(kw * 64) - (int64_t)g0) & mask;
      if (nb == 64) a.words[kw] = val;
      else if (val) atomicOr(reinterpret_cast<unsigned long long*>(&a.words[kw]), (unsigned long long)val);
    }
    wave_sync_lds();
    cur = nxt;
  }
}

constexpr uint64_t FMR_BITS = 448;
// one lane per output word: slot 0 of block b = rank1(448 b), slots 1..7 = words 7b..7b+6
__global__ void k_huff_fmr(const uint64_t* __restrict__ D, const uint64_t* __restrict__ words, uint64_t nw,
                           uint64_t nblocks, uint64_t* __restrict__ fmr) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblocks * 8) return;
  const uint64_t b = k >> 3, slot = k & 7;
  if (slot == 0) {
    fmr[k] = rank9_dev(D, words, b * FMR_BITS);
  } else {
    const uint64_t wi = 7 * b + slot - 1;
    fmr[k] = wi < nw ? words[wi] : 0ull;
  }
}

__global__ void k_huff_rbase(HuffTab* tab, const uint64_t* __restrict__ D, const uint64_t* __restrict__ words) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < tab->n_inner) tab->rbase[v] = rank9_dev(D, words, tab->bvpos[v]);
}

// ------------------------------------------------------------------ queries
struct QLds {
  uint64_t code[256];
  uint64_t bvpos[256];
  uint64_t rbase[256];
  uint16_t child[256][2];
  uint8_t len[256];
};
__device__ __forceinline__ void load_q(QLds& Q, const HuffTab* __restrict__ T) {
  for (int k = threadIdx.x; k < 256; k += blockDim.x) {
    Q.code[k] = T->code[k];
    Q.bvpos[k] = T->bvpos[k];
    Q.rbase[k] = T->rbase[k];
    Q.child[k][0] = T->child[k][0];
    Q.child[k][1] = T->child[k][1];
    Q.len[k] = T->len[k];
  }
}

struct HQ {
  const HuffTab* tab;
  const uint64_t* words;
  const uint64_t* D;
  const uint64_t* fmr;
  SelQ s1, s0;
  uint64_t n;
  uint32_t sigma, single;
};

// rank1(i) of the tree bits from the 64-byte query blocks (i <= bv_size)
__device__ __forceinline__ uint64_t fmr_rank(const uint64_t* __restrict__ fmr, uint64_t i) {
  const uint64_t b = i / FMR_BITS;
  const uint32_t r = (uint32_t)(i - b * FMR_BITS);
  const uint4* blk = reinterpret_cast<const uint4*>(fmr + 8 * b);
  const uint4 q0 = blk[0], q1 = blk[1], q2 = blk[2], q3 = blk[3];
  const uint64_t w[8] = {((uint64_t)q0.y << 32) | q0.x, ((uint64_t)q0.w << 32) | q0.z,
                         ((uint64_t)q1.y << 32) | q1.x, ((uint64_t)q1.w << 32) | q1.z,
                         ((uint64_t)q2.y << 32) | q2.x, ((uint64_t)q2.w << 32) | q2.z,
                         ((uint64_t)q3.y << 32) | q3.x, ((uint64_t)q3.w << 32) | q3.z};
  uint64_t cnt = w[0];
  const uint32_t wq = r >> 6, sh = r & 63;
#pragma unroll
  for (uint32_t q = 0; q < 7; q++) {
    const uint64_t m = q < wq ? ~0ull : (q == wq ? ((1ull << sh) - 1) : 0ull);
    cnt += popc64(w[1 + q] & m);
  }
  return cnt;
}

// rank_c(i) by walking c's code (wt_pc_par.hpp:462-485)
__device__ __forceinline__ uint64_t huff_rank(const QLds& Q, const HQ& h, uint64_t i, uint32_t c) {
  const uint32_t len = Q.len[c];
  if (!len) return h.sigma == 1 && c == h.single ? i : 0;
  uint64_t code = Q.code[c];
  uint32_t v = 0;
  for (uint32_t l = 0; l < len && i; l++, code >>= 1) {
    const uint64_t r = fmr_rank(h.fmr, Q.bvpos[v] + i) - Q.rbase[v];
    const uint32_t b = (uint32_t)code & 1u;
    i = b ? r : i - r;
    v = Q.child[v][b];
  }
  return i;
}

__global__ __launch_bounds__(256) void k_huff_rank(HQ h, const uint64_t* __restrict__ qi,
                                                   const uint32_t* __restrict__ qc, uint64_t nq,
                                                   uint64_t* __restrict__ out) {
  __shared__ QLds Q;
  load_q(Q, h.tab);
  __syncthreads();
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = qi[q];
    const uint32_t c = qc[q];
    out[q] = i > h.n ? WTG_NONE : c > 255 ? 0 : huff_rank(Q, h, i, c);
  }
}

// operator[] / inverse_select (wt_pc_par.hpp:430-448, :498-516)
__global__ __launch_bounds__(256) void k_huff_access(HQ h, const uint64_t* __restrict__ qi, uint64_t nq,
                                                     uint64_t* __restrict__ out_sym,
                                                     uint64_t* __restrict__ out_rank) {
  __shared__ QLds Q;
  load_q(Q, h.tab);
  __syncthreads();
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t i = qi[q];
    uint64_t sym = WTG_NONE, rk = WTG_NONE;
    if (i < h.n) {
      if (h.sigma == 1) {
        sym = h.single;
        rk = i;
      } else {
        uint32_t v = 0;
        for (;;) {
          const uint64_t p = Q.bvpos[v] + i;
          const uint64_t r = fmr_rank(h.fmr, p) - Q.rbase[v];
          const uint32_t b = (uint32_t)((h.words[p >> 6] >> (p & 63)) & 1ull);
          i = b ? r : i - r;
          const uint32_t nx = Q.child[v][b];
          if (nx >= HF_LEAF) { sym = nx - HF_LEAF; break; }
          v = nx;
        }
        rk = i;
      }
    }
    out_sym[q] = sym;
    if (out_rank) out_rank[q] = rk;
  }
}

// select(i, c): from the leaf up (wt_pc_par.hpp:527-554).  The inner node at
// depth l of c's path is found by walking the code from the root.
__global__ __launch_bounds__(256) void k_huff_select(HQ h, const uint64_t* __restrict__ qi,
                                                     const uint32_t* __restrict__ qc, uint64_t nq,
                                                     uint64_t* __restrict__ out) {
  __shared__ QLds Q;
  __shared__ uint64_t s_C[257];
  load_q(Q, h.tab);
  for (int k = threadIdx.x; k < 257; k += blockDim.x) s_C[k] = h.tab->C[k];
  __syncthreads();
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = qi[q];
    const uint32_t c = qc[q];
    uint64_t res = WTG_NONE;
    if (c <= 255 && i >= 1 && i <= s_C[c + 1] - s_C[c]) {
      if (h.sigma == 1) {
        res = i - 1;
      } else {
        const uint32_t len = Q.len[c];
        const uint64_t code = Q.code[c];
        res = i - 1;
        for (int l = (int)len - 1; l >= 0; l--) {
          uint32_t v = 0;
          for (int t = 0; t < l; t++) v = Q.child[v][(code >> t) & 1u];
          const uint64_t bp = Q.bvpos[v], br = Q.rbase[v];
          if ((code >> l) & 1u) res = select_plain(h.s1, br + res + 1) - bp;
          else res = select_plain(h.s0, bp - br + res + 1) - bp;
        }
      }
    }
    out[q] = res;
  }
}

// FM-index count: backward search, [sp, ep) <- C[c] + rank_c([sp, ep)) from the
// last pattern byte; rank_c of both ends walks the same nodes, and both rank
// blocks of a level are loaded before either is used.  NP patterns per lane
// (NP t .. NP t + NP-1) interleave their searches step by step.
template <int NP>
__global__ __launch_bounds__(256) void k_fm_count(HQ h, const uint8_t* __restrict__ pats, uint32_t m,
                                                  uint64_t npat, uint64_t* __restrict__ out) {
  __shared__ QLds Q;
  __shared__ uint64_t s_C[257];
  load_q(Q, h.tab);
  for (int k = threadIdx.x; k < 257; k += blockDim.x) s_C[k] = h.tab->C[k];
  __syncthreads();
  const uint64_t ngroup = (npat + NP - 1) / NP;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ngroup;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint8_t* P[NP];
    uint64_t sp[NP], ep[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) {  // the first step needs no rank
      const uint64_t q = NP * t + p < npat ? NP * t + p : NP * t;
      P[p] = pats + q * m;
      const uint32_t c = P[p][m - 1];
      sp[p] = s_C[c];
      ep[p] = NP * t + p < npat ? s_C[c + 1] : s_C[c];
    }
    for (int j = (int)m - 2; j >= 0; j--) {
      bool any = false;
#pragma unroll
      for (int p = 0; p < NP; p++) any |= sp[p] < ep[p];
      if (!any) break;
      uint32_t cc[NP], len[NP], v[NP];
      uint64_t code[NP], a[NP], b[NP];
      uint32_t lmax = 0;
#pragma unroll
      for (int p = 0; p < NP; p++) {
        cc[p] = P[p][j];
        len[p] = 0;
        v[p] = 0;
        a[p] = sp[p];
        b[p] = ep[p];
        code[p] = Q.code[cc[p]];
        if (sp[p] < ep[p]) {
          len[p] = Q.len[cc[p]];
          if (!len[p] && !(h.sigma == 1 && cc[p] == h.single)) a[p] = b[p] = 0;  // absent symbol
        }
        lmax = len[p] > lmax ? len[p] : lmax;
      }
      for (uint32_t l = 0; l < lmax; l++) {
        bool go[NP];
        uint64_t ra[NP], rbv[NP];
#pragma unroll
        for (int p = 0; p < NP; p++) {
          go[p] = l < len[p] && a[p] < b[p];
          const uint64_t bp = Q.bvpos[v[p]];
          ra[p] = fmr_rank(h.fmr, go[p] ? bp + a[p] : 0);
          rbv[p] = fmr_rank(h.fmr, go[p] ? bp + b[p] : 0);
        }
#pragma unroll
        for (int p = 0; p < NP; p++) {
          if (!go[p]) continue;
          const uint64_t rb = Q.rbase[v[p]];
          const uint64_t x = ra[p] - rb, y = rbv[p] - rb;
          const uint32_t bit = (uint32_t)code[p] & 1u;
          if (bit) { a[p] = x; b[p] = y; } else { a[p] -= x; b[p] -= y; }
          v[p] = Q.child[v[p]][bit];
          code[p] >>= 1;
        }
      }
#pragma unroll
      for (int p = 0; p < NP; p++) {
        if (sp[p] >= ep[p]) continue;
        if (a[p] >= b[p]) { sp[p] = ep[p] = 0; continue; }
        sp[p] = s_C[cc[p]] + a[p];
        ep[p] = s_C[cc[p]] + b[p];
      }
    }
#pragma unroll
    for (int p = 0; p < NP; p++)
      if (NP * t + p < npat) out[NP * t + p] = ep[p] > sp[p] ? ep[p] - sp[p] : 0;
  }
}
// patterns per lane: with the loads of a level issued before their use, one
// pattern per lane (full occupancy) measured fastest on 10^6 length-20 DNA
// patterns: 0.66 ms, against 0.75 ms for two and 0.78 ms for four
#define FM_NP 1

// --------------------------------------------------------------- host side
// Huffman shape (wt_huff/wt_huff_par.hpp:92-117) and the BFS layout of the
// byte tree; fills the public shape and the device table
void build_shape(const uint64_t* cnt, uint64_t n, wtg_huff_shape& S, HuffTab& T) {
  std::memset(&S, 0, sizeof(S));
  std::memset(&T, 0, sizeof(T));
  S.n = n;
  for (int c = 0; c < 256; c++) S.C[c + 1] = S.C[c] + cnt[c];
  for (int c = 0; c < 256; c++) { S.c_to_leaf[c] = -1; T.C[c] = S.C[c]; }
  T.C[256] = S.C[256];
  for (int v = 0; v < WTG_HUFF_MAX_NODES; v++) {
    S.child[v][0] = S.child[v][1] = -1;
    S.parent[v] = -1;
    S.sym[v] = -1;
  }
  struct Tmp { uint64_t f; int sym, c0, c1; };
  std::vector<Tmp> t;
  typedef std::pair<uint64_t, uint64_t> PII;  // (freq, node id), smallest first
  std::priority_queue<PII, std::vector<PII>, std::greater<PII>> pq;
  for (int c = 0; c < 256; c++)
    if (cnt[c]) { pq.push(PII(cnt[c], t.size())); t.push_back(Tmp{cnt[c], c, -1, -1}); }
  S.sigma = (uint32_t)t.size();
  if (t.empty()) return;
  while (pq.size() > 1) {
    const PII a = pq.top(); pq.pop();
    const PII b = pq.top(); pq.pop();
    pq.push(PII(a.first + b.first, t.size()));
    t.push_back(Tmp{a.first + b.first, -1, (int)a.second, (int)b.second});
  }
  // breadth-first numbering from the root, child[0] first
  std::vector<int> order(1, (int)t.size() - 1), id(t.size(), -1);
  id[t.size() - 1] = 0;
  for (size_t h = 0; h < order.size(); h++) {
    const Tmp& x = t[order[h]];
    if (x.c0 >= 0) {
      id[x.c0] = (int)order.size(); order.push_back(x.c0);
      id[x.c1] = (int)order.size(); order.push_back(x.c1);
    }
  }
  S.n_nodes = (uint32_t)order.size();
  uint64_t bv = 0;
  std::vector<int> inner_of(order.size(), -1);
  uint32_t ni = 0;
  for (size_t k = 0; k < order.size(); k++) {
    const Tmp& x = t[order[k]];
    S.bv_pos[k] = bv;
    S.sym[k] = x.sym;
    if (x.c0 >= 0) {
      S.child[k][0] = id[x.c0];
      S.child[k][1] = id[x.c1];
      S.parent[id[x.c0]] = (int)k;
      S.parent[id[x.c1]] = (int)k;
      T.bvpos[ni] = bv;
      inner_of[k] = (int)ni++;
      bv += x.f;
    }
  }
  S.bv_size = bv;
  S.n_inner = ni;
  T.n_inner = ni;
  T.sigma = S.sigma;
  for (size_t k = 0; k < order.size(); k++) {
    if (inner_of[k] < 0) continue;
    for (int b = 0; b < 2; b++) {
      const int ch = S.child[k][b];
      T.child[inner_of[k]][b] = (uint16_t)(inner_of[ch] >= 0 ? (uint32_t)inner_of[ch]
                                                              : HF_LEAF + (uint32_t)S.sym[ch]);
    }
  }
  // per-depth id ranges of the inner nodes (a range per depth when the ids are
  // in BFS order; otherwise every level takes the full id width from 0)
  {
    std::vector<int> dep(ni, 0);
    for (size_t k = 0; k < order.size(); k++) {
      if (inner_of[k] < 0) continue;
      int d = 0;
      for (int x = (int)k; S.parent[x] >= 0; x = S.parent[x]) d++;
      dep[inner_of[k]] = d;
    }
    bool bfs = true;
    for (uint32_t v = 1; v < ni; v++) bfs = bfs && dep[v] >= dep[v - 1];
    uint32_t full = 1;
    while ((1u << full) - 1 < ni) full++;
    for (int l = 0; l < 64; l++) {
      uint32_t first = 0, cnt = 0;
      for (uint32_t v = 0; v < ni; v++)
        if (dep[v] == l) { if (!cnt) first = v; cnt++; }
      uint32_t kb = 1;
      while ((1u << kb) - 1 < cnt) kb++;
      T.lvl_first[l] = bfs ? (uint8_t)first : 0;
      T.lvl_bits[l] = bfs ? (uint8_t)kb : (uint8_t)full;
    }
  }
  uint32_t maxd = 0;
  for (size_t k = 0; k < order.size(); k++) {
    if (S.sym[k] < 0) continue;
    const int c = S.sym[k];
    S.c_to_leaf[c] = (int)k;
    uint64_t path = 0;
    uint32_t len = 0;
    for (int x = (int)k; S.parent[x] >= 0; x = S.parent[x]) {
      const int p = S.parent[x];
      path = (path << 1) | (S.child[p][1] == x ? 1u : 0u);
      len++;
    }
    S.path[c] = path;
    S.path_len[c] = len;
    T.code[c] = path;
    T.len[c] = (uint8_t)len;
    if (len > maxd) maxd = len;
  }
  S.max_depth = maxd;
  T.max_depth = maxd;
  if (S.sigma == 1) T.single = (uint32_t)t[0].sym;
}

HQ make_hq(const wtg_huff* wt) {
  HQ h{};
  h.tab = wt->d_tab;
  h.words = wt->words;
  h.D = wt->dir;
  h.fmr = wt->fmr;
  const uint64_t W = (wt->shape.bv_size + 63) / 64 + 1;
  h.s1 = SelQ{wt->s1.superblock, wt->s1.mini, wt->s1.long_index, wt->s1.long_pool, wt->words, W, 1};
  h.s0 = SelQ{wt->s0.superblock, wt->s0.mini, wt->s0.long_index, wt->s0.long_pool, wt->words, W, 0};
  h.n = wt->shape.n;
  h.sigma = wt->shape.sigma;
  h.single = wt->tab.single;
  return h;
}

unsigned query_grid(uint64_t nq) {
  const uint64_t b = wtg_ceil_div(nq, 256);
  return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}
}  // namespace

extern "C" int wtg_huff_destroy(wtg_ctx* ctx, wtg_huff* wt) {
  if (!wt) return WTG_OK;
  WTG_REQUIRE(ctx, "null ctx");
  wtg_free(ctx, wt->d_tab);
  wtg_free(ctx, wt->words);
  wtg_free(ctx, wt->dir);
  wtg_free(ctx, wt->d_hist);
  wtg_free(ctx, wt->fmr);
  wtg_select_free(ctx, &wt->s1);
  wtg_select_free(ctx, &wt->s0);
  delete wt;
  return WTG_OK;
}

// After the tree bits are in wt->words (and the shape in wt->shape / wt->tab,
// the table on the device when there are inner nodes): rank9 + select1 +
// select0, the query blocks and the per-node rank bases.
static int huff_finish(wtg_ctx* ctx, wtg_huff* wt, const uint64_t* hist) {
  const wtg_huff_shape& S = wt->shape;
  const uint64_t nw = (S.bv_size + 63) / 64 + 1;
  // the ones of the tree bits: each byte contributes the ones of its code
  uint64_t ones = 0;
  for (int c = 0; c < 256; c++) ones += hist[c] * (uint64_t)__builtin_popcountll(wt->tab.code[c]);
  wtg_rs_pending pend;
  WTG_TRY(wtg_rank_select_begin(ctx, wt->words, S.bv_size, wt->dir, true, &wt->s1, &wt->s0, ones, &pend));
  const uint64_t nblocks = S.bv_size / FMR_BITS + 1;
  if (S.bv_size > 0) WTG_TRY(wtg_malloc(ctx, (void**)&wt->fmr, nblocks * 64));
  // one host read for the long-superblock counts and the per-node rank bases
  uint32_t nl[2] = {0, 0};
  uint64_t rb[256];
  if (S.n_inner > 0) {
    WTG_LAUNCH(ctx, "huff_rbase", k_huff_rbase, dim3(1), dim3(256), 0, wt->d_tab, wt->dir, wt->words);
    WTG_LAUNCH_CHECK();
  } else {
    WTG_TRY(wtg_stage_h2d(ctx, wt->d_tab, &wt->tab, sizeof(HuffTab)));
  }
  // both reads into the pinned staging area, one sync
  uint8_t* st = ctx->h_stage + WTG_STAGE_D2H;
  static_assert(64 + sizeof(rb) <= WTG_STAGE_D2H_MAX, "staging area");
  if ((pend.active && hipMemcpyAsync(st, pend.d_nl, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
      (S.n_inner > 0 &&
       hipMemcpyAsync(st + 64, wt->d_tab->rbase, sizeof(rb), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    wtg_set_error("wt_huff: rank base copy failed");
    return WTG_ERR_HIP;
  }
  // the query blocks (rank9 directory done): queued after the read, so the
  // device has them to run while the host takes the counts
  if (S.bv_size > 0) {
    WTG_LAUNCH(ctx, "huff_fmr", k_huff_fmr, dim3((unsigned)wtg_ceil_div(nblocks * 8, 256)), dim3(256), 0,
               wt->dir, wt->words, nw, nblocks, wt->fmr);
    WTG_LAUNCH_CHECK();
  }
  if (pend.active) std::memcpy(nl, st, 8);
  if (S.n_inner > 0) std::memcpy(rb, st + 64, sizeof(rb));
  WTG_TRY(wtg_rank_select_end(ctx, &pend, nl));
  if (S.n_inner > 0) {
    std::memcpy(wt->tab.rbase, rb, sizeof(rb));
    uint32_t ni = 0;
    for (uint32_t v = 0; v < wt->shape.n_nodes; v++)
      if (wt->shape.child[v][0] >= 0) wt->shape.bv_pos_rank[v] = rb[ni++];
  }
  return WTG_OK;
}

extern "C" int wtg_huff_build(wtg_ctx* ctx, const uint8_t* d_text, uint64_t n, wtg_huff** out) {
  WTG_REQUIRE(ctx && out && (n == 0 || d_text), "wtg_huff_build: null argument");
  WTG_REQUIRE(n < (1ull << 40), "wtg_huff_build: n >= 2^40");
  WTG_TRY(wtg_set_device(ctx));
  *out = nullptr;
  wtg_huff* wt = new wtg_huff();
  int rc = WTG_OK;
  uint64_t hist[256] = {0};
  wt->shape.n = n;
  if (n == 0) {  // wt_pc_par.hpp:323-324: empty tree
    for (int c = 0; c < 256; c++) wt->shape.c_to_leaf[c] = -1;
    *out = wt;
    return WTG_OK;
  }
  if ((rc = wtg_malloc(ctx, (void**)&wt->d_hist, 256 * 8)) != WTG_OK) goto fail;
  if ((rc = wtg_malloc(ctx, (void**)&wt->d_tab, sizeof(HuffTab))) != WTG_OK) goto fail;
  if ((rc = wtg_fill_u64(ctx, wt->d_hist, 256, 0)) != WTG_OK) goto fail;
  {
    uint64_t hgrid = wtg_ceil_div(n, (uint64_t)(HF_THREADS / 64) * HF_HIST_PIECES * HF_PIECE);
    if (hgrid > HF_HIST_BLOCKS) hgrid = HF_HIST_BLOCKS;
    const unsigned grid = (unsigned)hgrid;
    WTG_LAUNCH(ctx, "huff_hist", k_huff_hist, dim3(grid), dim3(HF_THREADS), 0, d_text, n,
               (unsigned long long*)wt->d_hist, ((uintptr_t)d_text & 15) == 0);
    WTG_LAUNCH_CHECK();
  }
  if ((rc = wtg_stage_d2h(ctx, hist, wt->d_hist, sizeof(hist))) != WTG_OK) goto fail;
  build_shape(hist, n, wt->shape, wt->tab);
  if (wt->shape.max_depth > 64) {
    wtg_set_error("wtg_huff_build: Huffman code longer than 64 bits");
    rc = WTG_ERR_UNSUPPORTED;
    goto fail;
  }
  {
    const wtg_huff_shape& S = wt->shape;
    const uint64_t nw = (S.bv_size + 63) / 64 + 1;  // + zero padding word
    if ((rc = wtg_malloc(ctx, (void**)&wt->words, nw * 8)) != WTG_OK) goto fail;
    if ((rc = wtg_malloc(ctx, (void**)&wt->dir, wtg_rank9_words(S.bv_size) * 8)) != WTG_OK) goto fail;
    if ((rc = wtg_fill_u64(ctx, wt->words, nw, 0)) != WTG_OK) goto fail;
    if (S.n_inner > 0) {
      // segment: 1024..4096 bytes so the count table stays <= n/4 bytes
      uint32_t S_bytes = HF_PIECE;
      while (S_bytes < 4 * HF_PIECE && (uint64_t)S.n_inner * 16 > S_bytes) S_bytes *= 2;
      const uint64_t nseg = wtg_ceil_div(n, S_bytes);
      const uint64_t entries = (uint64_t)S.n_inner * nseg;
      const uint64_t scan_tiles = wtg_ceil_div(entries, 4096);
      const size_t o_cnt = 0;
      const size_t o_pos = wtg_align256(o_cnt + entries * 4);
      const size_t o_st = wtg_align256(o_pos + (entries + 1) * 8);
      const size_t o_ctr = wtg_align256(o_st + scan_tiles * 8);
      const size_t total = o_ctr + 256;
      if ((rc = wtg_ws_reserve(ctx, total)) != WTG_OK) goto fail;
      char* ws = (char*)ctx->ws;
      if ((rc = wtg_fill_bytes_zero(ctx, ws + o_st, total - o_st)) != WTG_OK) goto fail;
      if ((rc = wtg_stage_h2d(ctx, wt->d_tab, &wt->tab, sizeof(HuffTab))) != WTG_OK) goto fail;
      SegArgs a{};
      a.text = d_text;
      a.n = n;
      a.S = S_bytes;
      a.nseg = nseg;
      a.n_inner = S.n_inner;
      a.stage_words = (uint32_t)(((uint64_t)S_bytes * S.max_depth + 63) / 64 + 2);
      a.tab = wt->d_tab;
      a.cnt = (uint32_t*)(ws + o_cnt);
      a.pos = (const uint64_t*)(ws + o_pos);
      a.words = wt->words;
      a.aligned = ((uintptr_t)d_text & 15) == 0;
      const unsigned grid = (unsigned)std::min<uint64_t>(wtg_ceil_div(nseg, 4), 8192);
      // S_bytes stays HF_PIECE for these tree sizes: 16 bytes per lane per segment
      if (S.n_inner <= 8 && S_bytes <= 64 * 255) {
        WTG_LAUNCH(ctx, "huff_count", k_huff_count_swar<1>, dim3(grid), dim3(HF_THREADS), 0, a);
      } else if (S.n_inner <= 16 && S_bytes <= 64 * 255) {
        WTG_LAUNCH(ctx, "huff_count", k_huff_count_swar<2>, dim3(grid), dim3(HF_THREADS), 0, a);
      } else if (S.n_inner <= 32 && S_bytes <= 64 * 255) {
        WTG_LAUNCH(ctx, "huff_count", k_huff_count_swar<4>, dim3(grid), dim3(HF_THREADS), 0, a);
      } else {
        WTG_LAUNCH(ctx, "huff_count", k_huff_count_hist, dim3(grid), dim3(HF_THREADS), 0, a);
      }
      WTG_LAUNCH_CHECK();
      WTG_LAUNCH(ctx, "huff_scan", k_scan_u32<uint64_t>, dim3((unsigned)scan_tiles), dim3(256), 0,
                 (const uint32_t*)(ws + o_cnt), (uint64_t*)(ws + o_pos), entries, (uint64_t*)(ws + o_st),
                 (uint32_t*)(ws + o_ctr), ctx->d_flag);
      WTG_LAUNCH_CHECK();
      const size_t lds = (size_t)4 * (a.stage_words + HF_PIECE / 8) * 8;
      if (lds > 150 * 1024) {
        wtg_set_error("wtg_huff_build: staging area exceeds LDS");
        rc = WTG_ERR_UNSUPPORTED;
        goto fail;
      }
      if (S.n_inner <= 32 && S_bytes == HF_PIECE) {
        WTG_LAUNCH(ctx, "huff_emit", k_huff_emit_lane, dim3(grid), dim3(HF_THREADS), (size_t)4 * a.stage_words * 8,
                   a);
      } else {
        WTG_LAUNCH(ctx, "huff_emit", k_huff_emit, dim3(grid), dim3(HF_THREADS), lds, a);
      }
      WTG_LAUNCH_CHECK();
    }
    if ((rc = huff_finish(ctx, wt, hist)) != WTG_OK) goto fail;
  }
  if ((rc = wtg_finish(ctx)) != WTG_OK) goto fail;
  *out = wt;
  return WTG_OK;
fail:
  wtg_huff_destroy(ctx, wt);
  return rc;
}

extern "C" int wtg_huff_from_bits(wtg_ctx* ctx, const uint64_t* counts, const uint64_t* d_words, uint64_t nbits,
                                  wtg_huff** out) {
  WTG_REQUIRE(ctx && out && counts, "wtg_huff_from_bits: null argument");
  WTG_TRY(wtg_set_device(ctx));
  *out = nullptr;
  uint64_t hist[256], n = 0;
  for (int c = 0; c < 256; c++) n += (hist[c] = counts[c]);
  WTG_REQUIRE(n < (1ull << 40), "wtg_huff_from_bits: n >= 2^40");
  wtg_huff* wt = new wtg_huff();
  int rc = WTG_OK;
  wt->shape.n = n;
  if (n == 0) {
    for (int c = 0; c < 256; c++) wt->shape.c_to_leaf[c] = -1;
    *out = wt;
    return WTG_OK;
  }
  build_shape(hist, n, wt->shape, wt->tab);
  if (wt->shape.max_depth > 64) {
    wtg_set_error("wtg_huff_from_bits: Huffman code longer than 64 bits");
    wtg_huff_destroy(ctx, wt);
    return WTG_ERR_UNSUPPORTED;
  }
  {
    const uint64_t nb = wt->shape.bv_size, nw = (nb + 63) / 64 + 1;
    if (nb != nbits || (nb && !d_words)) {
      wtg_set_error("wtg_huff_from_bits: the bits do not match the Huffman tree of the counts");
      wtg_huff_destroy(ctx, wt);
      return WTG_ERR_ARG;
    }
    if ((rc = wtg_malloc(ctx, (void**)&wt->d_tab, sizeof(HuffTab))) != WTG_OK) goto fail;
    if ((rc = wtg_malloc(ctx, (void**)&wt->words, nw * 8)) != WTG_OK) goto fail;
    if ((rc = wtg_malloc(ctx, (void**)&wt->dir, wtg_rank9_words(nb) * 8)) != WTG_OK) goto fail;
    if ((rc = wtg_fill_u64(ctx, wt->words, nw, 0)) != WTG_OK) goto fail;
    if (nb && hipMemcpyAsync(wt->words, d_words, (nb + 63) / 64 * 8, hipMemcpyDeviceToDevice, ctx->stream) !=
                  hipSuccess) {
      wtg_set_error("wtg_huff_from_bits: copy failed");
      rc = WTG_ERR_HIP;
      goto fail;
    }
    if (nb && hipMemcpyAsync(wt->d_tab, &wt->tab, sizeof(HuffTab), hipMemcpyHostToDevice, ctx->stream) !=
                  hipSuccess) {
      wtg_set_error("wtg_huff_from_bits: table copy failed");
      rc = WTG_ERR_HIP;
      goto fail;
    }
    if ((rc = huff_finish(ctx, wt, hist)) != WTG_OK) goto fail;
  }
  if ((rc = wtg_finish(ctx)) != WTG_OK) goto fail;
  *out = wt;
  return WTG_OK;
fail:
  wtg_huff_destroy(ctx, wt);
  return rc;
}

extern "C" int wtg_huff_get_shape(const wtg_huff* wt, wtg_huff_shape* out) {
  WTG_REQUIRE(wt && out, "null argument");
  *out = wt->shape;
  return WTG_OK;
}
extern "C" int wtg_huff_data(const wtg_huff* wt, const uint64_t** d_words, const uint64_t** d_dir) {
  WTG_REQUIRE(wt, "null wt");
  if (d_words) *d_words = wt->words;
  if (d_dir) *d_dir = wt->dir;
  return WTG_OK;
}
extern "C" const wtg_select* wtg_huff_select1(const wtg_huff* wt) { return wt ? &wt->s1 : nullptr; }
extern "C" const wtg_select* wtg_huff_select0(const wtg_huff* wt) { return wt ? &wt->s0 : nullptr; }

static int hq_check(wtg_ctx* ctx, const wtg_huff* wt, uint64_t nq, const void* a, const void* b) {
  WTG_REQUIRE(ctx && wt && (nq == 0 || (a && b)), "null argument");
  if (wt->shape.n == 0 && nq) {
    wtg_set_error("query on an empty wt_huff");
    return WTG_ERR_ARG;
  }
  return WTG_OK;
}

extern "C" int wtg_huff_rank_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint64_t* d_i,
                                   const uint32_t* d_c, uint64_t nq, uint64_t* d_out) {
  WTG_TRY(hq_check(ctx, wt, nq, d_i, d_out));
  WTG_REQUIRE(nq == 0 || d_c, "null d_c");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "huff_rank", k_huff_rank, dim3(query_grid(nq)), dim3(256), 0, make_hq(wt), d_i, d_c, nq,
             d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_huff_access_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint64_t* d_i, uint64_t nq,
                                     uint64_t* d_sym, uint64_t* d_rank) {
  WTG_TRY(hq_check(ctx, wt, nq, d_i, d_sym));
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "huff_access", k_huff_access, dim3(query_grid(nq)), dim3(256), 0, make_hq(wt), d_i, nq,
             d_sym, d_rank);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_huff_select_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint64_t* d_i,
                                     const uint32_t* d_c, uint64_t nq, uint64_t* d_out) {
  WTG_TRY(hq_check(ctx, wt, nq, d_i, d_out));
  WTG_REQUIRE(nq == 0 || d_c, "null d_c");
  if (!nq) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "huff_select", k_huff_select, dim3(query_grid(nq)), dim3(256), 0, make_hq(wt), d_i, d_c,
             nq, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
extern "C" int wtg_fm_count_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint8_t* d_pats, uint32_t m,
                                  uint64_t npat, uint64_t* d_out) {
  WTG_TRY(hq_check(ctx, wt, npat, d_pats, d_out));
  WTG_REQUIRE(m >= 1, "pattern length 0");
  if (!npat) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_LAUNCH(ctx, "fm_count", k_fm_count<FM_NP>, dim3(query_grid((npat + FM_NP - 1) / FM_NP)), dim3(256), 0,
             make_hq(wt), d_pats, m, npat, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
