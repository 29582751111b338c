// wtg_rank_select.hip -- (b) rank_support_v<1,1> directory and (b') the
// select_support_mcl<b,1> samples, built in one pass over the bit words.
//
// Directory (rank/rank_support_v_par.hpp:78-113): entry b covers words
// [8b, 8b+8): D[2b] = ones in words [0,8b); D[2b+1] = sum_j c_j << (63-9j) for
// j = 1..min(7, m-1), plus c_m << (63-9m) when the entry has m < 8 words.
//
// Select content (select/select_support_mcl_par.hpp:227-276 / :483-561): for
// superblock k of 4096 arguments, a = pos(4096k), e = pos(min(4096k+4095,m-1));
// e-a > logn^4 -> "long" (all 4096 positions) else "mini" (pos(4096k+64t)-a).
//
// Two passes over the words: ones per 512-word tile (k_rank_tile_sums) and,
// after their exclusive scan, one lane per directory entry (k_rank_entries):
// the lane's 8 word counts give the entry's 9-bit fields, a DPP scan of the
// entry totals its prefix, and for each word the lane emits the sample of the
// argument rank j%64==0 the word holds (at most one per bit value): the low
// 32 bits of pos(j) straight into slot j/64 of the final mini
// array (mini offsets are < logn^4 < 2^32, so subtracting the low 32 bits of
// the superblock start later gives the exact offset), and the superblock
// starts/ends.  A second kernel classifies long/mini superblocks and rebases
// the mini offsets; a third fills the rare long superblocks.
#include "wtg_internal.hpp"
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#define FLAG_LOOKBACK 1u

struct SelOut {
  uint32_t* mini;        // worst-case sized: ceil(nbits/4096)*64 entries
  uint64_t* sbstart;     // ceil(nbits/4096)+1
  uint64_t* sbend;       // ceil(nbits/4096)+1
  unsigned long long* last;  // 1 + position of the last argument (0 = none)
};

struct RankSelArgs {
  const uint64_t* words;
  uint64_t nbits;
  uint64_t W;            // words
  uint64_t S;            // capacity>>9 (entries 0..S)
  uint64_t* D;           // out (BUILD) or in (FROM_D)
  uint64_t* states;
  uint32_t* counter;
  uint32_t* err;
  const uint64_t* tpre;  // BUILD: ones before each tile (exclusive scan of k_rank_tile_sums)
  unsigned long long* ones_total;  // out: total ones
  SelOut sel[2];         // [0] = select0, [1] = select1
  int want[2];
};

// ones per tile of 512 words (64 directory entries), one wave per tile,
// lane-consecutive loads
constexpr int RS_TILE_WORDS = 512;
__global__ __launch_bounds__(256) void k_rank_tile_sums(const uint64_t* __restrict__ words, uint64_t W,
                                                        uint64_t ntiles, uint32_t* __restrict__ tsum, bool aligned) {
  const uint64_t waves = (uint64_t)gridDim.x * 4;
  for (uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += waves) {
    const uint64_t q0 = t * (RS_TILE_WORDS / 2) + lane_id();  // 16-byte pairs of words
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < RS_TILE_WORDS / 128; k++) {
      const uint64_t wi = 2 * (q0 + (uint64_t)k * 64);
      if (aligned && wi + 2 <= W) {
        const uint4 v = reinterpret_cast<const uint4*>(words)[wi >> 1];
        c += popc64(((uint64_t)v.y << 32) | v.x) + popc64(((uint64_t)v.w << 32) | v.z);
      } else {
        c += (wi < W ? popc64(words[wi]) : 0u) + (wi + 1 < W ? popc64(words[wi + 1]) : 0u);
      }
    }
    c = wave_sum(c);
    if (lane_id() == 0) tsum[t] = c;
  }
}

// Directory entries and select samples, one lane per word, one wave per tile
// of 512 words.  Lanes 8i..8i+7 of a round hold directory entry i of the round;
// a word holds at most one argument rank that is a multiple of 64 (per bit
// value), so each lane emits at most one sample per select structure.
__device__ __forceinline__ void word_sample(uint64_t x, uint64_t r, uint32_t c, uint64_t w, const SelOut& o) {
  if (c == 0) return;
  const uint64_t j = (r + 63) & ~63ull;
  if (j < r + c) {
    const uint64_t pos = (w << 6) + select_in_word(x, (uint32_t)(j - r));
    o.mini[j >> 6] = (uint32_t)pos;
    if ((j & 4095) == 0) o.sbstart[j >> 12] = pos;
  }
  uint64_t je = ((r >> 12) << 12) + 4095;
  if (je < r) je += 4096;
  if (je < r + c) o.sbend[je >> 12] = (w << 6) + select_in_word(x, (uint32_t)(je - r));
}

// inclusive scan over the 64 lanes: DPP shifts inside each row of 16, then the
// row totals read back as scalars
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  const uint32_t row = (uint32_t)lane_id() >> 4;
  return v + (row == 0 ? 0u : row == 1 ? r0 : row == 2 ? r0 + r1 : r0 + r1 + r2);
}

// Directory entries and select samples: one lane per directory entry (8
// words), one wave per tile of 512 words.  The lane's word counts give the
// entry's 9-bit fields directly; one DPP scan of the entry totals gives each
// entry's prefix inside the tile.
__global__ __launch_bounds__(256) void k_rank_entries(RankSelArgs a, uint64_t ntiles, bool write_dir, bool aligned) {
  const uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= ntiles) return;
  const int lane = lane_id();
  const uint64_t e = t * 64 + (uint64_t)lane;  // directory entry
  const uint64_t w0 = e * 8;
  uint64_t x[8];
  if (aligned && w0 + 8 <= a.W) {
    const uint4* q = reinterpret_cast<const uint4*>(a.words + w0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = q[k];
      x[2 * k] = ((uint64_t)v.y << 32) | v.x;
      x[2 * k + 1] = ((uint64_t)v.w << 32) | v.z;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = w0 + k < a.W ? a.words[w0 + k] : 0ull;
  }
  uint32_t f[9];  // f[j] = ones in words [w0, w0 + j)
  f[0] = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) f[k + 1] = f[k] + popc64(x[k]);
  const uint32_t incl = wave_incl_scan_dpp(f[8]);
  const uint64_t pre = a.tpre[t] + (incl - f[8]);  // ones before word w0
  if (write_dir && e <= a.S) {
    const uint64_t m = a.W > w0 ? (a.W - w0 < 8 ? a.W - w0 : 8) : 0;  // words of the entry
    uint64_t F = 0;
#pragma unroll
    for (int j = 1; j < 8; j++)
      if ((uint64_t)j <= m) F |= (uint64_t)f[j] << (63 - 9 * j);
    a.D[2 * e] = pre;
    a.D[2 * e + 1] = F;
  }
  uint64_t last1 = 0, last0 = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint64_t w = w0 + k;
    if (w < a.W) {
      const uint32_t c = f[k + 1] - f[k];
      const uint64_t r = pre + f[k];
      const uint64_t lo = w * 64;
      if (a.want[1]) {
        word_sample(x[k], r, c, w, a.sel[1]);
        if (c) last1 = lo + (63 - __clzll(x[k])) + 1;
      }
      if (a.want[0]) {
        const uint64_t vm = a.nbits - lo >= 64 ? ~0ull : ((1ull << (a.nbits - lo)) - 1);
        const uint64_t z = ~x[k] & vm;
        word_sample(z, lo - r, popc64(z), w, a.sel[0]);
        if (z) last0 = lo + (63 - __clzll(z)) + 1;
      }
      if (w + 1 == a.W) *a.ones_total = r + c;
    }
  }
  // last argument position (+1): only the tile holding the last one (the last
  // zero) writes it, a plain store; the scanned tile counts identify that tile
  const uint64_t T = a.tpre[ntiles], tp0 = a.tpre[t], tp1 = a.tpre[t + 1];
  const uint64_t v0 = t * RS_TILE_WORDS * 64 < a.nbits ? t * RS_TILE_WORDS * 64 : a.nbits;
  const uint64_t v1 = (t + 1) * RS_TILE_WORDS * 64 < a.nbits ? (t + 1) * RS_TILE_WORDS * 64 : a.nbits;
  const bool has_last1 = tp0 < T && tp1 == T;
  const bool has_last0 = v0 - tp0 < a.nbits - T && v1 - tp1 == a.nbits - T;
  if (has_last1 || has_last0) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t o1 = (uint64_t)__shfl_xor((unsigned long long)last1, d);
      const uint64_t o0 = (uint64_t)__shfl_xor((unsigned long long)last0, d);
      last1 = o1 > last1 ? o1 : last1;
      last0 = o0 > last0 ? o0 : last0;
    }
    if (lane == 0) {
      if (a.want[1] && has_last1) *a.sel[1].last = last1;
      if (a.want[0] && has_last0) *a.sel[0].last = last0;
    }
  }
}

// ----------------------------------------------------- select assembly
struct AsmArgs {
  uint32_t* mini;
  uint8_t* mini_width;
  uint8_t* long_width;
  uint32_t* long_flag;
  const uint64_t* sbstart;
  const uint64_t* sbend;
  const unsigned long long* last;
  uint64_t m;          // arguments
  uint64_t sb;
  uint32_t logn4;
};

__device__ __forceinline__ uint32_t hi_dev(uint64_t x) { return x ? 63 - __clzll(x) : 0; }

// thread per 4 mini slots (one 16-byte load and store): rebase the samples of
// mini superblocks to their start, clear those of long superblocks; the first
// thread of a superblock writes its widths and long flag
__global__ __launch_bounds__(256) void k_select_assemble(AsmArgs a) {
  const uint64_t idx4 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx4 >= a.sb * 16) return;
  const uint64_t k = idx4 >> 4;  // superblock (64 slots = 16 threads)
  const uint64_t st = a.sbstart[k];
  const uint64_t en = (k + 1 < a.sb) ? a.sbend[k] : (uint64_t)(*a.last) - 1;
  const bool lng = (en - st) > (uint64_t)a.logn4;
  const uint64_t nsamp = (a.m + 63) / 64;
  const uint64_t i0 = idx4 * 4;
  uint4 v = reinterpret_cast<const uint4*>(a.mini)[idx4];
  const uint32_t s32 = (uint32_t)st;
  v.x = (!lng && i0 + 0 < nsamp) ? v.x - s32 : 0u;
  v.y = (!lng && i0 + 1 < nsamp) ? v.y - s32 : 0u;
  v.z = (!lng && i0 + 2 < nsamp) ? v.z - s32 : 0u;
  v.w = (!lng && i0 + 3 < nsamp) ? v.w - s32 : 0u;
  reinterpret_cast<uint4*>(a.mini)[idx4] = v;
  if ((idx4 & 15) == 0) {
    a.mini_width[k] = lng ? 0 : (uint8_t)(hi_dev(en - st) + 1);
    a.long_width[k] = lng ? (uint8_t)(hi_dev(en) + 1) : 0;
    a.long_flag[k] = lng ? 1u : 0u;
  }
}

// also lists the long superblocks (list[ex[k]] = k) and copies their count
// ex[sb] to *nl_out
__global__ void k_long_index(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ ex,
                             int64_t* __restrict__ li, uint64_t sb, uint32_t* __restrict__ list,
                             uint32_t* __restrict__ nl_out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < sb) {
    li[k] = flag[k] ? (int64_t)ex[k] : -1;
    if (flag[k]) list[ex[k]] = (uint32_t)k;
  }
  if (k == 0) *nl_out = ex[sb];
}

// one block of LF_THREADS per long superblock (list[blockIdx.x]): enumerate the
// arguments in [a, e], one word per thread per round (a long superblock can
// hold dense stretches, and each thread walks its word's set bits serially)
constexpr int LF_THREADS = 1024;
__global__ __launch_bounds__(LF_THREADS) void k_long_fill(const uint64_t* __restrict__ words, uint64_t nbits,
                                                          int bit, const uint64_t* __restrict__ sbstart,
                                                          const uint64_t* __restrict__ sbend,
                                                          const unsigned long long* last,
                                                          const uint32_t* __restrict__ list, uint64_t sb,
                                                          uint64_t* __restrict__ pool) {
  constexpr int NWV = LF_THREADS / 64;
  __shared__ uint32_t s_wave[NWV];
  const uint64_t k = list[blockIdx.x];
  const uint64_t st = sbstart[k];
  const uint64_t en = (k + 1 < sb) ? sbend[k] : (uint64_t)(*last) - 1;
  uint64_t* out = pool + (uint64_t)blockIdx.x * 4096;
  uint32_t run = 0;
  const uint64_t w0 = st >> 6, w1 = en >> 6;
  for (uint64_t base = w0; base <= w1; base += LF_THREADS) {
    const uint64_t wi = base + threadIdx.x;
    uint64_t x = 0;
    if (wi <= w1) {
      x = bit ? words[wi] : ~words[wi];
      const uint64_t lo = wi * 64;
      if (lo < st) x &= ~0ull << (st - lo);
      if (lo + 63 > en) x &= (en - lo == 63) ? ~0ull : ((1ull << (en - lo + 1)) - 1);
      if (lo + 64 > nbits) x &= (nbits > lo) ? ((nbits - lo >= 64) ? ~0ull : ((1ull << (nbits - lo)) - 1)) : 0ull;
    }
    const uint32_t c = popc64(x);
    const int w = threadIdx.x >> 6, l = lane_id();
    const uint32_t incl = wave_incl_scan(c);
    if (l == 63) s_wave[w] = incl;
    __syncthreads();
    uint32_t off = 0, tt = 0;
#pragma unroll
    for (int q = 0; q < NWV; q++) { const uint32_t t = s_wave[q]; if (q < w) off += t; tt += t; }
    __syncthreads();
    uint32_t r = run + incl - c + off;
    while (x) {
      if (r < 4096) out[r] = wi * 64 + (uint64_t)__builtin_ctzll(x);
      r++;
      x &= x - 1;
    }
    run += tt;
  }
}

// ------------------------------------------------------------- queries
// rank1(i), or rank0(i) = i - rank1(i) for BIT = 0, from the rank1 directory
template <int BIT>
__global__ void k_rank9_query(const uint64_t* __restrict__ words, const uint64_t* __restrict__ D,
                              uint64_t nbits, const uint64_t* __restrict__ idx, uint64_t nq,
                              uint64_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint64_t i = idx[q];
  if (i > nbits) { out[q] = WTG_NONE; return; }
  const uint64_t b = i >> 9;
  uint64_t r = D[2 * b] + ((D[2 * b + 1] >> (63 - 9 * ((i & 511) >> 6))) & 0x1FF);
  if (i & 63) r += popc64(words[i >> 6] & ((1ull << (i & 63)) - 1));
  out[q] = BIT ? r : i - r;
}

__global__ void k_select_query(SelQ s, const uint64_t* __restrict__ iv, uint64_t nq, uint64_t m,
                               uint64_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint64_t i = iv[q];
  out[q] = (i >= 1 && i <= m) ? select_plain(s, i) : WTG_NONE;
}

// ================================================================== host
extern "C" uint64_t wtg_rank9_words(uint64_t nbits) {
  if (nbits == 0) return 2;
  const uint64_t cap = (nbits + 63) / 64 * 64;
  return ((cap >> 9) + 1) * 2;
}

namespace {
struct SelTmp {
  uint64_t* sbend = nullptr;
  unsigned long long* last = nullptr;
};

// device memory of a select structure: its own allocation, or the next
// 256-byte-aligned piece of the arena
int sel_take(wtg_ctx* ctx, wtg_arena* ar, void** p, uint64_t bytes) {
  if (!ar) return wtg_malloc(ctx, p, bytes);
  const size_t a = wtg_align256(bytes);
  WTG_REQUIRE(ar->used + a <= ar->size, "select arena exhausted");
  *p = ar->base + ar->used;
  ar->used += a;
  return WTG_OK;
}

int sel_alloc_worst(wtg_ctx* ctx, uint64_t nbits, int bit, wtg_select* s, wtg_arena* ar) {
  std::memset(s, 0, sizeof(*s));
  s->nbits = nbits;
  s->bit = bit;
  const uint64_t sbmax = wtg_ceil_div(nbits, 4096) + 1;
  WTG_TRY(sel_take(ctx, ar, (void**)&s->mini, sbmax * 64 * 4));
  WTG_TRY(sel_take(ctx, ar, (void**)&s->superblock, sbmax * 8));
  return WTG_OK;
}

// scratch of one select structure's finish: long flags, their exclusive scan,
// the scan's tile states and counter, and the list of long superblocks
struct FinScratch {
  uint32_t* flag;
  uint32_t* ex;
  uint64_t* st;
  uint32_t* ctr;
  uint32_t* list;
};
FinScratch fin_scratch(char* scratch, uint64_t sb) {
  FinScratch f;
  f.flag = (uint32_t*)scratch;
  f.ex = (uint32_t*)(scratch + wtg_align256(sb * 4));
  f.st = (uint64_t*)(scratch + wtg_align256(sb * 4) + wtg_align256((sb + 1) * 4));
  f.ctr = (uint32_t*)((char*)f.st + wtg_align256((wtg_ceil_div(sb, 4096) + 1) * 8));
  f.list = (uint32_t*)((char*)f.ctr + 256);
  return f;
}

// finish one select structure once the arguments are counted, in two halves
// so that both structures share one host read of their long-superblock
// counts: sel_count stages the widths, the long flags and their scan, and
// writes the number of long superblocks to *d_nl; sel_fill allocates and
// fills the long pool once that number is on the host
int sel_count(wtg_ctx* ctx, uint64_t nbits, wtg_select* s, uint64_t m, const uint64_t* sbend,
              const unsigned long long* last, char* scratch, uint32_t* d_nl, wtg_arena* ar) {
  const uint64_t cap = (nbits + 63) / 64 * 64;
  s->arg_cnt = m;
  s->logn = (uint32_t)wtg_hi_host(cap) + 1;
  s->logn4 = (s->logn * s->logn) * (s->logn * s->logn);
  s->sb = wtg_ceil_div(m, 4096);
  s->n_long = 0;
  const uint64_t sb = s->sb;
  WTG_TRY(sel_take(ctx, ar, (void**)&s->mini_width, sb ? sb : 1));
  WTG_TRY(sel_take(ctx, ar, (void**)&s->long_width, sb ? sb : 1));
  WTG_TRY(sel_take(ctx, ar, (void**)&s->long_index, (sb ? sb : 1) * 8));
  if (sb == 0) return wtg_fill_bytes_zero(ctx, d_nl, 4);
  const FinScratch f = fin_scratch(scratch, sb);
  uint32_t* flag = f.flag;
  uint32_t* ex = f.ex;
  uint64_t* st = f.st;
  uint32_t* ctr = f.ctr;
  WTG_TRY(wtg_fill_bytes_zero(ctx, st, (char*)ctr + 256 - (char*)st));
  AsmArgs aa{};
  aa.mini = s->mini; aa.mini_width = s->mini_width; aa.long_width = s->long_width;
  aa.long_flag = flag; aa.sbstart = s->superblock; aa.sbend = sbend; aa.last = last;
  aa.m = m; aa.sb = sb; aa.logn4 = s->logn4;
  WTG_LAUNCH(ctx, "select_assemble", k_select_assemble, dim3((unsigned)wtg_ceil_div(sb * 16, 256)),
             dim3(256), 0, aa);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "select_scan", k_scan_u32<uint32_t>, dim3((unsigned)wtg_ceil_div(sb, 4096)), dim3(256), 0,
             flag, ex, sb, st, ctr, ctx->d_flag);
  WTG_LAUNCH_CHECK();
  WTG_LAUNCH(ctx, "select_longidx", k_long_index, dim3((unsigned)wtg_ceil_div(sb, 256)), dim3(256), 0,
             flag, ex, s->long_index, sb, f.list, d_nl);
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}

int sel_fill(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, wtg_select* s, uint32_t nl,
             const uint64_t* sbend, const unsigned long long* last, const char* scratch, wtg_arena* ar) {
  const uint64_t sb = s->sb;
  s->n_long = nl;
  if (sb == 0 || nl == 0) return WTG_OK;
  const FinScratch f = fin_scratch(const_cast<char*>(scratch), sb);
  WTG_TRY(sel_take(ctx, ar, (void**)&s->long_pool, (uint64_t)nl * 4096 * 8));
  WTG_TRY(wtg_fill_u64(ctx, s->long_pool, (uint64_t)nl * 4096, 0));
  WTG_LAUNCH(ctx, "select_longfill", k_long_fill, dim3(nl), dim3(LF_THREADS), 0, d_words, nbits, s->bit,
             s->superblock, sbend, last, f.list, sb, s->long_pool);
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}

// workspace: counter | ones_total | last[2] | long counts[2] (256 bytes, first)
// | states | tile sums | tile prefixes | scan states | sbend[2] | finish scratch[2]
struct RsLayout {
  uint64_t W, S, tiles, wtiles, stiles, sbmax;
  size_t o_ctr, o_states, o_tsum, o_tpre, o_scst, o_end0, o_end1, fin_bytes, o_fin, total;
};
RsLayout rs_layout(uint64_t nbits) {
  RsLayout L{};
  const uint64_t cap = (nbits + 63) / 64 * 64;
  L.W = cap >> 6;
  L.S = cap >> 9;
  L.tiles = wtg_ceil_div(L.S + 1, 256);
  L.sbmax = wtg_ceil_div(nbits, 4096) + 1;
  // word tiles: every entry 0..S has its first word 8b inside them
  L.wtiles = wtg_ceil_div(8 * (L.S + 1), RS_TILE_WORDS);
  L.stiles = wtg_ceil_div(L.wtiles, 4096);
  size_t o = 0;
  L.o_ctr = o; o = wtg_align256(o + 256);
  L.o_states = o; o = wtg_align256(o + (L.tiles + 1) * 8);
  L.o_tsum = o; o = wtg_align256(o + (L.wtiles + 1) * 4);
  L.o_tpre = o; o = wtg_align256(o + (L.wtiles + 1) * 8);
  L.o_scst = o; o = wtg_align256(o + (L.stiles + 1) * 8);
  L.o_end0 = o; o = wtg_align256(o + L.sbmax * 8);
  L.o_end1 = o; o = wtg_align256(o + L.sbmax * 8);
  L.fin_bytes = wtg_align256(4 * wtg_align256((L.sbmax + 1) * 8) + 4096);
  L.o_fin = o; o = wtg_align256(o + 2 * L.fin_bytes);
  L.total = o;
  return L;
}
static_assert(WTG_RS_ONES_OFF == 64 && WTG_RS_NL_OFF == 192, "counter block layout");

}  // namespace

size_t wtg_rank_select_ws_bytes(uint64_t nbits) { return rs_layout(nbits).total; }
size_t wtg_select_arena_worst(uint64_t nbits) {
  const uint64_t sbmax = wtg_ceil_div(nbits, 4096) + 1;
  return wtg_align256(sbmax * 64 * 4) + wtg_align256(sbmax * 8);
}
size_t wtg_select_arena_count(uint64_t ones) {
  const uint64_t sb = std::max<uint64_t>(wtg_ceil_div(ones, 4096), 1);
  return 2 * wtg_align256(sb) + wtg_align256(sb * 8);
}
size_t wtg_select_arena_pool(uint32_t n_long) { return wtg_align256((uint64_t)n_long * 4096 * 8); }

int wtg_rank_select_begin(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir, bool build,
                          wtg_select* sel1, wtg_select* sel0, uint64_t known_ones, wtg_rs_pending* pend,
                          char* ws_ext, wtg_arena* ar) {
  *pend = wtg_rs_pending{};
  const RsLayout L = rs_layout(nbits);
  const uint64_t W = L.W, S = L.S, wtiles = L.wtiles, stiles = L.stiles;
  const size_t o_ctr = L.o_ctr, o_states = L.o_states, o_tsum = L.o_tsum, o_tpre = L.o_tpre;
  const size_t o_scst = L.o_scst, o_end0 = L.o_end0, o_end1 = L.o_end1, o_fin = L.o_fin;
  const size_t fin_bytes = L.fin_bytes;
  if (!ws_ext) WTG_TRY(wtg_ws_reserve(ctx, L.total));
  char* ws = ws_ext ? ws_ext : (char*)ctx->ws;
  WTG_TRY(wtg_fill_bytes_zero(ctx, ws, o_end0));
  uint32_t* ctr = (uint32_t*)(ws + o_ctr);
  unsigned long long* ones_total = (unsigned long long*)(ws + o_ctr + WTG_RS_ONES_OFF);
  unsigned long long* last = (unsigned long long*)(ws + o_ctr + 128);  // [0]=zeros, [1]=ones
  wtg_select* outs[2] = {sel0, sel1};
  RankSelArgs ra{};
  ra.words = d_words; ra.nbits = nbits; ra.W = W; ra.S = S; ra.D = d_dir;
  ra.states = (uint64_t*)(ws + o_states); ra.counter = ctr; ra.err = ctx->d_flag;
  ra.ones_total = ones_total;
  for (int t = 0; t < 2; t++) {
    ra.want[t] = outs[t] != nullptr && nbits > 0;
    if (ra.want[t]) {
      WTG_TRY(sel_alloc_worst(ctx, nbits, t, outs[t], ar));
      ra.sel[t].mini = outs[t]->mini;
      ra.sel[t].sbstart = outs[t]->superblock;
      ra.sel[t].sbend = (uint64_t*)(ws + (t ? o_end1 : o_end0));
      ra.sel[t].last = last + t;
    } else if (outs[t]) {
      std::memset(outs[t], 0, sizeof(wtg_select));
      outs[t]->bit = t;
    }
  }
  if (nbits == 0) {
    if (build) WTG_TRY(wtg_fill_u64(ctx, d_dir, 2, 0));
    for (int t = 0; t < 2; t++)
      if (outs[t]) {
        WTG_TRY(sel_take(ctx, ar, (void**)&outs[t]->mini_width, 1));
        WTG_TRY(sel_take(ctx, ar, (void**)&outs[t]->long_width, 1));
        WTG_TRY(sel_take(ctx, ar, (void**)&outs[t]->long_index, 8));
        outs[t]->logn = 1;
        outs[t]->logn4 = 1;
      }
    return WTG_OK;
  }
  {
    // ones per 512-word tile, their exclusive scan, then one pass writing the
    // directory entries and the select samples (a wave per tile)
    ra.tpre = (const uint64_t*)(ws + o_tpre);
    const unsigned g4 = (unsigned)wtg_ceil_div(wtiles, 4);
    WTG_LAUNCH(ctx, "rank_sums", k_rank_tile_sums, dim3(g4), dim3(256), 0, d_words, W, wtiles,
               (uint32_t*)(ws + o_tsum), ((uintptr_t)d_words & 15) == 0);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "rank_scan", k_scan_u32<uint64_t>, dim3((unsigned)stiles), dim3(256), 0,
               (const uint32_t*)(ws + o_tsum), (uint64_t*)(ws + o_tpre), wtiles, (uint64_t*)(ws + o_scst), ctr + 1,
               ctx->d_flag);
    WTG_LAUNCH_CHECK();
    WTG_LAUNCH(ctx, "rank_select", k_rank_entries, dim3(g4), dim3(256), 0, ra, wtiles, build,
               ((uintptr_t)d_words & 15) == 0);
  }
  WTG_LAUNCH_CHECK();
  if (!ra.want[0] && !ra.want[1]) return WTG_OK;  // rank9 alone: no count of ones needed
  pend->active = true;
  pend->d_words = d_words;
  pend->nbits = nbits;
  pend->d_nl = (const uint32_t*)(ws + o_ctr + WTG_RS_NL_OFF);
  pend->ar = ar;
  for (int t = 0; t < 2; t++) {
    pend->want[t] = ra.want[t];
    pend->outs[t] = ra.want[t] ? outs[t] : nullptr;
    pend->sbend[t] = ra.sel[t].sbend;
    pend->last[t] = last + t;
    pend->fin[t] = ws + o_fin + t * fin_bytes;
  }
  WTG_REQUIRE(known_ones != WTG_ONES_DEFER || (build && ws_ext), "rank/select: a deferred count needs its own workspace");
  if (known_ones == WTG_ONES_DEFER) return WTG_OK;
  uint64_t m1 = 0;
  if (build && known_ones != WTG_ONES_UNKNOWN) {
    m1 = known_ones;
  } else if (build) {
    unsigned long long t = 0;
    WTG_TRY(wtg_stage_d2h(ctx, &t, ones_total, 8));
    m1 = t;
  } else {
    // ones = D[2S] + ones in the last entry
    uint64_t e[2];
    WTG_CHECK_HIP(hipMemcpyAsync(e, d_dir + 2 * S, 16, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint64_t> tailw(8, 0);
    const uint64_t wb = 8 * S;
    const uint64_t nt = W > wb ? W - wb : 0;
    if (nt) WTG_CHECK_HIP(hipMemcpyAsync(tailw.data(), d_words + wb, nt * 8, hipMemcpyDeviceToHost, ctx->stream));
    WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    m1 = e[0];
    for (uint64_t k = 0; k < nt; k++) m1 += __builtin_popcountll(tailw[k]);
  }
  return wtg_rank_select_count(ctx, pend, m1);
}

int wtg_rank_select_count(wtg_ctx* ctx, wtg_rs_pending* pend, uint64_t ones) {
  if (!pend->active) return WTG_OK;
  const uint64_t m[2] = {pend->nbits - ones, ones};
  uint32_t* d_nl = const_cast<uint32_t*>(pend->d_nl);
  for (int t = 0; t < 2; t++)
    if (pend->want[t])
      WTG_TRY(sel_count(ctx, pend->nbits, pend->outs[t], m[t], pend->sbend[t], pend->last[t],
                        const_cast<char*>(pend->fin[t]), d_nl + t, pend->ar));
  return WTG_OK;
}

int wtg_rank_select_end(wtg_ctx* ctx, const wtg_rs_pending* pend, const uint32_t nl[2]) {
  if (!pend->active) return WTG_OK;
  for (int t = 0; t < 2; t++)
    if (pend->outs[t])
      WTG_TRY(sel_fill(ctx, pend->d_words, pend->nbits, pend->outs[t], nl[t], pend->sbend[t], pend->last[t],
                       pend->fin[t], pend->ar));
  return WTG_OK;
}

int wtg_rank_select_run(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir, bool build,
                        wtg_select* sel1, wtg_select* sel0) {
  wtg_rs_pending pend;
  WTG_TRY(wtg_rank_select_begin(ctx, d_words, nbits, d_dir, build, sel1, sel0, WTG_ONES_UNKNOWN, &pend));
  if (!pend.active) return WTG_OK;
  uint32_t nl[2] = {0, 0};
  WTG_TRY(wtg_stage_d2h(ctx, nl, pend.d_nl, 8));
  return wtg_rank_select_end(ctx, &pend, nl);
}

extern "C" int wtg_rank9_build(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir) {
  WTG_REQUIRE(ctx && d_dir && (nbits == 0 || d_words), "wtg_rank9_build: null argument");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(wtg_rank_select_run(ctx, d_words, nbits, d_dir, true, nullptr, nullptr));
  return wtg_finish(ctx);
}

extern "C" int wtg_rank9_select_build(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits,
                                      uint64_t* d_dir, wtg_select* sel1, wtg_select* sel0) {
  WTG_REQUIRE(ctx && d_dir && (nbits == 0 || d_words), "wtg_rank9_select_build: null argument");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(wtg_rank_select_run(ctx, d_words, nbits, d_dir, true, sel1, sel0));
  return wtg_finish(ctx);
}

extern "C" int wtg_select_build(wtg_ctx* ctx, const uint64_t* d_words, const uint64_t* d_dir,
                                uint64_t nbits, int bit, wtg_select* out) {
  WTG_REQUIRE(ctx && out && d_dir && (nbits == 0 || d_words), "wtg_select_build: null argument");
  WTG_REQUIRE(bit == 0 || bit == 1, "bit must be 0 or 1");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(wtg_rank_select_run(ctx, d_words, nbits, (uint64_t*)d_dir, false, bit ? out : nullptr,
                          bit ? nullptr : out));
  return wtg_finish(ctx);
}

namespace {
// argument words of the two-bit patterns `10` (pattern 10) and `01` (pattern 1):
// bit i = (bit i-1, bit i) == the pattern; the previous word's top bit is the
// carry, and before the vector sits the bit that rules out position 0
__global__ __launch_bounds__(256) void k_pattern_words(const uint64_t* __restrict__ w, uint64_t W, uint32_t pattern,
                                                       uint64_t tail_mask, uint64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < W; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = w[i];
    const uint64_t c = i ? w[i - 1] >> 63 : (pattern == 10 ? 0ull : 1ull);
    const uint64_t prev = (x << 1) | c;
    uint64_t r = pattern == 10 ? prev & ~x : ~prev & x;
    if (i + 1 == W) r &= tail_mask;
    out[i] = r;
  }
}
}  // namespace

extern "C" int wtg_pattern_words(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint32_t pattern, int clip,
                                 uint64_t* d_out) {
  WTG_REQUIRE(ctx && (nbits == 0 || (d_words && d_out)), "wtg_pattern_words: null argument");
  WTG_REQUIRE(pattern == 10 || pattern == 1, "wtg_pattern_words: pattern must be 10 or 1 (01)");
  if (nbits == 0) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  const uint64_t W = (nbits + 63) / 64;
  const uint64_t tail = (clip && (nbits & 63)) ? (1ull << (nbits & 63)) - 1 : ~0ull;
  const unsigned g = (unsigned)std::min<uint64_t>(wtg_ceil_div(W, 256), 65536);
  WTG_LAUNCH(ctx, "pattern_words", k_pattern_words, dim3(g), dim3(256), 0, d_words, W, pattern, tail, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

extern "C" int wtg_select_free(wtg_ctx* ctx, wtg_select* s) {
  WTG_REQUIRE(ctx && s, "null argument");
  wtg_free(ctx, s->superblock);
  wtg_free(ctx, s->mini);
  wtg_free(ctx, s->mini_width);
  wtg_free(ctx, s->long_index);
  wtg_free(ctx, s->long_width);
  wtg_free(ctx, s->long_pool);
  std::memset(s, 0, sizeof(*s));
  return WTG_OK;
}

extern "C" int wtg_rank9_batch(wtg_ctx* ctx, const uint64_t* d_words, const uint64_t* d_dir,
                               uint64_t nbits, const uint64_t* d_idx, uint64_t nq, uint64_t* d_out) {
  return wtg_rank9_batch_bit(ctx, d_words, d_dir, nbits, 1, d_idx, nq, d_out);
}

extern "C" int wtg_rank9_batch_bit(wtg_ctx* ctx, const uint64_t* d_words, const uint64_t* d_dir, uint64_t nbits,
                                   int bit, const uint64_t* d_idx, uint64_t nq, uint64_t* d_out) {
  WTG_REQUIRE(ctx && d_dir && (nq == 0 || (d_idx && d_out)), "wtg_rank9_batch: null argument");
  WTG_REQUIRE(bit == 0 || bit == 1, "wtg_rank9_batch_bit: bit must be 0 or 1");
  if (nq == 0) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  const dim3 g((unsigned)wtg_ceil_div(nq, 256));
  if (bit)
    WTG_LAUNCH(ctx, "rank9_query", k_rank9_query<1>, g, dim3(256), 0, d_words, d_dir, nbits, d_idx, nq, d_out);
  else
    WTG_LAUNCH(ctx, "rank9_query", k_rank9_query<0>, g, dim3(256), 0, d_words, d_dir, nbits, d_idx, nq, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

extern "C" int wtg_select_batch(wtg_ctx* ctx, const wtg_select* s, const uint64_t* d_words,
                                const uint64_t* d_i, uint64_t nq, uint64_t* d_out) {
  WTG_REQUIRE(ctx && s && (nq == 0 || (d_i && d_out)), "wtg_select_batch: null argument");
  if (nq == 0) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  SelQ q{s->superblock, s->mini, s->long_index, s->long_pool, d_words,
         (s->nbits + 63) / 64, s->bit};
  WTG_LAUNCH(ctx, "select_query", k_select_query, dim3((unsigned)wtg_ceil_div(nq, 256)), dim3(256), 0,
             q, d_i, nq, s->arg_cnt, d_out);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}

// ------------------------------------------------- allocation-free rank9 build
// wtg_rank9_build_noalloc runs the kernels of wtg_rank9_build (tile sums, their
// exclusive scan, the directory pass) with their scratch in one file-scope
// device array instead of the context's workspace, so the call makes no device
// allocation.  Layout for a vector of at most WTG_RANK9_NOALLOC_MAX_BITS bits:
// [scan counter + ones total: 256 B][scan look-back states][tile sums][tile prefixes].
namespace {
constexpr uint64_t na_words(uint64_t nbits) { return (nbits + 63) / 64; }
constexpr uint64_t na_wtiles(uint64_t nbits) {
  return (8 * ((na_words(nbits) * 64 >> 9) + 1) + RS_TILE_WORDS - 1) / RS_TILE_WORDS;
}
constexpr uint64_t na_stiles(uint64_t nbits) { return (na_wtiles(nbits) + 4095) / 4096; }
constexpr uint64_t na_al(uint64_t x) { return (x + 255) & ~uint64_t(255); }
constexpr uint64_t NA_O_STATES = 256;
constexpr uint64_t na_o_tsum(uint64_t nbits) { return NA_O_STATES + na_al((na_stiles(nbits) + 1) * 8); }
constexpr uint64_t na_o_tpre(uint64_t nbits) { return na_o_tsum(nbits) + na_al((na_wtiles(nbits) + 1) * 4); }
constexpr uint64_t na_bytes(uint64_t nbits) { return na_o_tpre(nbits) + na_al((na_wtiles(nbits) + 1) * 8); }
constexpr uint64_t NA_SCRATCH_BYTES = na_bytes(WTG_RANK9_NOALLOC_MAX_BITS);
static_assert(NA_SCRATCH_BYTES <= (13ull << 20), "scratch of the allocation-free rank9 build is 13 MiB at most");
}  // namespace

static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

// device addresses of `scratch`, one per device (each device holds its own copy)
static std::mutex g_na_mu;
static std::map<int, uint8_t*> g_na_ptr;

// Called by wtg_ctx_create with `device` current: resolves the device address
// of `scratch` once per device.
int wtg_noalloc_init_device(int device) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr.count(device)) return WTG_OK;
  void* ptr = nullptr;
  WTG_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) {
    wtg_set_error("hipGetSymbolAddress(scratch) returned a null address");
    return WTG_ERR_HIP;
  }
  g_na_ptr[device] = (uint8_t*)ptr;
  return WTG_OK;
}

__global__ void k_na_zero(uint64_t* __restrict__ d, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = 0;
}

extern "C" int wtg_rank9_build_noalloc(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir) {
  WTG_REQUIRE(ctx && d_dir && (nbits == 0 || d_words), "wtg_rank9_build_noalloc: null argument");
  if (nbits > WTG_RANK9_NOALLOC_MAX_BITS) {
    wtg_set_error("wtg_rank9_build_noalloc: nbits above WTG_RANK9_NOALLOC_MAX_BITS (2^35); use wtg_rank9_build");
    return WTG_ERR_UNSUPPORTED;
  }
  WTG_TRY(wtg_set_device(ctx));
  uint8_t* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    auto it = g_na_ptr.find(ctx->device);
    if (it != g_na_ptr.end()) ws = it->second;
  }
  if (!ws) {
    wtg_set_error("wtg_rank9_build_noalloc: scratch address of this device not resolved (context not created)");
    return WTG_ERR_HIP;
  }
  if (nbits == 0) {
    hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(64), 0, ctx->stream, d_dir, (uint64_t)2);
    WTG_LAUNCH_CHECK();
    return wtg_finish(ctx);
  }
  const uint64_t cap = (nbits + 63) / 64 * 64;
  const uint64_t W = cap >> 6, S = cap >> 9;
  const uint64_t wtiles = na_wtiles(nbits), stiles = na_stiles(nbits);
  uint32_t* ctr = (uint32_t*)ws;
  unsigned long long* ones_total = (unsigned long long*)(ws + 64);
  uint64_t* scst = (uint64_t*)(ws + NA_O_STATES);
  uint32_t* tsum = (uint32_t*)(ws + na_o_tsum(nbits));
  uint64_t* tpre = (uint64_t*)(ws + na_o_tpre(nbits));
  // the scan's counter and look-back states start at zero (the previous call's
  // values stay in the array)
  const uint64_t nz = na_o_tsum(nbits) / 8;
  hipLaunchKernelGGL(k_na_zero, dim3((unsigned)wtg_ceil_div(nz, 256)), dim3(256), 0, ctx->stream, (uint64_t*)ws, nz);
  WTG_LAUNCH_CHECK();
  const bool aligned = ((uintptr_t)d_words & 15) == 0;
  const unsigned g4 = (unsigned)wtg_ceil_div(wtiles, 4);
  hipLaunchKernelGGL(k_rank_tile_sums, dim3(g4), dim3(256), 0, ctx->stream, d_words, W, wtiles, tsum, aligned);
  WTG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_u32<uint64_t>, dim3((unsigned)stiles), dim3(256), 0, ctx->stream,
                     (const uint32_t*)tsum, tpre, wtiles, scst, ctr, ctx->d_flag);
  WTG_LAUNCH_CHECK();
  RankSelArgs ra{};
  ra.words = d_words; ra.nbits = nbits; ra.W = W; ra.S = S; ra.D = d_dir;
  ra.err = ctx->d_flag; ra.tpre = tpre; ra.ones_total = ones_total;
  ra.want[0] = ra.want[1] = 0;
  hipLaunchKernelGGL(k_rank_entries, dim3(g4), dim3(256), 0, ctx->stream, ra, wtiles, true, aligned);
  WTG_LAUNCH_CHECK();
  return wtg_finish(ctx);
}
