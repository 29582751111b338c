/*
 * wtg_oracle.c -- CPU oracle (plain C99) for the wavelet-tree / rank / select /
 * Huffman-WT / FM-count path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product library never links or calls it.
 *
 * Each function restates the reference behaviour at the cited file:line
 * (paths relative to the reference repository root).  How each one is pinned:
 *   orc_levels_serial  -- compiled reference serialWT/levelWT (oracle/_ref),
 *                         the independent closed form orc_levels_closed, the
 *                         access-reconstruction invariant.
 *   orc_rank9_*        -- brute-force prefix popcount for every index, plus the
 *                         field-by-field closed form in the tests.
 *   orc_select_*       -- brute-force position of the i-th argument; structure
 *                         content cross-checked against an independent Python
 *                         restatement of init_slow in the tests.
 *   orc_pattern_words  -- brute force on the bits: position i > 0 with bits
 *                         (i-1, i) equal to the pattern, every word boundary
 *                         and both clip modes (tests).
 *   orc_wtint_*        -- brute force on the text (count / i-th occurrence / value;
 *                         lex counts: values below / above c in [i, j); range
 *                         count: positions in [lb, rb] with values in [vlb, vrb]).
 *   orc_huff_*         -- answers by brute force on the text; the shape against an
 *                         independent (freq, id) heap merge on tie-heavy
 *                         alphabets, the BFS numbering and bv_pos against the
 *                         reference's own statement (wt_pc_par.hpp:134-135) and
 *                         the bits against the path rule (tests); sdsl's
 *                         byte_tree internals beyond that statement are
 *                         "parity unpinned" (sdsl absent).
 *   orc_fm_count       -- brute-force substring count and SA binary search.
 */
#include "wtg_oracle.h"
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ utils */
int orc_log2up(uint64_t x) { /* common/utils.h:54-60 */
  int a = 0;
  uint64_t b = x - 1;
  if (x == 0) return 0; /* reference: b = (T)-1 > 0 loops 64x for unsigned; unused */
  while (b > 0) { b >>= 1; a++; }
  return a;
}
int orc_hi(uint64_t x) { /* sdsl bits::hi: index of the most significant 1, hi(0)=0 */
  int r = 0;
  if (x == 0) return 0;
  while (x >>= 1) r++;
  return r;
}
uint32_t orc_pbbs_hash(uint32_t a) { /* common/utils.h:76-85 */
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}
static inline uint64_t rd(const void* s, int w, uint64_t i) {
  switch (w) {
    case 1: return ((const uint8_t*)s)[i];
    case 2: return ((const uint16_t*)s)[i];
    case 4: return ((const uint32_t*)s)[i];
    default: return ((const uint64_t*)s)[i];
  }
}
static inline void wr(void* s, int w, uint64_t i, uint64_t v) {
  switch (w) {
    case 1: ((uint8_t*)s)[i] = (uint8_t)v; break;
    case 2: ((uint16_t*)s)[i] = (uint16_t)v; break;
    case 4: ((uint32_t*)s)[i] = (uint32_t)v; break;
    default: ((uint64_t*)s)[i] = v; break;
  }
}
static inline int pc(uint64_t x) { return __builtin_popcountll(x); }
static inline int getbit(const uint64_t* w, uint64_t i) { return (int)((w[i >> 6] >> (i & 63)) & 1); }
static inline void setbit(uint64_t* w, uint64_t i) { w[i >> 6] |= 1ull << (i & 63); }

/* ------------------------------------------------------------------ remap */
uint64_t orc_remap(void* s, int w, uint64_t n) { /* wavelet/levelWT/WTTime.C:80-94 */
  uint64_t k = 0, i, sigma = 0;
  for (i = 0; i < n; i++) { uint64_t v = rd(s, w, i); if (v > k) k = v; }
  k += 1; /* k = 1 + max */
  uint64_t* A = (uint64_t*)calloc(k + 1, sizeof(uint64_t));
  for (i = 0; i < n; i++) A[rd(s, w, i)] = 1;
  for (i = 0; i < k + 1; i++) { uint64_t t = A[i]; A[i] = sigma; sigma += t; } /* exclusive plusScan */
  for (i = 0; i < n; i++) wr(s, w, i, A[rd(s, w, i)]);
  free(A);
  return sigma;
}

/* --------------------------------------------------------- serialWT levels */
typedef struct { uint64_t start, length, nodeID, range; } chunk;

int orc_levels_serial(const void* s, int w, uint64_t n, uint32_t L, uint64_t* words,
                      orc_node* nodes) {
  /* wavelet/serialWT/WT.C:32-103: per level, per node chunk in order, count the
   * zeros (:65-66), then one stable pass writing bits and the two halves (:67-73). */
  if (L == 0 || L > 63) return -1;
  uint64_t nw = (n * L + 63) / 64, i, j;
  memset(words, 0, nw * 8);
  void* s1 = malloc(n ? n * w : 1);
  void* s2 = malloc(n ? n * w : 1);
  uint64_t nslots = 1ull << L;
  uint64_t nchunk = (nslots / 2 + 1 < n + 2) ? nslots / 2 + 1 : n + 2; /* <= min(2^l, n) live */
  chunk* off = (chunk*)malloc(sizeof(chunk) * nchunk);
  chunk* off2 = (chunk*)malloc(sizeof(chunk) * nchunk);
  if (nodes) {
    memset(nodes, 0, sizeof(orc_node) * nslots);
    nodes[0].parent = ORC_NONE;
  }
  off[0].start = 0; off[0].length = n; off[0].nodeID = 0; off[0].range = 1ull << L;
  uint64_t nch = 1;
  for (uint32_t l = 0; l < L; l++) {
    uint64_t curr = (uint64_t)l * n;
    uint64_t mask = 1ull << (L - l - 1);
    const void* src = (l == 0) ? s : s1;
    uint64_t nnext = 0;
    for (j = 0; j < nch; j++) {
      uint64_t start = off[j].start, len = off[j].length, v = off[j].nodeID, range = off[j].range;
      if (nodes) { nodes[v].bitmapPtr = curr; nodes[v].length = len; nodes[v].reached = 1; }
      if (range <= 2) { /* :56-63 */
        for (i = start; i < start + len; i++) {
          if (rd(src, w, i) & mask) setbit(words, curr);
          curr++;
        }
        if (nodes) nodes[v].leftChild = nodes[v].rightChild = ORC_NONE;
      } else {
        uint64_t zeros = 0, left = 0, right;
        for (i = start; i < start + len; i++) if (!(rd(src, w, i) & mask)) zeros++;
        right = zeros;
        for (i = start; i < start + len; i++) {
          uint64_t x = rd(src, w, i);
          if (x & mask) { setbit(words, curr); wr(s2, w, start + right++, x); }
          else wr(s2, w, start + left++, x);
          curr++;
        }
        uint64_t ones = len - zeros;
        if (zeros) {
          if (nodes) { nodes[2 * v + 1].parent = v; nodes[v].leftChild = 2 * v + 1; }
          off2[nnext].start = start; off2[nnext].length = zeros;
          off2[nnext].nodeID = 2 * v + 1; off2[nnext].range = range >> 1; nnext++;
        } else if (nodes) nodes[v].leftChild = ORC_NONE;
        if (ones) {
          if (nodes) { nodes[2 * v + 2].parent = v; nodes[v].rightChild = 2 * v + 2; }
          off2[nnext].start = start + zeros; off2[nnext].length = ones;
          off2[nnext].nodeID = 2 * v + 2; off2[nnext].range = range >> 1; nnext++;
        } else if (nodes) nodes[v].rightChild = ORC_NONE;
      }
    }
    nch = nnext;
    { chunk* t = off; off = off2; off2 = t; }
    { void* t = s1; s1 = s2; s2 = t; }
  }
  free(s1); free(s2); free(off); free(off2);
  return 0;
}

/* -------------------------------------------------- closed-form levels */
static void stable_sort_idx_by_key(uint64_t* idx, const uint64_t* key, uint64_t n, uint64_t* tmp) {
  /* bottom-up merge sort of idx by key[idx]; stable */
  uint64_t width, i;
  uint64_t* a = idx;
  uint64_t* b = tmp;
  for (width = 1; width < n; width *= 2) {
    for (i = 0; i < n; i += 2 * width) {
      uint64_t lo = i, mid = i + width < n ? i + width : n, hi = i + 2 * width < n ? i + 2 * width : n;
      uint64_t p = lo, q = mid, k = lo;
      while (p < mid && q < hi) b[k++] = (key[a[q]] < key[a[p]]) ? a[q++] : a[p++];
      while (p < mid) b[k++] = a[p++];
      while (q < hi) b[k++] = a[q++];
    }
    uint64_t* t = a; a = b; b = t;
  }
  if (a != idx) memcpy(idx, a, n * 8);
}

int orc_levels_closed(const void* s, int w, uint64_t n, uint32_t L, uint64_t* words) {
  if (L == 0 || L > 63) return -1;
  uint64_t nw = (n * L + 63) / 64, i;
  memset(words, 0, nw * 8);
  uint64_t* idx = (uint64_t*)malloc((n ? n : 1) * 8);
  uint64_t* tmp = (uint64_t*)malloc((n ? n : 1) * 8);
  uint64_t* key = (uint64_t*)malloc((n ? n : 1) * 8);
  for (uint32_t l = 0; l < L; l++) {
    for (i = 0; i < n; i++) { idx[i] = i; key[i] = (l == 0) ? 0 : (rd(s, w, i) >> (L - l)); }
    stable_sort_idx_by_key(idx, key, n, tmp);
    for (i = 0; i < n; i++)
      if ((rd(s, w, idx[i]) >> (L - 1 - l)) & 1) setbit(words, (uint64_t)l * n + i);
  }
  free(idx); free(tmp); free(key);
  return 0;
}

/* access query, wavelet/levelWT/WTTime.C:37-76 */
static uint64_t naive_rank(const uint64_t* bm, uint64_t s, uint64_t f, int bit) {
  uint64_t r = 0, i;
  for (i = s; i < f; i++) if (getbit(bm, i) == bit) r++;
  return r;
}
uint64_t orc_wt_query(const orc_node* nodes, const uint64_t* bm, uint64_t index, uint64_t sigma) {
  orc_node curr = nodes[0];
  uint64_t mn = 0, mx = 1ull << orc_log2up(sigma);
  while (mx - mn > 2 && (curr.leftChild != ORC_NONE || curr.rightChild != ORC_NONE)) {
    uint64_t offset = curr.bitmapPtr;
    if (getbit(bm, index + offset)) {
      index -= naive_rank(bm, offset, offset + index, 0);
      mn += 1ull << (orc_log2up(mx - mn) - 1);
      curr = nodes[curr.rightChild];
    } else {
      index -= naive_rank(bm, offset, offset + index, 1);
      mx = mn + (1ull << (orc_log2up(mx - mn) - 1));
      curr = nodes[curr.leftChild];
    }
  }
  return getbit(bm, index + curr.bitmapPtr) ? mn + 1 : mn;
}

/* ------------------------------------------------------------------ rank9 */
uint64_t orc_rank9_entries(uint64_t nbits) {
  uint64_t cap = (nbits + 63) / 64 * 64;
  if (nbits == 0) return 2;
  return ((cap >> 9) + 1) * 2;
}
void orc_rank9_build(const uint64_t* words, uint64_t nbits, uint64_t* D) {
  /* rank/rank_support_v_par.hpp:78-113 (t_b=1: args_in_the_word = popcount,
   * carry unused).  The word after the last one (sdsl's padding word) reads 0. */
  uint64_t cap = (nbits + 63) / 64 * 64, W = cap >> 6, S = cap >> 9, b, i;
  if (nbits == 0) { D[0] = D[1] = 0; return; } /* :74-76 */
  for (b = 0; b <= S; b++) {
    uint64_t w0 = (8 * b < W) ? words[8 * b] : 0;
    uint64_t sum = (uint64_t)pc(w0), cnt = 0;
    uint64_t end = (8 * b + 8 < W) ? 8 * b + 8 : W;
    for (i = 8 * b + 1; i < end; i++) {
      cnt |= sum << (63 - 9 * (i & 7));
      sum += (uint64_t)pc(words[i]);
    }
    D[2 * b] = sum;
    if (i % 8 != 0) cnt |= sum << (63 - 9 * (i & 7)); /* :107-108 */
    D[2 * b + 1] = cnt;
  }
  uint64_t acc = 0; /* exclusive scan over D[2b], :112-113 */
  for (b = 0; b <= S; b++) { uint64_t t = D[2 * b]; D[2 * b] = acc; acc += t; }
}
uint64_t orc_rank9_rank(const uint64_t* D, const uint64_t* words, uint64_t idx) {
  /* rank/rank_support_v_par.hpp:131-142 */
  const uint64_t* p = D + ((idx >> 8) & ~1ull);
  uint64_t r = p[0] + ((p[1] >> (63 - 9 * ((idx & 0x1FF) >> 6))) & 0x1FF);
  if (idx & 63) r += (uint64_t)pc(words[idx >> 6] & ((1ull << (idx & 63)) - 1));
  return r;
}

void orc_rank9_rank_batch(const uint64_t* D, const uint64_t* words, uint64_t nbits, const uint64_t* idx,
                          uint64_t nq, uint64_t* out) {
  for (uint64_t q = 0; q < nq; q++) out[q] = idx[q] <= nbits ? orc_rank9_rank(D, words, idx[q]) : ORC_NONE;
}

/* ------------------------------------------------------ two-bit patterns */
void orc_pattern_words(const uint64_t* words, uint64_t nbits, uint32_t pattern, int clip, uint64_t* out) {
  /* sdsl-lite's rank_support_trait / select_support_trait <10,2> and <01,2>
   * (bits::map10 / map01 with the carry), as the reference's block pass calls
   * them (rank/rank_support_v_par.hpp:91-105: carry = the previous word's top
   * bit, init_carry before word 0) and as init_slow's found_arg scan
   * (select/select_support_mcl_par.hpp:240-245) counts them below size. */
  uint64_t W = (nbits + 63) / 64, k, carry = pattern == 10 ? 0 : 1;
  for (k = 0; k < W; k++) {
    uint64_t x = words[k], prev = (x << 1) | carry, r;
    r = pattern == 10 ? (x ^ prev) & ~x : (x ^ prev) & x; /* map10 / map01 */
    if (clip && k + 1 == W && (nbits & 63)) r &= (1ull << (nbits & 63)) - 1;
    out[k] = r;
    carry = x >> 63;
  }
}

/* ----------------------------------------------------------------- select */
static inline uint64_t argword(const uint64_t* words, uint64_t j, int bit) {
  return bit ? words[j] : ~words[j];
}
/* iterate the argument positions in [0,nbits) */
#define FOR_ARGS(words, nbits, bit, POS, BODY)                                     \
  do {                                                                              \
    uint64_t _W = ((nbits) + 63) / 64, _j;                                          \
    for (_j = 0; _j < _W; _j++) {                                                   \
      uint64_t _x = argword(words, _j, bit);                                        \
      if (_j == _W - 1 && ((nbits) & 63)) _x &= (1ull << ((nbits) & 63)) - 1;       \
      while (_x) {                                                                  \
        uint64_t POS = _j * 64 + (uint64_t)__builtin_ctzll(_x);                     \
        _x &= _x - 1;                                                               \
        BODY                                                                        \
      }                                                                             \
    }                                                                               \
  } while (0)

static void select_walk(const uint64_t* words, uint64_t nbits, int bit, orc_select* S, int fill) {
  /* select/select_support_mcl_par.hpp:227-276 (init_slow) */
  uint64_t m = 0;
  FOR_ARGS(words, nbits, bit, p_, { (void)p_; m++; });
  uint64_t cap = (nbits + 63) / 64 * 64;
  S->nbits = nbits;
  S->arg_cnt = m;
  S->logn = (uint32_t)orc_hi(cap) + 1;                          /* :626 */
  S->logn4 = (S->logn * S->logn) * (S->logn * S->logn);         /* :627-628, uint32 */
  S->sb = (m + 4095) / 4096;
  if (!fill) S->n_long = 0;
  if (m == 0) return;
  uint64_t* pos = (uint64_t*)malloc(4096 * 8);
  uint64_t cnt = 0, k = 0, nl = 0;
  FOR_ARGS(words, nbits, bit, p, {
    pos[cnt % 4096] = p;
    cnt++;
    if (cnt % 4096 == 0 || cnt == m) {
      uint64_t last = (cnt - 1) % 4096;
      uint64_t diff = pos[last] - pos[0];
      if (diff > S->logn4) { /* long superblock, :261-265 */
        if (fill) {
          S->superblock[k] = pos[0];
          S->mini_width[k] = 0;
          S->long_index[k] = (int64_t)nl;
          S->long_width[k] = (uint8_t)(orc_hi(pos[last]) + 1);
          for (uint64_t t = 0; t < 64; t++) S->mini[k * 64 + t] = 0;
          for (uint64_t jj = 0; jj < 4096; jj++)
            S->long_pool[nl * 4096 + jj] = (jj <= last) ? pos[jj] : 0;
        }
        nl++;
      } else { /* mini block, :266-270 */
        if (fill) {
          S->superblock[k] = pos[0];
          S->mini_width[k] = (uint8_t)(orc_hi(diff) + 1);
          S->long_index[k] = -1;
          S->long_width[k] = 0;
          for (uint64_t t = 0; t < 64; t++)
            S->mini[k * 64 + t] = (uint32_t)((64 * t <= last) ? pos[64 * t] - pos[0] : 0);
        }
      }
      k++;
    }
  });
  if (!fill) S->n_long = nl;
  free(pos);
}
void orc_select_plan(const uint64_t* words, uint64_t nbits, int bit, orc_select* S) {
  select_walk(words, nbits, bit, S, 0);
}
void orc_select_fill(const uint64_t* words, uint64_t nbits, int bit, orc_select* S) {
  select_walk(words, nbits, bit, S, 1);
}
static inline uint64_t sel_in_word(uint64_t x, uint64_t i) { /* 1-based i-th set bit */
  while (--i) x &= x - 1;
  return (uint64_t)__builtin_ctzll(x);
}
uint64_t orc_select_query(const orc_select* S, const uint64_t* words, int bit, uint64_t i) {
  /* select/select_support_mcl_par.hpp:564-611 */
  i = i - 1;
  uint64_t sb = i >> 12, off = i & 0xFFF;
  if (S->long_index[sb] >= 0) return S->long_pool[(uint64_t)S->long_index[sb] * 4096 + off];
  if ((off & 63) == 0) return S->superblock[sb] + S->mini[sb * 64 + (off >> 6)];
  i = i - (sb << 12) - ((off >> 6) << 6); /* 1..63 */
  uint64_t pos = S->superblock[sb] + S->mini[sb * 64 + (off >> 6)] + 1;
  uint64_t wp = pos >> 6, wo = pos & 63;
  uint64_t x = argword(words, wp, bit) >> wo;
  uint64_t args = (uint64_t)pc(x);
  if (args >= i) return (wp << 6) + wo + sel_in_word(x, i);
  uint64_t sum = args;
  for (;;) {
    wp++;
    x = argword(words, wp, bit);
    args = (uint64_t)pc(x);
    if (sum + args >= i) return (wp << 6) + sel_in_word(x, i - sum);
    sum += args;
  }
}
void orc_select_query_batch(const orc_select* S, const uint64_t* words, int bit, const uint64_t* idx,
                            uint64_t nq, uint64_t* out) {
  for (uint64_t q = 0; q < nq; q++)
    out[q] = (idx[q] >= 1 && idx[q] <= S->arg_cnt) ? orc_select_query(S, words, bit, idx[q]) : ORC_NONE;
}
uint64_t orc_select_brute(const uint64_t* words, uint64_t nbits, int bit, uint64_t i) {
  uint64_t c = 0, res = ORC_NONE;
  FOR_ARGS(words, nbits, bit, p, {
    if (++c == i) { res = p; goto done; }
  });
done:
  return res;
}

/* ---------------------------------------------------------------- wt_int */
#define R9(x) orc_rank9_rank(D, words, (x))
uint64_t orc_wtint_rank(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                        uint64_t i, uint64_t c) {
  /* wt_int/wt_int_par.hpp:553-577 */
  if (L < 64 && (1ull << L) <= c) return 0;
  uint64_t offset = 0, mask = 1ull << (L - 1), node_size = n;
  for (uint32_t k = 0; k < L && i; k++) {
    uint64_t o = R9(offset);
    uint64_t oi = R9(offset + i) - o;
    uint64_t oe = R9(offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; i = oi; }
    else { node_size = node_size - oe; i = i - oi; }
    offset += n;
    mask >>= 1;
  }
  return i;
}
uint64_t orc_wtint_access(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                          uint64_t i) {
  /* wt_int/wt_int_par.hpp:519-541 */
  uint64_t offset = 0, res = 0, node_size = n;
  for (uint32_t k = 0; k < L; k++) {
    res <<= 1;
    uint64_t o = R9(offset);
    uint64_t oi = R9(offset + i) - o;
    uint64_t oe = R9(offset + node_size) - o;
    if (getbit(words, offset + i)) { offset += node_size - oe; node_size = oe; i = oi; res |= 1; }
    else { node_size = node_size - oe; i = i - oi; }
    offset += n;
  }
  return res;
}
uint64_t orc_wtint_select(const uint64_t* D, const uint64_t* words, const orc_select* S1,
                          const orc_select* S0, uint64_t n, uint32_t L, uint64_t i, uint64_t c) {
  /* wt_int/wt_int_par.hpp:622-660 */
  uint64_t path_off[65], path_rank[65];
  uint64_t offset = 0, mask = 1ull << (L - 1), node_size = n;
  path_off[0] = path_rank[0] = 0;
  for (uint32_t k = 0; k < L && node_size; k++) {
    uint64_t o = R9(offset);
    path_rank[k] = o;
    uint64_t oe = R9(offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; }
    else node_size = node_size - oe;
    offset += n;
    path_off[k + 1] = offset;
    mask >>= 1;
  }
  if (node_size == 0 || node_size < i) return ORC_NONE; /* :644-646 throws */
  mask = 1;
  for (uint32_t k = L; k > 0; k--) {
    offset = path_off[k - 1];
    uint64_t o = path_rank[k - 1];
    if (c & mask) i = orc_select_query(S1, words, 1, o + i) - offset + 1;
    else i = orc_select_query(S0, words, 0, offset - o + i) - offset + 1;
    mask <<= 1;
  }
  return i - 1;
}

void orc_wtint_lex_count(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L, uint64_t i,
                         uint64_t j, uint64_t c, uint64_t* out3) {
  /* wt_int/wt_int_par.hpp:712-741: (rank(i,c), #smaller, #greater in [i,j)) */
  if (L < 64 && (1ull << L) <= c) { out3[0] = 0; out3[1] = j - i; out3[2] = 0; return; }
  uint64_t offset = 0, smaller = 0, greater = 0, mask = 1ull << (L - 1), node_size = n;
  for (uint32_t k = 0; k < L; k++) {
    uint64_t o = R9(offset);
    uint64_t oi = R9(offset + i) - o;
    uint64_t oj = R9(offset + j) - o;
    uint64_t oe = R9(offset + node_size) - o;
    if (c & mask) {
      offset += node_size - oe;
      node_size = oe;
      smaller += j - i - oj + oi;
      i = oi;
      j = oj;
    } else {
      node_size -= oe;
      greater += oj - oi;
      i -= oi;
      j -= oj;
    }
    offset += n;
    mask >>= 1;
  }
  out3[0] = i; out3[1] = smaller; out3[2] = greater;
}
void orc_wtint_lex_smaller(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L, uint64_t i,
                           uint64_t c, uint64_t* out2) {
  /* wt_int/wt_int_par.hpp:753-780: (rank(i,c), #smaller in [0,i)) */
  if (L < 64 && (1ull << L) <= c) { out2[0] = 0; out2[1] = i; return; }
  uint64_t offset = 0, result = 0, mask = 1ull << (L - 1), node_size = n;
  for (uint32_t k = 0; k < L && i; k++) {
    uint64_t o = R9(offset);
    uint64_t oi = R9(offset + i) - o;
    uint64_t oe = R9(offset + node_size) - o;
    if (c & mask) { offset += node_size - oe; node_size = oe; result += i - oi; i = oi; }
    else { node_size -= oe; i -= oi; }
    offset += n;
    mask >>= 1;
  }
  out2[0] = i; out2[1] = result;
}
/* wt_int/wt_int_par.hpp:816-870 (_range_search_2d, report = false): count of the
 * positions in [lb, rb] of the node at (level, ilb) whose values lie in [vlb, vrb] */
static void range2d(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L, uint64_t lb, uint64_t rb,
                    uint64_t vlb, uint64_t vrb, uint32_t level, uint64_t ilb, uint64_t node_size, uint64_t offset,
                    uint64_t* cnt) {
  if (lb > rb) return;
  if (level == L) { *cnt += rb - lb + 1; return; }
  uint64_t irb = ilb + (1ull << (L - level));
  uint64_t mid = (irb + ilb) >> 1;
  uint64_t o = R9(offset), olb = R9(offset + lb), orb = R9(offset + rb + 1), oe = R9(offset + node_size);
  uint64_t zo = offset - o, zlb = offset + lb - olb, zrb = offset + rb + 1 - orb, ze = offset + node_size - oe;
  if (vlb < mid && mid) {
    uint64_t nlb = zlb - zo, nrb = zrb - zo;
    if (nrb) range2d(D, words, n, L, nlb, nrb - 1, vlb, vrb < mid - 1 ? vrb : mid - 1, level + 1, ilb, ze - zo,
                     offset + n, cnt);
  }
  if (vrb >= mid) {
    uint64_t nlb = olb - o, nrb = orb - o;
    if (nrb) range2d(D, words, n, L, nlb, nrb - 1, mid > vlb ? mid : vlb, vrb, level + 1, mid, oe - o,
                     offset + n + (ze - zo), cnt);
  }
}
uint64_t orc_wtint_range_count(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L, uint64_t lb,
                               uint64_t rb, uint64_t vlb, uint64_t vrb) {
  /* wt_int/wt_int_par.hpp:801-814 */
  if (L < 64 && vrb > (1ull << L)) vrb = 1ull << L;
  if (vlb > vrb) return 0;
  uint64_t cnt = 0;
  range2d(D, words, n, L, lb, rb, vlb, vrb, 0, 0, n, 0, &cnt);
  return cnt;
}

/* --------------------------------------------------------------- Huffman */
void orc_huff_shape(const uint64_t* C, orc_huff* H) {
  /* wt_huff/wt_huff_par.hpp:92-117: leaves in symbol order with ids 0..s-1,
   * repeatedly pop (freq,id)-minimum twice, the parent gets the next id,
   * child[0]=first pop, child[1]=second pop.  Then BFS renumbering from the root
   * (sdsl byte_tree; parity unpinned: sdsl is not available). */
  uint64_t tf[511]; int32_t tc0[511], tc1[511], tsym[511];
  int alive[511];
  int nt = 0, i;
  for (i = 0; i < 256; i++)
    if (C[i] > 0) { tf[nt] = C[i]; tc0[nt] = tc1[nt] = -1; tsym[nt] = i; alive[nt] = 1; nt++; }
  int nleaves = nt;
  memset(H, 0, sizeof(*H));
  H->sigma = (uint32_t)nleaves;
  for (i = 0; i < 256; i++) { H->c_to_leaf[i] = -1; }
  if (nleaves == 0) { H->n_nodes = 0; return; }
  int remaining = nleaves;
  while (remaining > 1) {
    int v[2];
    for (int t = 0; t < 2; t++) { /* pop min by (freq,id) */
      int best = -1;
      for (i = 0; i < nt; i++)
        if (alive[i] && (best < 0 || tf[i] < tf[best] || (tf[i] == tf[best] && i < best))) best = i;
      alive[best] = 0;
      v[t] = best;
    }
    tf[nt] = tf[v[0]] + tf[v[1]]; tc0[nt] = v[0]; tc1[nt] = v[1]; tsym[nt] = -1; alive[nt] = 1;
    nt++;
    remaining--;
  }
  /* BFS from the root (last created node) */
  int map_q[511]; int head = 0, tail = 0;
  int newid[511];
  map_q[tail++] = nt - 1; newid[nt - 1] = 0;
  int cnt = 1;
  H->n_nodes = (uint32_t)nt;
  uint64_t bv = 0;
  H->parent[0] = -1;
  while (head < tail) {
    int old = map_q[head++];
    int id = newid[old];
    H->freq[id] = tf[old];
    H->sym[id] = tsym[old];
    H->bv_pos[id] = bv;
    if (tc0[old] >= 0) bv += tf[old];
    H->child0[id] = H->child1[id] = -1;
    if (tc0[old] >= 0) {
      int a = tc0[old], b = tc1[old];
      newid[a] = cnt++; H->child0[id] = newid[a]; H->parent[newid[a]] = id; map_q[tail++] = a;
      newid[b] = cnt++; H->child1[id] = newid[b]; H->parent[newid[b]] = id; map_q[tail++] = b;
    }
  }
  H->bv_size = bv;
  for (int v = 0; v < nt; v++) {
    if (H->sym[v] >= 0) {
      int c = H->sym[v];
      H->c_to_leaf[c] = v;
      uint64_t path = 0; uint32_t len = 0;
      int x = v;
      while (H->parent[x] >= 0) {
        int p = H->parent[x];
        path = (path << 1) | (H->child1[p] == x ? 1u : 0u);
        len++;
        x = p;
      }
      H->path[c] = path; H->path_len[c] = len;
    }
  }
}
void orc_huff_bits(const uint8_t* t, uint64_t n, const orc_huff* H, uint64_t* words) {
  /* wt_huff/wt_pc_par.hpp:193-257: the bit of an occurrence at a node of depth l
   * is bit l of its code path; each node keeps its occurrences in text order. */
  uint64_t* fill = (uint64_t*)calloc(511, 8);
  uint64_t nw = (H->bv_size + 63) / 64, i;
  memset(words, 0, (nw ? nw : 1) * 8);
  for (i = 0; i < n; i++) {
    int c = t[i];
    uint64_t p = H->path[c];
    int v = 0;
    for (uint32_t l = 0; l < H->path_len[c]; l++) {
      uint64_t pos = H->bv_pos[v] + fill[v]++;
      int d = (int)((p >> l) & 1);
      if (d) setbit(words, pos);
      v = d ? H->child1[v] : H->child0[v];
    }
  }
  free(fill);
}
uint64_t orc_huff_rank(const orc_huff* H, const uint64_t* D, const uint64_t* words,
                       uint64_t i, uint32_t c) {
  if (c > 255 || H->c_to_leaf[c] < 0) return 0;
  if (H->sigma == 1) return i;
  uint64_t p = H->path[c], res = i;
  uint32_t len = H->path_len[c];
  int v = 0;
  for (uint32_t l = 0; l < len && res; l++, p >>= 1) {
    uint64_t r = R9(H->bv_pos[v] + res) - R9(H->bv_pos[v]);
    if (p & 1) res = r; else res -= r;
    v = (p & 1) ? H->child1[v] : H->child0[v];
  }
  return res;
}
uint64_t orc_huff_access(const orc_huff* H, const uint64_t* D, const uint64_t* words, uint64_t i) {
  int v = 0;
  while (H->child0[v] >= 0) {
    uint64_t r = R9(H->bv_pos[v] + i) - R9(H->bv_pos[v]);
    if (getbit(words, H->bv_pos[v] + i)) { i = r; v = H->child1[v]; }
    else { i -= r; v = H->child0[v]; }
  }
  return (uint64_t)H->sym[v];
}
uint64_t orc_huff_select(const orc_huff* H, const uint64_t* D, const orc_select* S1,
                         const orc_select* S0, const uint64_t* words, uint64_t n,
                         uint64_t i, uint32_t c) {
  if (c > 255 || H->c_to_leaf[c] < 0) return n;
  if (H->sigma == 1) return (i - 1 < n) ? i - 1 : n;
  int v = H->c_to_leaf[c];
  uint64_t res = i - 1, p = H->path[c];
  uint32_t len = H->path_len[c];
  p <<= (64 - len);
  for (uint32_t l = 0; l < len; l++, p <<= 1) {
    v = H->parent[v];
    uint64_t bp = H->bv_pos[v], bpr = R9(bp);
    if ((p & 0x8000000000000000ull) == 0)
      res = orc_select_query(S0, words, 0, bp - bpr + res + 1) - bp;
    else
      res = orc_select_query(S1, words, 1, bpr + res + 1) - bp;
  }
  return res;
}

/* ------------------------------------------------------------- BWT / FM */
void orc_bwt_from_sa(const uint8_t* t, const int64_t* sa, uint64_t n, uint8_t* bwt) {
  /* fm-index/par/WT.C:38-41 */
  for (uint64_t i = 0; i < n; i++) bwt[i] = (sa[i] == 0) ? t[n - 1] : t[sa[i] - 1];
}
static const uint8_t* g_sa_text; static uint64_t g_sa_n;
static int sa_cmp(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  uint64_t lx = g_sa_n - (uint64_t)x, ly = g_sa_n - (uint64_t)y;
  uint64_t l = lx < ly ? lx : ly;
  int r = memcmp(g_sa_text + x, g_sa_text + y, l);
  if (r) return r;
  return (lx < ly) ? -1 : (lx > ly);
}
void orc_sa_brute(const uint8_t* t, uint64_t n, int64_t* sa) {
  for (uint64_t i = 0; i < n; i++) sa[i] = (int64_t)i;
  g_sa_text = t; g_sa_n = n;
  qsort(sa, n, sizeof(int64_t), sa_cmp);
}
void orc_fm_count(const uint8_t* bwt, uint64_t n, const uint8_t* pats, uint32_t m,
                  uint64_t npat, uint64_t* out) {
  /* backward search: [sp,ep) <- C[c] + rank_c([sp,ep)) from the last pattern char */
  uint64_t C[257] = {0}, i;
  const uint64_t step = 256;
  uint64_t nchk = n / step + 1;
  uint32_t* chk = (uint32_t*)calloc(nchk * 256, 4); /* counts of each byte in bwt[0, k*step) */
  uint32_t run[256] = {0};
  for (i = 0; i < n; i++) {
    if (i % step == 0) memcpy(chk + (i / step) * 256, run, sizeof(run));
    run[bwt[i]]++;
  }
  if (n % step == 0) memcpy(chk + (n / step) * 256, run, sizeof(run));
  for (int c = 0; c < 256; c++) C[c + 1] = C[c] + run[c];
  for (uint64_t q = 0; q < npat; q++) {
    const uint8_t* P = pats + q * m;
    uint64_t sp = 0, ep = n;
    for (int64_t j = (int64_t)m - 1; j >= 0 && sp < ep; j--) {
      int c = P[j];
      uint64_t a = 0, b = 0, k;
      for (k = (sp / step) * step; k < sp; k++) a += (bwt[k] == c);
      for (k = (ep / step) * step; k < ep; k++) b += (bwt[k] == c);
      sp = C[c] + chk[(sp / step) * 256 + c] + a;
      ep = C[c] + chk[(ep / step) * 256 + c] + b;
    }
    out[q] = (ep > sp) ? ep - sp : 0;
  }
  free(chk);
}
static int cmp_suffix_pat(const uint8_t* t, uint64_t n, int64_t s, const uint8_t* P, uint32_t m) {
  /* compare suffix s truncated to m with P; <0 if suffix < P */
  uint64_t l = n - (uint64_t)s;
  uint64_t k = l < m ? l : m;
  int r = memcmp(t + s, P, k);
  if (r) return r;
  return (l < m) ? -1 : 0;
}
void orc_sa_count(const uint8_t* t, const int64_t* sa, uint64_t n, const uint8_t* pats,
                  uint32_t m, uint64_t npat, uint64_t* out) {
  for (uint64_t q = 0; q < npat; q++) {
    const uint8_t* P = pats + q * m;
    uint64_t lo = 0, hi = n;
    while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (cmp_suffix_pat(t, n, sa[mid], P, m) < 0) lo = mid + 1; else hi = mid; }
    uint64_t a = lo; hi = n;
    while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (cmp_suffix_pat(t, n, sa[mid], P, m) <= 0) lo = mid + 1; else hi = mid; }
    out[q] = lo - a;
  }
}
