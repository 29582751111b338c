/*
 * wtg_oracle.h -- CPU oracle for the wavelet-tree / rank / select / FM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is the checker used by tests/, by the smoke() entry and
 * by bench.py's cpu_baseline leg.  Every function restates the behaviour of
 * the reference file:line given in its comment in plain C.
 */
#ifndef WTG_ORACLE_H
#define WTG_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NONE UINT64_MAX

typedef struct {
  uint64_t bitmapPtr, length, parent, leftChild, rightChild;
  uint64_t reached; /* 1 iff the build touched this slot */
} orc_node;

int orc_log2up(uint64_t x);        /* common/utils.h:54-60 */
int orc_hi(uint64_t x);            /* sdsl bits::hi, hi(0)=0 */
uint32_t orc_pbbs_hash(uint32_t a); /* common/utils.h:76-85 */

/* Dense alphabet remap, wavelet/levelWT/WTTime.C:80-94.  In place. Returns sigma. */
uint64_t orc_remap(void* s, int w, uint64_t n);

/* serialWT restatement, wavelet/serialWT/WT.C:32-103.  levels = L bits, the root
 * range is 2^L (levelWT passes L = max(1,log2Up(sigma))).  words: ceil(n*L/64)
 * zero-initialised by the callee.  nodes: 2^L slots or NULL. */
int orc_levels_serial(const void* s, int w, uint64_t n, uint32_t L,
                      uint64_t* words, orc_node* nodes);
/* Independent closed form: P_l = stable sort by x>>(L-l), bit = (P_l[i]>>(L-1-l))&1 */
int orc_levels_closed(const void* s, int w, uint64_t n, uint32_t L, uint64_t* words);
/* Access-reconstruction query, wavelet/levelWT/WTTime.C:37-76 */
uint64_t orc_wt_query(const orc_node* nodes, const uint64_t* words, uint64_t index,
                      uint64_t sigma);

/* rank_support_v<1,1> directory, rank/rank_support_v_par.hpp:78-113 */
uint64_t orc_rank9_entries(uint64_t nbits); /* number of uint64 words in D */
void orc_rank9_build(const uint64_t* words, uint64_t nbits, uint64_t* D);
uint64_t orc_rank9_rank(const uint64_t* D, const uint64_t* words, uint64_t idx); /* :131-142 */
/* orc_rank9_rank for idx[0..nq): ORC_NONE where idx > nbits */
void orc_rank9_rank_batch(const uint64_t* D, const uint64_t* words, uint64_t nbits, const uint64_t* idx,
                          uint64_t nq, uint64_t* out);

/* argument words of sdsl's two-bit patterns `10` (pattern 10) and `01` (pattern 1):
 * bit i = (bit i-1, bit i) == the pattern, position 0 never; clip: none at or past
 * nbits (select's count) instead of whole words (rank's block pass) */
void orc_pattern_words(const uint64_t* words, uint64_t nbits, uint32_t pattern, int clip, uint64_t* out);

/* select_support_mcl<b,1> content, select/select_support_mcl_par.hpp:227-276 (init_slow)
 * and :483-561 (init_fast) produce the same content. */
typedef struct {
  uint64_t nbits, arg_cnt, sb, n_long;
  uint32_t logn, logn4;
  uint64_t* superblock;  /* sb entries, width logn */
  uint32_t* mini;        /* sb*64 entries (long superblocks hold zeros here) */
  uint8_t*  mini_width;  /* sb: hi(e-a)+1, or 0 for a long superblock */
  int64_t*  long_index;  /* sb: index into long_pool, -1 for mini */
  uint8_t*  long_width;  /* sb: hi(e)+1, or 0 */
  uint64_t* long_pool;   /* n_long*4096 */
} orc_select;
/* pass 1: fills nbits, arg_cnt, sb, n_long, logn, logn4 only */
void orc_select_plan(const uint64_t* words, uint64_t nbits, int bit, orc_select* S);
/* pass 2: arrays must be allocated by the caller with the sizes of pass 1 */
void orc_select_fill(const uint64_t* words, uint64_t nbits, int bit, orc_select* S);
uint64_t orc_select_query(const orc_select* S, const uint64_t* words, int bit, uint64_t i); /* :564-611 */
uint64_t orc_select_brute(const uint64_t* words, uint64_t nbits, int bit, uint64_t i);
/* orc_select_query for idx[0..nq): out[q] = select(idx[q]), ORC_NONE outside [1, arg_cnt] */
void orc_select_query_batch(const orc_select* S, const uint64_t* words, int bit, const uint64_t* idx,
                            uint64_t nq, uint64_t* out);

/* wt_int queries over the concatenated tree, wt_int/wt_int_par.hpp:519-660 */
uint64_t orc_wtint_rank(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                        uint64_t i, uint64_t c);
uint64_t orc_wtint_access(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                          uint64_t i);
/* returns ORC_NONE when c does not occur i times (the reference throws) */
uint64_t orc_wtint_select(const uint64_t* D, const uint64_t* words,
                          const orc_select* S1, const orc_select* S0,
                          uint64_t n, uint32_t L, uint64_t i, uint64_t c);
void orc_wtint_lex_count(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                         uint64_t i, uint64_t j, uint64_t c, uint64_t* out3);
void orc_wtint_lex_smaller(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                           uint64_t i, uint64_t c, uint64_t* out2);
uint64_t orc_wtint_range_count(const uint64_t* D, const uint64_t* words, uint64_t n, uint32_t L,
                               uint64_t lb, uint64_t rb, uint64_t vlb, uint64_t vrb);

/* Huffman-shaped wavelet tree over a byte text (wt_huff/wt_huff_par.hpp:92-117,
 * wt_huff/wt_pc_par.hpp:193-257).  Node order is BFS from the root, child[0]
 * before child[1]. */
typedef struct {
  uint32_t n_nodes;           /* 2*sigma'-1 */
  uint32_t sigma;             /* symbols with freq>0 */
  uint64_t bv_size;           /* total bits */
  uint64_t bv_pos[511];
  uint64_t freq[511];
  int32_t child0[511], child1[511], parent[511];
  int32_t sym[511];           /* leaf symbol, -1 for inner */
  int32_t c_to_leaf[256];     /* -1 if absent */
  uint64_t path[256];         /* bit l = direction at depth l */
  uint32_t path_len[256];
} orc_huff;
void orc_huff_shape(const uint64_t* C /*256*/, orc_huff* H);
void orc_huff_bits(const uint8_t* t, uint64_t n, const orc_huff* H, uint64_t* words);
uint64_t orc_huff_rank(const orc_huff* H, const uint64_t* D, const uint64_t* words,
                       uint64_t i, uint32_t c);                     /* wt_pc_par.hpp:462-485 */
uint64_t orc_huff_access(const orc_huff* H, const uint64_t* D, const uint64_t* words,
                         uint64_t i);                                /* :430-448 */
uint64_t orc_huff_select(const orc_huff* H, const uint64_t* D,
                         const orc_select* S1, const orc_select* S0, const uint64_t* words,
                         uint64_t n, uint64_t i, uint32_t c);        /* :527-554 */

/* BWT from a suffix array, fm-index/par/WT.C:38-41 */
void orc_bwt_from_sa(const uint8_t* t, const int64_t* sa, uint64_t n, uint8_t* bwt);
/* Brute-force suffix array (comparison sort), small n only */
void orc_sa_brute(const uint8_t* t, uint64_t n, int64_t* sa);
/* Backward-search count over a byte BWT (occurrence counts by direct scan of a
 * sampled table); the reference has no count query. */
void orc_fm_count(const uint8_t* bwt, uint64_t n, const uint8_t* pats, uint32_t m,
                  uint64_t npat, uint64_t* out);
/* Count by binary search over the suffix array (independent check) */
void orc_sa_count(const uint8_t* t, const int64_t* sa, uint64_t n, const uint8_t* pats,
                  uint32_t m, uint64_t npat, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
