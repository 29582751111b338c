/*
 * wtg.h -- C ABI of the MI355X (gfx950) wavelet-tree / rank / select / FM path.
 *
 * Plain C types only: device buffers are `void*` / `uint64_t*` obtained from
 * wtg_malloc (or any hipMalloc'd pointer on the context's device), host buffers
 * are ordinary pointers.  Every function returns WTG_OK (0) or a negative
 * status; wtg_last_error() gives the message of the last failure on the
 * calling thread.  Nothing throws across this boundary.
 *
 * Bit layout everywhere: bit j of 64-bit word k is bit position 64k+j; bits
 * past the end of a vector are zero.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the reference repository jlabeit/wavelet-suffix-fm-index).
 */
#ifndef WTG_H
#define WTG_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WTG_OK 0
#define WTG_ERR_ARG (-1)          /* invalid argument (null pointer, size, width) */
#define WTG_ERR_HIP (-2)          /* a HIP runtime call failed */
#define WTG_ERR_NOMEM (-3)        /* device allocation failed */
#define WTG_ERR_UNSUPPORTED (-4)  /* valid input outside what the library handles */
#define WTG_ERR_ABSENT (-5)       /* select(i,c): c does not occur i times */
#define WTG_ERR_DEVICE (-6)       /* a kernel's bounded wait expired (device-side error flag) */

#define WTG_NONE UINT64_MAX       /* absent node / child (reference: UINT_T_MAX) */

typedef struct wtg_ctx wtg_ctx;

/* ------------------------------------------------------------ context */
const char* wtg_version(void);
const char* wtg_last_error(void);
int wtg_device_count(int* out);
/* One context = one device + one HIP stream + a growable device workspace.
 * Contexts are independent; a context is used by one host thread at a time. */
int wtg_ctx_create(int device, wtg_ctx** out);
int wtg_ctx_destroy(wtg_ctx* ctx);
/* Wait for all work queued on the context's stream; reports device-side errors. */
int wtg_ctx_sync(wtg_ctx* ctx);
/* 1 (default): every entry point returns after its work finished.
 * 0: entry points only enqueue; errors surface at the next wtg_ctx_sync. */
int wtg_ctx_set_blocking(wtg_ctx* ctx, int blocking);
/* How the level-wise trees of this context are built (same output either way):
 * 0 (default): multi-level direct passes, up to 8 levels per pass over the keys;
 * 1: one stable split pass per level (the reference's own schedule,
 *    wavelet/levelWT/WT.C:43-227).  The environment variable WTG_LEVEL_SPLIT=1
 *    selects 1 for every context.  No counterpart in the reference, which has
 *    one schedule. */
int wtg_ctx_set_level_method(wtg_ctx* ctx, int method);
/* The HIP stream (hipStream_t) the context launches on, as an opaque pointer. */
void* wtg_ctx_stream(wtg_ctx* ctx);

int wtg_malloc(wtg_ctx* ctx, void** dptr, uint64_t bytes);
int wtg_free(wtg_ctx* ctx, void* dptr);
int wtg_memcpy_h2d(wtg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int wtg_memcpy_d2h(wtg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int wtg_memcpy_d2d(wtg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int wtg_memset_zero(wtg_ctx* ctx, void* dst, uint64_t bytes); /* fill kernel, no hipMemset */

/* Stream timing with HIP events on the context's stream. */
int wtg_timer_start(wtg_ctx* ctx);
int wtg_timer_stop(wtg_ctx* ctx, float* ms);
/* Per-kernel timing: when enabled, each launch of a library kernel is
 * bracketed by HIP events on the context's stream.  wtg_kernel_time reports the
 * summed milliseconds and number of launches of kernel `name` since the last
 * reset (it synchronizes the stream). */
int wtg_kernel_timing(wtg_ctx* ctx, int enable);
int wtg_kernel_time(wtg_ctx* ctx, const char* name, double* total_ms, uint64_t* launches);
int wtg_kernel_timing_reset(wtg_ctx* ctx);
/* Restrict per-kernel timing to launches whose name starts with name_prefix
 * (NULL or "" = every launch), so the other launches carry no events. */
int wtg_kernel_timing_only(wtg_ctx* ctx, const char* name_prefix);

/* ------------------------------------------------ (a) level-wise wavelet tree */
typedef struct {
  uint64_t bitmapPtr, length, parent, leftChild, rightChild;
} wtg_node;

/* Number of 64-bit words of an n*levels-bit tree. */
uint64_t wtg_levels_words(uint64_t n, uint32_t levels);
/* Level-wise stable bit split.  Replaces
 *   pair<WTnode*,long*> WT(symbol* s, uintT n, uintT sigma)
 *   (wavelet/levelWT/WT.h:43, wavelet/levelWT/WT.C:43-227; same output as
 *   serialWT/ddWT/recWT WT.C) with levels = max(1, log2Up(sigma)), and the tree
 *   bit_vector of sdsl::wt_int (wt_int/wt_int_par.hpp:166-244,368-406) with
 *   levels = max_level.
 * d_sym: n device symbols of sym_bytes (1, 2, 4 or 8), each < 2^levels.
 * d_words: wtg_levels_words(n, levels) device words (fully overwritten).
 * d_nodes: NULL, or node_slots device records indexed by heap id (root 0,
 *   children 2v+1 / 2v+2); reached nodes as in the reference, other slots
 *   {0, 0, WTG_NONE, WTG_NONE, WTG_NONE}.
 * Limits: 1 <= levels <= 64, n < 2^32; d_nodes only up to 24 levels (above, pass
 * NULL: WTG_ERR_UNSUPPORTED otherwise).  Above 24 levels the top 24 come from the
 * multi-level passes over the keys' top 24 bits and the rest level by level, with
 * nodes numbered by their rank within the level (linear work per level). */
int wtg_levels_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n,
                     uint32_t levels, uint64_t* d_words, wtg_node* d_nodes,
                     uint64_t node_slots);
/* Dense alphabet remap of the driver (wavelet/levelWT/WTTime.C:80-94), in place
 * on the device; *sigma = number of distinct values. */
int wtg_remap(wtg_ctx* ctx, void* d_sym, int sym_bytes, uint64_t n, uint64_t* sigma);
/* Host-buffer convenience used by the C++ WT() drop-in: copies h_sym to the
 * device, builds with levels = max(1, log2Up(sigma)), copies back
 * wtg_levels_words(n, levels) words and 2*sigma nodes (h_nodes may be NULL). */
int wtg_wt_host(wtg_ctx* ctx, const void* h_sym, int sym_bytes, uint64_t n, uint64_t sigma,
                uint64_t* h_words, wtg_node* h_nodes);

/* ------------------------------------------------ (b) rank directory */
/* Number of 64-bit words of the rank_support_v<1,1> directory:
 * 2*((capacity>>9)+1), capacity = 64*ceil(nbits/64) (2 for nbits = 0). */
uint64_t wtg_rank9_words(uint64_t nbits);
/* Builds the interleaved superblock/block directory of
 *   sdsl::rank_support_v<1,1>(const bit_vector*)  (rank/rank_support_v_par.hpp:69-123). */
int wtg_rank9_build(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir);
/* rank(idx) for a batch of idx (0 <= idx <= nbits; larger idx -> WTG_NONE)
 * (rank_support_v_par.hpp:131-142). */
int wtg_rank9_batch(wtg_ctx* ctx, const uint64_t* d_words, const uint64_t* d_dir,
                    uint64_t nbits, const uint64_t* d_idx, uint64_t nq, uint64_t* d_out);
/* The same for rank_support_v<bit,1> (bit 0: rank0(idx) = idx - rank1(idx), the
 * count rank_support_v<0,1>::rank returns, rank_support_v_par.hpp:131-142), from
 * the rank1 directory d_dir. */
int wtg_rank9_batch_bit(wtg_ctx* ctx, const uint64_t* d_words, const uint64_t* d_dir, uint64_t nbits,
                        int bit, const uint64_t* d_idx, uint64_t nq, uint64_t* d_out);

/* Allocation-free variant of wtg_rank9_build.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * directory, same kernels; their scratch (tile sums, their scan and its
 * look-back state, 13 MiB at most) lives in one static device array per GPU
 * instead of the context's workspace, so the call makes no hipMalloc / hipFree.
 * Limit: nbits <= WTG_RANK9_NOALLOC_MAX_BITS (larger -> WTG_ERR_UNSUPPORTED).
 * Every call on one GPU shares that array: calls that overlap on one GPU (two
 * contexts, or two host threads) race and must be serialized by the caller. */
#define WTG_RANK9_NOALLOC_MAX_BITS (1ull << 35)
int wtg_rank9_build_noalloc(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint64_t* d_dir);

/* ------------------------------------------------ (b') select samples */
/* Content of sdsl::select_support_mcl<bit,1>  (select/select_support_mcl_par.hpp:72-117,
 * :227-276 init_slow, :483-561 init_fast).  All arrays are device arrays owned
 * by the structure (release with wtg_select_free). */
typedef struct {
  uint64_t nbits, arg_cnt, sb, n_long;
  uint32_t logn, logn4;
  int32_t bit;
  int32_t _pad;
  uint64_t* superblock;   /* sb: position of argument 4096k (width logn) */
  uint32_t* mini;         /* sb*64: pos(4096k+64t) - superblock[k]; 0 past the end / for long */
  uint8_t* mini_width;    /* sb: hi(end-start)+1, 0 for a long superblock */
  int64_t* long_index;    /* sb: row of long_pool, -1 for a mini superblock */
  uint8_t* long_width;    /* sb: hi(end)+1, 0 for a mini superblock */
  uint64_t* long_pool;    /* n_long*4096 argument positions (0 past the end), NULL if none */
} wtg_select;
/* Builds select1 (bit=1) or select0 (bit=0) samples.  d_dir: the rank9
 * directory of the same words (wtg_rank9_build), used to locate samples. */
int wtg_select_build(wtg_ctx* ctx, const uint64_t* d_words, const uint64_t* d_dir,
                     uint64_t nbits, int bit, wtg_select* out);
/* rank9 directory + select1 + select0 in one pass over the words. */
int wtg_rank9_select_build(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits,
                           uint64_t* d_dir, wtg_select* sel1, wtg_select* sel0);
int wtg_select_free(wtg_ctx* ctx, wtg_select* s);
/* select(i), 1 <= i <= arg_cnt, the sampled scan of select_support_mcl_par.hpp:564-611. */
int wtg_select_batch(wtg_ctx* ctx, const wtg_select* s, const uint64_t* d_words,
                     const uint64_t* d_i, uint64_t nq, uint64_t* d_out);

/* ------------------------------------------------ (b'') two-bit patterns */
/* The argument words of sdsl's two-bit patterns, for rank_support_v<t_b,2>
 * (rank/rank_support_v_par.hpp:54-58, :84-105) and select_support_mcl<t_b,2>
 * (select/select_support_mcl_par.hpp:72-77, :227-276): bit i of the output is 1
 * when bits i-1 and i of the vector read `10` (pattern = 10) or `01`
 * (pattern = 1, the template argument 01); before bit 0 sits the bit that
 * makes position 0 no occurrence (0 for `10`, 1 for `01`), each word's carry
 * is the previous word's top bit.  The rank9 directory and the select content
 * of the output words are those of the pattern: rank_support_v counts whole
 * words (clip = 0: an occurrence in the padding bits past nbits counts, as the
 * reference's block pass does), select_support_mcl counts positions below
 * nbits (clip = 1).  d_out: ceil(nbits/64) words. */
int wtg_pattern_words(wtg_ctx* ctx, const uint64_t* d_words, uint64_t nbits, uint32_t pattern, int clip,
                      uint64_t* d_out);

/* ------------------------------------------------ wt_int (integer wavelet tree) */
typedef struct wtg_wtint wtg_wtint;
/* sdsl::wt_int<>(int_vector<w>& buf, size, max_level)  (wt_int/wt_int_par.hpp:368-406):
 * max_level = 0 -> hi(max(1, max value))+1.  Builds the tree, rank9, select1, select0. */
int wtg_wtint_build(wtg_ctx* ctx, const void* d_sym, int sym_bytes, uint64_t n,
                    uint32_t max_level, wtg_wtint** out);
int wtg_wtint_destroy(wtg_ctx* ctx, wtg_wtint* wt);
/* A wt_int over an existing tree (d_tree: wtg_levels_words(n, levels) device
 * words, copied): builds rank_support_v + select_support_mcl<1> + <0> over it,
 * as wt_int::load (wt_int/wt_int_par.hpp:901-911) restores them; levels <= 64. */
int wtg_wtint_from_tree(wtg_ctx* ctx, const uint64_t* d_tree, uint64_t n, uint32_t levels, uint64_t sigma,
                        wtg_wtint** out);
/* size, max_level, sigma (#distinct values), device tree words and rank directory */
int wtg_wtint_info(const wtg_wtint* wt, uint64_t* n, uint32_t* levels, uint64_t* sigma,
                   const uint64_t** d_words, const uint64_t** d_dir);
const wtg_select* wtg_wtint_select1(const wtg_wtint* wt);
const wtg_select* wtg_wtint_select0(const wtg_wtint* wt);
/* rank(i,c) = #c in [0,i)  (wt_int_par.hpp:553-577) */
int wtg_wtint_rank_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                         const uint64_t* d_c, uint64_t nq, uint64_t* d_out);
/* operator[](i)  (wt_int_par.hpp:519-541) */
int wtg_wtint_access_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                           uint64_t nq, uint64_t* d_out);
/* select(i,c): position of the i-th c  (wt_int_par.hpp:622-660).  Where c does not
 * occur i times the output is WTG_NONE (the reference throws std::logic_error). */
int wtg_wtint_select_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                           const uint64_t* d_c, uint64_t nq, uint64_t* d_out);
/* inverse_select(i) = (rank(i, wt[i]), wt[i])  (wt_int_par.hpp:588-611) */
int wtg_wtint_inverse_select_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                   uint64_t nq, uint64_t* d_rank, uint64_t* d_sym);
/* lex_count(i,j,c) = (rank(i,c), #values < c in [i,j), #values > c in [i,j))
 * (wt_int_par.hpp:712-741; c >= 2^L gives (0, j-i, 0)).  i > j or j > size: WTG_NONE. */
int wtg_wtint_lex_count_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                              const uint64_t* d_j, const uint64_t* d_c, uint64_t nq,
                              uint64_t* d_rank, uint64_t* d_smaller, uint64_t* d_greater);
/* lex_smaller_count(i,c) = (rank(i,c), #values < c in [0,i))  (wt_int_par.hpp:753-780).
 * i > size: WTG_NONE. */
int wtg_wtint_lex_smaller_count_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_i,
                                      const uint64_t* d_c, uint64_t nq, uint64_t* d_rank,
                                      uint64_t* d_smaller);
/* range_search_2d(lb, rb, vlb, vrb, report=false).first: #positions in [lb,rb] with
 * values in [vlb,vrb]  (wt_int_par.hpp:801-870).  As the reference: vrb is clamped
 * to 2^L, vlb > vrb or lb > rb gives 0, and vlb = 2^L counts the value 2^L-1 (its
 * leaf test does not compare vlb).  rb >= size: WTG_NONE. */
int wtg_wtint_range_count_batch(wtg_ctx* ctx, const wtg_wtint* wt, const uint64_t* d_lb,
                                const uint64_t* d_rb, const uint64_t* d_vlb,
                                const uint64_t* d_vrb, uint64_t nq, uint64_t* d_out);

/* ------------------------------------- wt_huff (Huffman-shaped WT over bytes) */
/* Tree shape of sdsl::wt_huff<> (wt_huff/wt_huff_par.hpp:92-117: Huffman merge of
 * (freq, id) minima, child[0] = first pop; nodes numbered breadth-first from the
 * root, child[0] before child[1]).  Inner nodes hold bits at bv_pos[v] (BFS order);
 * a symbol's code bit l is the branch taken at depth l. */
#define WTG_HUFF_MAX_NODES 511
typedef struct {
  uint64_t n;                     /* text length */
  uint64_t bv_size;               /* total bits over all inner nodes */
  uint32_t sigma;                 /* symbols with nonzero count */
  uint32_t n_nodes;               /* 2*sigma-1 (0 for an empty text) */
  uint32_t n_inner;               /* sigma-1 */
  uint32_t max_depth;             /* longest code */
  uint64_t C[257];                /* C[c] = #bytes < c in the text */
  uint64_t bv_pos[WTG_HUFF_MAX_NODES];
  uint64_t bv_pos_rank[WTG_HUFF_MAX_NODES]; /* rank1 of the tree bits at bv_pos (inner) */
  int32_t child[WTG_HUFF_MAX_NODES][2];     /* -1 at leaves */
  int32_t parent[WTG_HUFF_MAX_NODES];       /* -1 at the root */
  int32_t sym[WTG_HUFF_MAX_NODES];          /* leaf symbol, -1 at inner nodes */
  int32_t c_to_leaf[256];                   /* -1 for absent symbols */
  uint64_t path[256];                       /* code bits, bit l = branch at depth l */
  uint32_t path_len[256];
} wtg_huff_shape;

typedef struct wtg_huff wtg_huff;
/* sdsl::wt_huff<>(int_vector<8>& text, size)  (wt_huff/wt_pc_par.hpp:321-361, tree
 * bits by build_recursive :193-257): counts, Huffman shape, all tree bits, rank9,
 * select1, select0.  d_text: n bytes in device memory. */
int wtg_huff_build(wtg_ctx* ctx, const uint8_t* d_text, uint64_t n, wtg_huff** out);
int wtg_huff_destroy(wtg_ctx* ctx, wtg_huff* wt);
/* A wt_huff from its byte counts (counts[256]) and existing tree bits (d_words:
 * nbits bits in ceil(nbits/64) device words, copied), as wt_pc::load restores
 * one (wt_huff/wt_pc_par.hpp:766-775): the Huffman shape of the counts, then
 * rank9 + select1 + select0 and the query blocks over the bits.  nbits must be
 * the bit count of that shape (else WTG_ERR_ARG). */
int wtg_huff_from_bits(wtg_ctx* ctx, const uint64_t* counts, const uint64_t* d_words, uint64_t nbits,
                       wtg_huff** out);
int wtg_huff_get_shape(const wtg_huff* wt, wtg_huff_shape* out);
/* device tree words (ceil(bv_size/64)+1, last = 0) and rank9 directory */
int wtg_huff_data(const wtg_huff* wt, const uint64_t** d_words, const uint64_t** d_dir);
const wtg_select* wtg_huff_select1(const wtg_huff* wt);
const wtg_select* wtg_huff_select0(const wtg_huff* wt);
/* rank(i,c) = #c in [0,i)  (wt_pc_par.hpp:462-485); i > n -> WTG_NONE */
int wtg_huff_rank_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint64_t* d_i,
                        const uint32_t* d_c, uint64_t nq, uint64_t* d_out);
/* operator[](i) (:430-448) and inverse_select(i) (:498-516); i >= n -> WTG_NONE.
 * d_rank may be null (access only). */
int wtg_huff_access_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint64_t* d_i,
                          uint64_t nq, uint64_t* d_sym, uint64_t* d_rank);
/* select(i,c) = position of the i-th c (:527-554); WTG_NONE when c does not occur
 * i times or i = 0 (the reference asserts 1 <= i <= rank(n, c)) */
int wtg_huff_select_batch(wtg_ctx* ctx, const wtg_huff* wt, const uint64_t* d_i,
                          const uint32_t* d_c, uint64_t nq, uint64_t* d_out);

/* ------------------------------------------------ FM-index over a BWT */
/* Suffix array of a byte text (the reference calls divsufsort,
 * fm-index/par/WT.C:35-36): suffixes in lexicographic order, a proper prefix
 * before its extensions.  Host memory, n < 2^40. */
int wtg_suffix_array(const uint8_t* text, uint64_t n, int64_t* sa);
/* The same suffix array on the GPU (device text and output, n < 2^31): prefix
 * doubling with device radix sorts, in place of the reference's divsufsort call
 * (fm-index/par/WT.C:35-36). */
int wtg_suffix_array_gpu(wtg_ctx* ctx, const uint8_t* d_text, uint64_t n, int64_t* d_sa);
/* wtg_bwt_from_sa on the GPU (device buffers); an entry of d_sa outside [0, n)
 * is reported as WTG_ERR_ARG. */
int wtg_bwt_from_sa_gpu(wtg_ctx* ctx, const uint8_t* d_text, const int64_t* d_sa, uint64_t n, uint8_t* d_bwt);
/* bwt[i] = text[sa[i]-1], or text[n-1] where sa[i] = 0  (fm-index/par/WT.C:38-41) */
int wtg_bwt_from_sa(const uint8_t* text, const int64_t* sa, uint64_t n, uint8_t* bwt);
/* Backward-search count of npat patterns of m bytes each (d_pats, pattern q at
 * q*m) over the wt_huff of a BWT (the index fm-index/par/WT.C:33-45 builds; the
 * reference stops after construction).  out[q] = occurrences of pattern q in
 * the text, exact when the text ends with a unique smallest byte. */
int wtg_fm_count_batch(wtg_ctx* ctx, const wtg_huff* bwt_wt, const uint8_t* d_pats,
                       uint32_t m, uint64_t npat, uint64_t* d_out);

/* ------------------------------------------------ batched level-wise trees */
/* Many WT(s_i, n_i, a_i) builds (wavelet/levelWT/WT.h:43, levels
 * L_i = max(1, log2Up(a_i)), a_i <= 256, byte symbols) in one go.  Sequence i is
 * d_sym[h_off[i], h_off[i] + n_i).  Sequences are grouped by L_i, at most
 * 2^cap_log[L_i] per group (default 2^(16-L_i): 16-bit keys); a group holds at most
 * 2^32 - 1 symbols.  Output (device, owned by the batch): level l of sequence i is the
 * bit range [first_bit + l*level_stride, ... + n_i) of the batch words (wtg_batch_seq). */
typedef struct wtg_batch wtg_batch;
/* Group sizes: cap_log[L] = log2 of the most sequences in one group of alphabet
 * width L (1..8; cap_log + L <= 24; keys are 8-bit when the group's sequence bits
 * + L <= 8, 16-bit when <= 16, else 32-bit); 0xFF = the default 16 - L.  cap_log[0]
 * is unused. */
typedef struct {
  uint8_t cap_log[9];
} wtg_batch_plan;
int wtg_batch_create(wtg_ctx* ctx, const uint8_t* d_sym, uint32_t nseq, const uint64_t* h_off,
                     const uint64_t* h_n, const uint32_t* h_sigma, wtg_batch** out);
int wtg_batch_create_planned(wtg_ctx* ctx, const uint8_t* d_sym, uint32_t nseq, const uint64_t* h_off,
                             const uint64_t* h_n, const uint32_t* h_sigma, const wtg_batch_plan* plan,
                             wtg_batch** out);
/* Autotuner: per alphabet width present, builds the batch's own sequences of that
 * width with each candidate group size (2^(8-L): 8-bit keys; 2^(14-L), 2^(16-L):
 * 16-bit keys; 2^(20-L), 2^(24-L): 32-bit keys; at most 64 groups), times the best of
 * `reps` graph replays and returns the fastest plan; ms_out (nullable, 45 floats) gets
 * the time of candidate c of width L at [L*5 + c] (-1: not tried). */
int wtg_batch_tune(wtg_ctx* ctx, const uint8_t* d_sym, uint32_t nseq, const uint64_t* h_off,
                   const uint64_t* h_n, const uint32_t* h_sigma, uint32_t reps, wtg_batch_plan* out,
                   float* ms_out);
int wtg_batch_destroy(wtg_ctx* ctx, wtg_batch* b);
/* 1 (default): the first run records every kernel of the batch as one hipGraph,
 * later runs replay it with a single launch; 0: plain stream launches */
int wtg_batch_set_graph(wtg_batch* b, int use_graph);
int wtg_batch_run(wtg_ctx* ctx, wtg_batch* b);
int wtg_batch_words(const wtg_batch* b, const uint64_t** d_words, uint64_t* nwords);
int wtg_batch_seq(const wtg_batch* b, uint32_t seq, uint32_t* levels, uint64_t* first_bit,
                  uint64_t* level_stride);
/* sequence `seq` in the WT() layout (level l at bit l*n_i): wtg_levels_words(n_i, L_i) words */
int wtg_batch_extract(wtg_ctx* ctx, const wtg_batch* b, uint32_t seq, uint64_t* d_out);

/* ------------------------------------------------ learned select (added form) */
/* No counterpart in the reference: select over a batch of bitvectors where a
 * per-vector piecewise-linear model predicts select(j) and the rank9 directory
 * (rank_support_v_par.hpp:131-142) recovers the exact position.  The model's
 * knots are placed by a greedy error cone over the vector's ones and trained on
 * the GPU (forward/backward kernels, Adam) on the answers of the plain sampled
 * select (select_support_mcl_par.hpp:564-611).  Answers always equal the plain
 * method's. */
typedef struct {
  uint32_t train_stride;  /* training samples select(T), select(2T), ...; power of two (16) */
  uint32_t max_gap;       /* longest knot interval in arguments; power of two, 2..2^24 (4096) */
  uint32_t tol;           /* knot placement tolerance in bits, 1 <= tol < eps_max - 1 (480) */
  uint32_t epochs;        /* Adam steps (50) */
  float lr;               /* Adam step in bit positions (4) */
  uint32_t eps_max;       /* accepted error window in bits, <= 512 (512) */
} wtg_lsel_params;
typedef struct {
  uint64_t nbits, ones;
  uint32_t learned;       /* 1: answered through the model; 0: by the plain sampled select */
  uint32_t nknots;
  uint64_t eps;           /* max |prediction - select(j)| over every j (exhaustive), model in use */
  uint64_t eps_init;      /* the same for the placed, untrained knots */
  double loss_init, loss_final;  /* mean squared residual on the training samples: placed knots /
                                    after training (the knots in use when trained = 1) */
  uint32_t trained;       /* 1: the Adam-trained knots are in use; 0: the placed ones (training
                             raised eps above eps_max) */
  uint32_t lut_log;       /* knot lookup table: one entry per 2^lut_log arguments */
} wtg_lsel_vec_info;
/* A vector's model on the device: d_knots = nknots pairs {uint32 j; float y}
 * (j strictly increasing from 1 to ones), pred(j) = alpha*j + beta + linear
 * interpolation of y; d_lut[i] = the knot interval holding argument i << lut_log. */
typedef struct {
  const void* d_knots;
  uint32_t nknots, lut_log;
  const uint32_t* d_lut;
  uint64_t lut_size;
  double alpha, beta;
  uint32_t learned, _pad;
} wtg_lsel_model_view;
typedef struct wtg_lsel wtg_lsel;
/* d_words[b]: device words of vector b (nbits[b] bits, zero padding word after
 * them), 1 <= nvec <= 256; params NULL = defaults.  Builds rank9 + select1 per
 * vector, places and trains the models, measures eps exhaustively. */
int wtg_lsel_build(wtg_ctx* ctx, uint32_t nvec, const uint64_t* const* d_words,
                   const uint64_t* nbits, const wtg_lsel_params* params, wtg_lsel** out);
int wtg_lsel_destroy(wtg_ctx* ctx, wtg_lsel* ls);
int wtg_lsel_info(const wtg_lsel* ls, uint32_t vec, wtg_lsel_vec_info* out);
int wtg_lsel_model(const wtg_lsel* ls, uint32_t vec, wtg_lsel_model_view* out);
/* out[q] = select1 of vector d_vec[q] at rank d_j[q] (1-based); WTG_NONE when
 * d_j[q] is 0 or above the vector's ones, or d_vec[q] >= nvec */
int wtg_lsel_select_batch(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec,
                          const uint64_t* d_j, uint64_t nq, uint64_t* d_out);
/* the same queries answered by the plain sampled select only */
int wtg_lsel_plain_batch(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec,
                         const uint64_t* d_j, uint64_t nq, uint64_t* d_out);
/* Access report: out[q] = the 64-byte HBM sectors query q reads on the learned
 * (learned = 1) or the plain path (select_support_mcl_par.hpp:564-611: sample
 * lookups + forward word scan), counted along the same addresses; d_ans = the
 * answers of a previous wtg_lsel_select_batch / wtg_lsel_plain_batch call on
 * the same queries.  0 for an invalid query. */
int wtg_lsel_sectors(wtg_ctx* ctx, const wtg_lsel* ls, const uint32_t* d_vec, const uint64_t* d_j,
                     const uint64_t* d_ans, uint64_t nq, int learned, uint32_t* d_out);
/* Training loss (mean squared residual) and its gradient for one model, by the
 * training kernels: d_knots = nknots pairs {uint32 j; float y} (j increasing),
 * samples d_pos[t] = select((t+1)*stride) (those outside [j_0, j_last] unseen),
 * baseline alpha*j + beta. */
int wtg_lsel_loss_grad(wtg_ctx* ctx, const void* d_knots, uint32_t nknots, uint32_t stride,
                       double alpha, double beta, const uint64_t* d_pos, uint64_t ns,
                       double* loss, float* d_grad);

/* ------------------------------------------------ synthetic workloads (host) */
/* Seeded generators of the benchmark inputs (splitmix64 streams). */
uint64_t wtg_splitmix64(uint64_t* state);
int wtg_gen_uniform(uint64_t seed, uint64_t n, uint64_t range, int sym_bytes, void* out);
int wtg_gen_geometric(uint64_t seed, uint64_t n, double p, uint64_t cap, int sym_bytes, void* out);
int wtg_gen_zipf_bytes(uint64_t seed, uint64_t n, double s, uint8_t* out);
/* Batched workload: nseq sequences, n_i = floor(2^U(log2_min, log2_max)),
 * alphabet a_i uniform over alphabets[0..nalpha); symbols (host, caller-sized
 * sum(n_i) bytes, sequence i at offset sum_{k<i} n_k) uniform in [0, a_i).
 * With sym == NULL only n and sigma are written. */
int wtg_gen_batch(uint64_t seed, uint32_t nseq, double log2_min, double log2_max,
                  const uint32_t* alphabets, uint32_t nalpha, uint64_t* n, uint32_t* sigma,
                  uint8_t* sym);
/* nvec bitvectors of nbits bits (vector b at words + b*ceil(nbits/64)), density
 * d_b log-uniform in [0.01, 0.5] (density may be NULL): even b Bernoulli(d_b);
 * odd b alternating runs, 0-runs Geometric(mean 64(1-d_b)/d_b), 1-runs
 * Geometric(mean 64) */
int wtg_gen_select_vectors(uint64_t seed, uint32_t nvec, uint64_t nbits, uint64_t* words,
                           double* density);
/* n bytes: order-5 Markov text over ACGT (Dirichlet(0.5) transition rows), 5% of
 * positions overwritten by copies of 1-10 kb segments, 0.1% 'N', text[n-1] = 0 */
int wtg_gen_dna(uint64_t seed, uint64_t n, uint8_t* out);
/* npat patterns of m bytes from text[0, n-1): even q exact substrings, odd q with
 * 1-2 substitutions by ACGT */
int wtg_gen_patterns(uint64_t seed, const uint8_t* text, uint64_t n, uint32_t m,
                     uint64_t npat, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif /* WTG_H */
