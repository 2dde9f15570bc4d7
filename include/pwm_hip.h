/* include/pwm_hip.h -- C-ABI of the MI355X (gfx950) wavelet matrix / wavelet
 * tree construction engine.
 *
 * Plain pointers and sizes only.  Every call returns an int status (0 = OK,
 * negative = error, see pwmhip_status_string), never aborts and never throws
 * across the ABI.  Each context owns one HIP stream, a device scratch arena
 * and an autotune table; calls on distinct contexts are thread-safe, calls on
 * one context must be serialised by the caller.
 *
 * Output layout (identical to the reference's bit_vectors contract):
 *   level l holds level_bits[l] bits in ceil(level_bits[l]/64) uint64 words;
 *   bit j lives in word j>>6 at mask 1<<(63-(j&63)) (MSB first); bits past
 *   level_bits[l] in the last word are 0.
 *   reference: include/util/common.hpp:60-68,
 *              include/construction/building_blocks.hpp:18-50,
 *              include/arrays/flat_two_dim_array.hpp:110-213.
 */
#ifndef PWM_HIP_H
#define PWM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PWMHIP_MAX_LEVELS 64

/* status codes */
#define PWMHIP_OK 0
#define PWMHIP_ERR_INVALID -1      /* bad argument (levels == 0, width, ...) */
#define PWMHIP_ERR_NO_DEVICE -2    /* no HIP device / device index invalid  */
#define PWMHIP_ERR_OOM -3          /* hipMalloc failed                      */
#define PWMHIP_ERR_HIP -4          /* other HIP runtime error               */
#define PWMHIP_ERR_SYMBOL_RANGE -5 /* a symbol >= 2^levels                  */
#define PWMHIP_ERR_UNSUPPORTED -6  /* input the reference leaves undefined  */
#define PWMHIP_ERR_BUDGET -7       /* request exceeds the device budget     */
#define PWMHIP_ERR_TIMEOUT -8      /* wait ended before the request did     */

/* Construction algorithms.  They produce identical level bits; they differ
 * only where the reference's outputs differ:
 *   - zeros[levels-1] is 0 for PPC, PPC_SS, PC_SS, DD_PC_SS (those never
 *     compute it: construction/ppc.hpp:19-90, pc_ss.hpp:13-33,
 *     ppc_ss.hpp:19-70) and the true count for the others;
 *   - the HUFF_* algorithms build Huffman-shaped structures and ignore
 *     `levels` (wx_huff_pc.hpp:45-80).  HUFF_PC, HUFF_PS and HUFF_DD_PC
 *     reproduce huff_pc's border bookkeeping (zeros included; the
 *     reference's huff_dd_pc matches it at one OpenMP thread, its output
 *     depends on the thread count above that); HUFF_NAIVE gives
 *     huff_naive's levels and true zero counts; HUFF_PC_SS gives
 *     huff_naive's levels with zeros[levels-1] = 0 (huff_pc_ss.hpp:18-121). */
enum pwmhip_algo {
  PWMHIP_ALGO_PC = 0,
  PWMHIP_ALGO_PS = 1,
  PWMHIP_ALGO_PPC = 2,
  PWMHIP_ALGO_PPS = 3,
  PWMHIP_ALGO_PC_SS = 4,
  PWMHIP_ALGO_PPC_SS = 5,
  PWMHIP_ALGO_DD_PC = 6,
  PWMHIP_ALGO_DD_PS = 7,
  PWMHIP_ALGO_DD_PC_SS = 8,
  PWMHIP_ALGO_NAIVE = 9,
  PWMHIP_ALGO_HUFF_PC = 10,
  PWMHIP_ALGO_HUFF_PS = 11,
  PWMHIP_ALGO_HUFF_PC_SS = 12,
  PWMHIP_ALGO_HUFF_NAIVE = 13,
  PWMHIP_ALGO_HUFF_DD_PC = 14
};

/* build flags */
#define PWMHIP_FLAG_TREE 1u            /* wavelet tree instead of matrix       */
#define PWMHIP_FLAG_TEXT_ON_DEVICE 2u  /* `text` is a device pointer           */
#define PWMHIP_FLAG_CHECK_RANGE 4u     /* verify all symbols < 2^levels        */
#define PWMHIP_FLAG_REUSE_OUTPUT 8u    /* `out` holds an earlier result whose   */
                                       /* device memory is reused if it fits   */
#define PWMHIP_FLAG_GRAPH 16u          /* batches: replay the captured launch   */
                                       /* sequence when the geometry repeats   */
#define PWMHIP_FLAG_AUTOTUNE 32u       /* batches: time the launch variants of  */
                                       /* a new length/level bucket, keep best */
#define PWMHIP_FLAG_STREAM 64u         /* host text copied in slices while the  */
                                       /* local build runs; only the output and */
                                       /* scratch live in HBM (2..4 levels;     */
                                       /* the analogue of wx_ppc_ie,            */
                                       /* external_memory/wx_ppc_ie.hpp:21-65)  */

typedef struct pwmhip_ctx pwmhip_ctx;

/* Result of a construction: device-resident levels plus host metadata. */
typedef struct pwmhip_ws {
  uint64_t levels;
  uint64_t text_size;
  int32_t is_tree;
  int32_t is_huffman;  /* 0 plain; Huffman-shaped: 1 with huff_pc's node
                          borders where they differ from huff_naive's (no
                          symbol queries), 2 with huff_naive's layout     */
  uint64_t* d_words;                          /* device allocation         */
  uint64_t level_offset[PWMHIP_MAX_LEVELS];   /* word offset of level l    */
  uint64_t level_bits[PWMHIP_MAX_LEVELS];     /* bits in level l           */
  uint64_t n_zeros;                           /* = levels for matrices     */
  uint64_t zeros[PWMHIP_MAX_LEVELS];
  uint64_t n_codes;                           /* Huffman: max_symbol + 1   */
  uint64_t* code_len;                         /* host arrays (n_codes)     */
  uint64_t* code_word;
  uint64_t alloc_bytes;
} pwmhip_ws;

/* ---- context ------------------------------------------------------------
 * Replaces nothing in the reference (it has no device state); every
 * wx_*::compute call of the reference (include/wx_*.hpp) becomes
 * pwmhip_build on a context. */
int pwmhip_create(pwmhip_ctx** ctx, int device);
int pwmhip_destroy(pwmhip_ctx* ctx);
const char* pwmhip_status_string(int status);
int pwmhip_device_count(int* count);

/* ---- construction --------------------------------------------------------
 * Replaces wx_<algo><A, is_tree>::compute(text, size, levels)
 * (wx_ppc.hpp:36-58, wx_pps.hpp:37-68, wx_dd_pc.hpp:42-119, wx_pc.hpp,
 *  wx_ps.hpp, wx_pc_ss.hpp, wx_ppc_ss.hpp, wx_dd_ps.hpp, wx_dd_pc_ss.hpp,
 *  wx_naive.hpp, wx_huff_pc.hpp:45-80, wx_huff_ps.hpp, wx_huff_pc_ss.hpp,
 *  wx_huff_naive.hpp, wx_huff_dd_pc.hpp).
 * word_width = sizeof(A) in {1, 2, 4, 8}.  size == 0 gives an empty
 * structure (levels = 0), as the reference does. */
int pwmhip_build(pwmhip_ctx* ctx, const void* text, uint64_t size, uint32_t word_width,
                 uint32_t levels, uint32_t algo, uint32_t flags, pwmhip_ws* out);

int pwmhip_ws_free(pwmhip_ctx* ctx, pwmhip_ws* ws);

/* ---- allocation-free construction ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same build as pwmhip_build for plain algorithms
 * (algo <= PWMHIP_ALGO_NAIVE), 2 to 4 levels, a device text, into the
 * caller's device memory `d_out` of `out_words` 64-bit words, with the
 * build's temporaries in a static device array of the library (128 MiB plus
 * a 64 KiB table, held on every device the library runs on until the process
 * ends).  It calls no hipMalloc, hipFree or other allocating routine.
 * Results equal pwmhip_build's word for word; `out` describes them (level l
 * at d_out + out->level_offset[l]; alloc_bytes = 0: pwmhip_ws_free frees
 * nothing).  Needed: out_words >= levels * ((ceil(size/64) + 2 + 31) / 32 * 32).
 * Flags: PWMHIP_FLAG_TREE, PWMHIP_FLAG_CHECK_RANGE (PWMHIP_FLAG_TEXT_ON_DEVICE
 * is implied).  Texts whose temporaries exceed the static array (about 2^29
 * symbols at 3 levels: (levels - 1) * size / 8 bytes plus per-64 Ki-symbol
 * tables) return PWMHIP_ERR_BUDGET before any work; other shapes return
 * PWMHIP_ERR_UNSUPPORTED.  Every call on one device shares its array, and
 * each call leaves its contents behind: calls that overlap on one device
 * (two contexts, two threads) race and must be serialised by the caller. */
int pwmhip_build_noalloc(pwmhip_ctx* ctx, const void* d_text, uint64_t size, uint32_t word_width,
                         uint32_t levels, uint32_t algo, uint32_t flags, uint64_t* d_out,
                         uint64_t out_words, pwmhip_ws* out);

/* Copies level `level` (ceil(level_bits/64) words) to host memory. */
int pwmhip_ws_download(pwmhip_ctx* ctx, const pwmhip_ws* ws, uint64_t level, uint64_t* host_words);

/* ---- rank / select support ----------------------------------------------
 * Replaces query_support<> (queries/query_support.hpp:18-186) with its
 * per-level bin_rank_popcnt (bin_rank_popcnt.hpp:54-86) and
 * bin_select_popcnt<0/1> (bin_select_popcnt.hpp:38-107) directories.
 * Tables: per level, 64 Ki-bit superblocks with absolute uint64 counts and
 * 512-bit blocks with uint16 counts relative to their superblock, plus
 * select samples every 8192 ones and zeros.  Valid for levels below 2^48
 * bits (superblock indices are 32-bit); longer levels return
 * PWMHIP_ERR_UNSUPPORTED.  The reference's directory is valid below 2^32-1
 * bits only (bin_rank_popcnt.hpp:54-86, popcnt_traits.hpp:16-20). */
typedef struct pwmhip_rs pwmhip_rs;
int pwmhip_rs_build(pwmhip_ctx* ctx, const pwmhip_ws* ws, pwmhip_rs** out);
int pwmhip_rs_free(pwmhip_ctx* ctx, pwmhip_rs* rs);

/* Bit-vector level queries (bin_rank_popcnt::rank1, bin_select_popcnt::select):
 * op 0: rank1(level, i)  = ones in [0, i)
 * op 1: rank0(level, i)
 * op 2: select1(level, k) = 0-based position of the k-th one (k >= 1; 0 -> 0)
 * op 3: select0(level, k)
 * arrays are host or device pointers (flag PWMHIP_FLAG_TEXT_ON_DEVICE). */
int pwmhip_bv_query(pwmhip_ctx* ctx, const pwmhip_rs* rs, const uint32_t* level,
                    const uint32_t* op, const uint64_t* arg, uint64_t n, uint64_t* result,
                    uint32_t flags);

/* Symbol queries over the whole structure (query_support::access / rank /
 * select, query_support.hpp:37-69), one GPU lane per query:
 * op 0: access(a)            -> symbol at position a
 * op 1: rank(symbol a, i b)  -> occurrences of a in [0, b)
 * op 2: select(symbol a, k b)-> 0-based position of the b-th a (b >= 1),
 *                               UINT64_MAX if a occurs fewer than b times
 * Works for matrices and trees, plain and Huffman-shaped (the reference has
 * none for Huffman shapes, query_support.hpp:43,55,67): a symbol's path is
 * its code; structures with is_huffman == 1 (huff_pc / huff_ps / huff_dd_pc
 * node borders that differ from huff_naive's) return PWMHIP_ERR_UNSUPPORTED
 * here and in pwmhip_rs_build. */
int pwmhip_query(pwmhip_ctx* ctx, const pwmhip_ws* ws, const pwmhip_rs* rs,
                 const uint32_t* op, const uint64_t* a, const uint64_t* b, uint64_t n,
                 uint64_t* result, uint32_t flags);

/* ---- bit copy -------------------------------------------------------------
 * Replaces copy_bits<uint64_t> (construction/merge.hpp:19-190): copies
 * `len` bits from src[src_off..] to dst[dst_off..] on the device (MSB-first
 * words); partially covered destination words are OR-ed, covered words are
 * overwritten.  Pointers are device pointers. */
int pwmhip_copy_bits(pwmhip_ctx* ctx, uint64_t* d_dst, const uint64_t* d_src, uint64_t dst_off,
                     uint64_t src_off, uint64_t len);

/* ---- alphabet reduction ----------------------------------------------------
 * Replaces reduce_alphabet<A> + levels_for_max_char (util/alphabet_util.hpp:
 * 18-45, 99-106): remaps symbols in first-occurrence order (0, 1, 2, ...),
 * in place on a device text, and returns the number of levels.  Unlike the
 * reference's 255-entry table, byte value 255 is handled. */
int pwmhip_reduce_alphabet(pwmhip_ctx* ctx, void* d_text, uint64_t size, uint32_t word_width,
                           uint64_t* levels_out);

/* ---- batched construction ------------------------------------------------
 * No counterpart in the reference (it builds one text per call).  Builds
 * `count` independent texts (host or device pointers, uniform word width,
 * levels and algorithm) in one persistent launch per phase from a device
 * work queue; results identical to one pwmhip_build per text.
 * Plain algorithms with levels <= 8: every structure lives in one device
 * allocation owned by outs[0] (alloc_bytes > 0); outs[1..] are views
 * (alloc_bytes == 0, pwmhip_ws_free frees nothing of theirs), so free
 * outs[0] last.  PWMHIP_FLAG_REUSE_OUTPUT takes over outs[0]'s earlier block
 * when it is large enough.  Other shapes (Huffman, more levels) are built
 * text by text and every structure owns its memory. */
int pwmhip_build_batch(pwmhip_ctx* ctx, const void* const* texts, const uint64_t* sizes,
                       uint64_t count, uint32_t word_width, uint32_t levels, uint32_t algo,
                       uint32_t flags, pwmhip_ws* outs);

/* ---- serving ---------------------------------------------------------------
 * No counterpart in the reference.  A request loop on one device: builds and
 * query batches are submitted from any thread and return at once; a
 * scheduler admits builds while the HBM budget allows (resident indices +
 * outputs of builds in flight + each worker's scratch allowance), groups
 * waiting builds of one shape (width, levels, algorithm, tree) into batched
 * launches, runs them on `workers` streams concurrently, keeps every built
 * index (levels + rank/select directories) resident until released, and
 * answers waiting queries of all ready indices together in one launch.
 * Each result equals what the same request gives alone (pwmhip_build +
 * pwmhip_query).  Plain algorithms only (a Huffman shape's size is known only
 * after its code plan, which the admission estimate would need).
 * Host buffers passed in (texts, query arrays, result arrays) must stay
 * valid until the request completes. */
typedef struct pwmhip_server pwmhip_server;

typedef struct pwmhip_server_config {
  int device;
  uint64_t hbm_budget_bytes;    /* 0: 200 GB                                    */
  uint32_t workers;             /* concurrent streams; 0: 2                     */
  uint32_t max_batch_texts;     /* texts per batched launch; 0: 1024            */
  uint64_t max_batch_bytes;     /* text bytes per batched launch; 0: 512 MiB    */
  uint64_t max_query_batch;     /* queries per launch; 0: 8 Mi                  */
} pwmhip_server_config;

typedef struct pwmhip_server_stats {
  uint64_t builds_done, builds_failed, queries_done, query_requests_done;
  uint64_t build_launches, query_launches;
  double build_p50_ms, build_p99_ms;   /* submit -> index ready            */
  double query_p50_ms, query_p99_ms;   /* submit -> results written         */
  uint64_t bytes_in_use, bytes_peak;   /* against the budget                */
  double mean_texts_per_launch;
} pwmhip_server_stats;

int pwmhip_server_create(const pwmhip_server_config* cfg, pwmhip_server** out);
/* Finishes all submitted work, then frees everything. */
int pwmhip_server_destroy(pwmhip_server* s);
/* Submits a build (arguments as pwmhip_build; host text).  *index_id names
 * the resident index.  PWMHIP_ERR_BUDGET if the request can never fit. */
int pwmhip_server_build(pwmhip_server* s, const void* text, uint64_t size, uint32_t word_width,
                        uint32_t levels, uint32_t algo, uint32_t flags, uint64_t* index_id);
/* Submits n symbol queries (as pwmhip_query) against an index, which may
 * still be building; results land in `result` (host). */
int pwmhip_server_query(pwmhip_server* s, uint64_t index_id, const uint32_t* op, const uint64_t* a,
                        const uint64_t* b, uint64_t n, uint64_t* result, uint64_t* request_id);
/* Waits until request `id` (an index id or a query request id) completes;
 * returns its status.  timeout_ms < 0 waits without limit; on timeout
 * returns PWMHIP_ERR_TIMEOUT. */
int pwmhip_server_wait(pwmhip_server* s, uint64_t id, int64_t timeout_ms);
/* Copies level `level` of a ready index to host memory. */
int pwmhip_server_download(pwmhip_server* s, uint64_t index_id, uint64_t level, uint64_t* host_words);
/* Metadata of a ready index (device pointers stay owned by the server). */
int pwmhip_server_index_info(pwmhip_server* s, uint64_t index_id, pwmhip_ws* info);
/* Drops a resident index (after its queries complete). */
int pwmhip_server_release(pwmhip_server* s, uint64_t index_id);
int pwmhip_server_stats_get(pwmhip_server* s, pwmhip_server_stats* out);

/* ---- timing -------------------------------------------------------------
 * Device time of the last build's kernels (HIP events on the context's
 * stream), per phase: [0] level-0 + local build, [1] offsets, [2] stitch,
 * [3] total.  Milliseconds. */
int pwmhip_last_timing(pwmhip_ctx* ctx, double* ms4);

/* ---- synthetic inputs (the bench's seeded generators, host side) ------- */
/* kind 0: i.i.d. uniform symbols < 2^levels (splitmix64(seed), LE bytes)
 * kind 1: DNA-like codes 0..4 (order-2 Markov over ACGT, Dirichlet(1) rows,
 *         N runs at rate 1e-5 with geometric length, mean 500)
 * kind 2: Zipf(1.1) over 2^levels ranks through a seeded permutation
 * Texts are generated in independent 1 Mi-symbol blocks, so the result does
 * not depend on the thread count. */
int pwmhip_generate(uint32_t kind, uint64_t seed, uint64_t size, uint32_t word_width,
                    uint32_t levels, void* host_out);

#ifdef __cplusplus
}
#endif

#endif /* PWM_HIP_H */
