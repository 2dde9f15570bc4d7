// pwm_amd/csrc/pwm_internal.hpp -- shared declarations of the gfx950 engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <unordered_map>
#include <utility>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pwm_hip.h"

namespace pwm {

// ---- status handling ------------------------------------------------------
struct hip_failure {
  int status;
  hipError_t err;
  const char* where;
};

#define PWM_HIP_TRY(expr)                                                  \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      throw ::pwm::hip_failure{e_ == hipErrorOutOfMemory ? PWMHIP_ERR_OOM \
                                                         : PWMHIP_ERR_HIP, \
                               e_, #expr};                                 \
    }                                                                      \
  } while (0)

struct status_error {
  int status;
  std::string what;
};

// Runs fn() and turns every exception into a status, so that nothing
// crosses the C-ABI (or leaves a thread body): HIP failures keep their
// status, host allocation failures (bad_alloc, a vector past max_size) are
// out-of-memory, anything else is a runtime error.
template <typename F>
int guarded(F&& fn) noexcept {
  try {
    fn();
    return PWMHIP_OK;
  } catch (const hip_failure& f) {
    (void)hipGetLastError();
    return f.status;
  } catch (const status_error& e) {
    return e.status;
  } catch (const std::bad_alloc&) {
    return PWMHIP_ERR_OOM;
  } catch (const std::length_error&) {
    return PWMHIP_ERR_OOM;
  } catch (...) {
    return PWMHIP_ERR_HIP;
  }
}

// ---- geometry ---------------------------------------------------------------
// A chunk is the unit of the local build (one wave): kChunkWords 64-bit
// words per bit plane (one per lane), i.e. kChunkWords*64 symbols.
constexpr int kChunkWords = 64;
constexpr uint64_t kChunkSyms = uint64_t(kChunkWords) * 64;
constexpr int kMaxGroup = 8;  // levels built per pass (2^8 buckets per chunk)

inline uint64_t words_for(uint64_t bits) { return (bits + 63) >> 6; }
// words between consecutive levels of one allocation: 256-byte aligned with
// at least two spare words so that funnel-shift reads past the last word of
// a level stay inside the allocation.
inline uint64_t level_stride(uint64_t bits) { return (words_for(bits) + 2 + 31) & ~uint64_t(31); }

// ---- device scratch arena (per context) ------------------------------------
struct arena {
  void* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  void reset() { used = 0; }
  void ensure(size_t bytes);  // reallocates (discarding contents) if needed
  template <typename T>
  T* take(size_t count) {
    size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
    if (used + bytes > cap) throw status_error{PWMHIP_ERR_OOM, "arena overflow"};
    T* p = reinterpret_cast<T*>(static_cast<char*>(base) + used);
    used += bytes;
    return p;
  }
  void release();
};

}  // namespace pwm

struct pwmhip_ctx {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_ms[4] = {0, 0, 0, 0};
  pwm::arena scratch;       // per-build temporaries
  pwm::arena text_scratch;  // device copy of a host text
  pwm::arena rs_scratch;    // query staging
  // pinned staging for small device->host results
  uint64_t* pinned = nullptr;
  size_t pinned_words = 0;
  // 64 KiB byte-compress table (lut[(m << 8) | x] = pext(x, m)), filled once
  uint8_t* d_lut = nullptr;
  // pinned host staging of batched builds (metadata up, zero counts down)
  void* hstage = nullptr;
  size_t hstage_bytes = 0;
  // batched builds: tuned (work-queue grab, gather windows) per length /
  // level bucket, and the last captured launch sequence with its geometry
  std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>> tune;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t gkey[12] = {};
  // streamed host texts: copy stream, slice events (copied, consumed)
  hipStream_t copy_stream = nullptr;
  hipEvent_t sev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace pwm {

// Code tables of a Huffman-shaped structure for symbol queries (device).
struct HuffQuery {
  const uint64_t* len;       // code length per symbol (0: absent)
  const uint64_t* word;      // code word per symbol
  uint64_t nsym;
  const uint64_t* leaf_key;  // (length << 48) | word, ascending
  const uint32_t* leaf_sym;
  uint32_t nleaf;
};

// One resident structure as the multi-structure query kernel reads it.
struct QueryIndexDesc {
  const void* lv;          // per-level rank/select descriptors (device)
  const uint64_t* zeros;   // true zero counts (device)
  uint64_t n;
  uint32_t L;
  int32_t tree;
  const HuffQuery* huff;   // device, Huffman-shaped structures only
};
QueryIndexDesc query_desc(const pwmhip_rs* rs);
// Answers n symbol queries (op 0 access, 1 rank, 2 select; as pwmhip_query)
// over several structures: query i reads d_desc[d_qidx[i]].  All device
// pointers; enqueued on stream s.
void launch_multi_query(hipStream_t s, const QueryIndexDesc* d_desc, const uint32_t* d_qidx, const uint32_t* op,
                        const uint64_t* a, const uint64_t* b, uint64_t* res, uint64_t n);

// The context's compress table, computed on first use.
const uint8_t* compress_table(pwmhip_ctx* ctx);

// ---- plain (non-Huffman) construction -------------------------------------
// Builds `levels` levels of a wavelet matrix (tree=false) or tree (tree=true)
// for a device-resident text into ws (which the caller has sized).  Fills
// ws->zeros with the true zero counts (matrix).  Throws on error.
// h_stream: a host text copied to the device in slices while the local
// build runs (d_text unused); only 2..4 levels (ERR_UNSUPPORTED otherwise).
// no_alloc: the allocation-free variant's static scratch (temporaries), its
// compress table and its capacity; the build then allocates nothing (device
// or pinned host memory) and fails with PWMHIP_ERR_BUDGET before any work if
// its temporaries do not fit.
struct NoAlloc {
  void* scratch;
  size_t cap;
  const uint8_t* lut;
};
void build_plain_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width,
                        uint32_t levels, bool tree, bool check_range, pwmhip_ws* ws,
                        const void* h_stream = nullptr, const NoAlloc* no_alloc = nullptr);

// Sets up the allocation-free build on the context's device (once per
// device): the static scratch's device address and its compress table.
void noalloc_setup(pwmhip_ctx* ctx);

// zeros[levels-1] is never computed by these reference algorithms
// (ppc.hpp:19-90, ppc_ss.hpp:19-70, pc_ss.hpp:13-33, wx_dd_pc_ss.hpp).
inline bool zeros_last_is_zero(uint32_t algo) {
  return algo == PWMHIP_ALGO_PPC || algo == PWMHIP_ALGO_PPC_SS || algo == PWMHIP_ALGO_PC_SS ||
         algo == PWMHIP_ALGO_DD_PC_SS;
}

// Builds T texts (device pointers) with one launch per phase.  All
// structures live in one allocation owned by outs[0] (alloc_bytes > 0, the
// others are views with alloc_bytes == 0).  levels <= kMaxGroup.  `reuse`
// (reuse_bytes) is an earlier batch's block, taken over when large enough
// and freed otherwise.
// flags: PWMHIP_FLAG_GRAPH / PWMHIP_FLAG_AUTOTUNE.
void build_batch_device(pwmhip_ctx* ctx, const void* const* d_texts, const uint64_t* sizes, uint64_t T,
                        uint32_t width, uint32_t L, bool tree, bool check_range, pwmhip_ws* outs, void* reuse,
                        uint64_t reuse_bytes, uint32_t flags);

// Allocates ws->d_words for `levels` levels of `bits[l]` bits each and fills
// the offsets.
// An earlier allocation (reuse, reuse_bytes) is kept if it is large enough
// and freed otherwise.
void alloc_structure(pwmhip_ws* ws, uint64_t levels, const uint64_t* bits,
                     uint64_t* reuse = nullptr, uint64_t reuse_bytes = 0);
// The same layout without the allocation: fills levels, offsets, level bits
// and returns the words it needs.
uint64_t layout_structure(pwmhip_ws* ws, uint64_t levels, const uint64_t* bits);

// Exclusive scan over chunks of counts[c*nslot + slot] (one block per slot).
// tile_scratch holds scan_scratch_words(nchunks, nslot) words.
// indicator: scan (count != 0) instead of count.
void launch_scan_slots(hipStream_t s, const uint32_t* counts, uint64_t* pref, uint64_t* totals,
                       uint64_t nchunks, uint32_t nslot, uint64_t* tile_scratch, bool indicator = false);
uint64_t scan_scratch_words(uint64_t nchunks, uint32_t nslot);

// Huffman-shaped local build for byte texts with codes of 2..4 bits
// (build.hip): level 0 to its final place, sub-levels 1..levels-1 per
// super-chunk (16 chunks) in slot order, counts per (super-chunk, slot).
struct HuffSuperParams {
  const uint8_t* text;
  uint64_t n, nchunks;
  uint32_t levels;            // 2..4
  const uint8_t* code;        // device, 256 entries: code left-aligned in `levels` bits
  uint32_t small_alpha;       // every symbol < 8 (byte permute instead of table reads)
  uint64_t* level0;
  uint64_t* temp;             // sub-level j at temp + (j - 1) * temp_stride
  uint64_t temp_stride;
  uint32_t* counts;           // [unit * nslot + slot]
  uint32_t nslot;
  uint32_t slot0[4];          // first slot of sub-level j
  uint8_t rank[4][8];         // slot of bucket q of sub-level j, minus slot0[j]; 0xff: no node
};
void launch_huff_super(pwmhip_ctx* ctx, const HuffSuperParams& hp);
constexpr uint64_t kSuperChunkSyms = uint64_t(16) * kChunkSyms;  // symbols per super-chunk

// Destination-major stitch of sub-levels 1..nsub from unit runs (build.hip's
// wide stitch), for callers with their own run tables: level j's entries
// (destination bit, source bit) in destination order from
// ((1 << j) - 2) * units, level_ne[j] of them; sub-level j's unit-local bits
// at src + (j - 1) * src_stride; nbits[j] bits written to dst[j].  `first`
// holds wide_stitch_first_words(max nbits, nsub) words.
struct WideStitch {
  const uint64_t* tab_dst;
  const uint64_t* tab_src;
  const uint64_t* level_ne;
  uint32_t* first;
  const uint64_t* src;
  uint64_t src_stride;
  uint64_t* dst[kMaxGroup];
  uint64_t nbits[kMaxGroup];
  uint64_t units;
  uint32_t nsub;
};
void launch_wide_stitch(pwmhip_ctx* ctx, const WideStitch& w);
uint64_t wide_stitch_first_words(uint64_t max_bits, uint32_t nsub);

// Huffman-shaped construction (huffman.hip).
// `reuse`: an earlier output block of `reuse_bytes` bytes, taken over when
// large enough, freed otherwise (also when the build fails).
void build_huffman_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width,
                          bool tree, bool pc_quirks, pwmhip_ws* ws, uint64_t* reuse = nullptr,
                          uint64_t reuse_bytes = 0);

// copy_bits on the device (bitops.hip).
void copy_bits_device(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t dst_off,
                      uint64_t src_off, uint64_t len);

// WT levels from WM levels by moving each level's buckets into numeric
// prefix order (bitops.hip).
void wm_to_wt_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width,
                     uint32_t levels, pwmhip_ws* ws);

}  // namespace pwm
