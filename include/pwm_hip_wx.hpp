// include/pwm_hip_wx.hpp -- C++ host side of the MI355X construction engine:
// the reference's construction-algorithm interface over the C-ABI in
// pwm_hip.h.
//
// Every wx_hip_<algo><A, is_tree> class carries the traits and the static
// compute(text, size, levels) of the reference's wx_<algo><A, is_tree>
// (include/construction/wx_base.hpp:13-23, include/benchmark/algorithm.hpp:
// 195-279, e.g. include/construction/wx_ppc.hpp:26-58), so the reference's
// CONSTRUCTION_REGISTER machinery can register it unchanged.  compute()
// returns a wavelet_structure owning host copies of the level words in the
// reference's layout (include/construction/wavelet_structure.hpp:22-199,
// include/arrays/flat_two_dim_array.hpp:110-213) plus the device-resident
// original, which rank/select/access batches (query_support's drop-in,
// include/queries/query_support.hpp:18-186) read without a copy.
//
// Errors: the reference has no error channel; here every failing C-ABI call
// throws pwm_hip::error carrying the status (a symbol >= 2^levels, levels ==
// 0, no device, out of memory, or an input the reference leaves undefined).
//
// Header-only; link with -lpwm_hip (pwm_amd/libpwm_hip.so).
#ifndef PWM_HIP_WX_HPP
#define PWM_HIP_WX_HPP

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pwm_hip.h"

namespace pwm_hip {

class error : public std::runtime_error {
 public:
  error(int status, const std::string& what)
      : std::runtime_error(what + ": " + pwmhip_status_string(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* what) {
  if (status != PWMHIP_OK) throw error(status, what);
}

// One HIP stream, scratch arena and compress table on one device.
class context {
 public:
  explicit context(int device = 0) { check(pwmhip_create(&ctx_, device), "pwmhip_create"); }
  ~context() {
    if (ctx_) pwmhip_destroy(ctx_);
  }
  context(const context&) = delete;
  context& operator=(const context&) = delete;
  pwmhip_ctx* get() const { return ctx_; }
  // Calls on one context are serialised by this lock.
  std::mutex& lock() { return mu_; }

  // The process-wide context of device 0, created on first use.
  static context& shared() {
    static context c(0);
    return c;
  }

 private:
  pwmhip_ctx* ctx_ = nullptr;
  std::mutex mu_;
};

// Result of a construction (reference: wavelet_structure).  Move-only.
class wavelet_structure {
 public:
  wavelet_structure() = default;
  wavelet_structure(context& ctx, pwmhip_ws ws, std::shared_ptr<void> block = nullptr)
      : ctx_(&ctx), ws_(ws), block_(std::move(block)) {
    bvs_.resize(ws_.levels);
    for (uint64_t l = 0; l < ws_.levels; ++l) {
      bvs_[l].assign((ws_.level_bits[l] + 63) / 64, 0);
      if (!bvs_[l].empty()) check(pwmhip_ws_download(ctx.get(), &ws_, l, bvs_[l].data()), "pwmhip_ws_download");
    }
  }
  ~wavelet_structure() { release(); }
  wavelet_structure(wavelet_structure&& o) noexcept { *this = std::move(o); }
  wavelet_structure& operator=(wavelet_structure&& o) noexcept {
    if (this != &o) {
      release();
      ctx_ = o.ctx_;
      ws_ = o.ws_;
      bvs_ = std::move(o.bvs_);
      block_ = std::move(o.block_);
      o.ctx_ = nullptr;
      o.ws_ = pwmhip_ws{};
    }
    return *this;
  }
  wavelet_structure(const wavelet_structure&) = delete;
  wavelet_structure& operator=(const wavelet_structure&) = delete;

  // -- the reference's accessors (wavelet_structure.hpp:160-199) ----------
  uint64_t levels() const { return ws_.levels; }
  uint64_t text_size() const { return ws_.text_size; }
  bool is_tree() const { return ws_.is_tree != 0; }
  bool is_huffman_shaped() const { return ws_.is_huffman != 0; }
  uint64_t level_bit_size(uint64_t level) const { return ws_.level_bits[level]; }
  // level l: ceil(level_bit_size(l)/64) words, MSB first, zero padded
  const std::vector<std::vector<uint64_t>>& bvs() const { return bvs_; }
  // true zero counts per level (matrices; zeros[levels-1] per algorithm)
  std::vector<uint64_t> zeros() const { return std::vector<uint64_t>(ws_.zeros, ws_.zeros + ws_.n_zeros); }
  // Huffman code (length, word) per symbol 0..max_symbol
  std::vector<std::pair<uint64_t, uint64_t>> codes() const {
    std::vector<std::pair<uint64_t, uint64_t>> c(ws_.n_codes);
    for (uint64_t s = 0; s < ws_.n_codes; ++s) c[s] = {ws_.code_len[s], ws_.code_word[s]};
    return c;
  }

  // -- device side ----------------------------------------------------------
  const pwmhip_ws& device() const { return ws_; }
  context* ctx() const { return ctx_; }

 private:
  void release() {
    // views into a batch (alloc_bytes == 0) free only their host arrays
    if (ctx_ && (ws_.d_words || ws_.code_len)) pwmhip_ws_free(ctx_->get(), &ws_);
    ctx_ = nullptr;
    block_.reset();
  }
  context* ctx_ = nullptr;
  pwmhip_ws ws_{};
  std::shared_ptr<void> block_;  // device block shared by a batch's structures
  std::vector<std::vector<uint64_t>> bvs_;
};

// Rank / select / access over a structure (query_support's drop-in, one GPU
// lane per query).  Queries of one batch are answered together.
class query_support {
 public:
  explicit query_support(const wavelet_structure& ws) : ws_(&ws) {
    check(pwmhip_rs_build(ws.ctx()->get(), &ws.device(), &rs_), "pwmhip_rs_build");
  }
  ~query_support() {
    if (rs_) pwmhip_rs_free(ws_->ctx()->get(), rs_);
  }
  query_support(const query_support&) = delete;
  query_support& operator=(const query_support&) = delete;

  enum op : uint32_t { ACCESS = 0, RANK = 1, SELECT = 2 };
  // a/b per op: access(a), rank(symbol a, position b), select(symbol a, k b)
  std::vector<uint64_t> batch(const std::vector<uint32_t>& ops, const std::vector<uint64_t>& a,
                              const std::vector<uint64_t>& b) const {
    std::vector<uint64_t> r(ops.size());
    if (!ops.empty())
      check(pwmhip_query(ws_->ctx()->get(), &ws_->device(), rs_, ops.data(), a.data(), b.data(), ops.size(),
                         r.data(), 0),
            "pwmhip_query");
    return r;
  }
  uint64_t access(uint64_t i) const { return batch({ACCESS}, {i}, {0})[0]; }
  uint64_t rank(uint64_t symbol, uint64_t i) const { return batch({RANK}, {symbol}, {i})[0]; }
  uint64_t select(uint64_t symbol, uint64_t k) const { return batch({SELECT}, {symbol}, {k})[0]; }

 private:
  const wavelet_structure* ws_;
  pwmhip_rs* rs_ = nullptr;
};

namespace detail {

template <typename A>
wavelet_structure compute(const A* text, uint64_t size, uint64_t levels, uint32_t algo, bool tree,
                          context& ctx = context::shared()) {
  std::lock_guard<std::mutex> g(ctx.lock());
  pwmhip_ws ws{};
  check(pwmhip_build(ctx.get(), text, size, sizeof(A), uint32_t(levels), algo, tree ? PWMHIP_FLAG_TREE : 0u, &ws),
        "pwmhip_build");
  return wavelet_structure(ctx, ws);
}

// Traits shared by every adapter (wx_in_out_external<false, false>,
// include/construction/wx_base.hpp:13-23).
struct in_out_internal {
  static constexpr bool external_in = false;
  static constexpr bool external_out = false;
  static constexpr bool stats_support = false;
  static constexpr bool is_inplace = false;
};

}  // namespace detail

// The adapters: one per reference algorithm, same traits (is_parallel as
// the reference declares it in wx_<algo>.hpp) and signature.
#define PWM_HIP_ADAPTER(NAME, ALGO, PAR, HUFF)                                                       \
  template <typename AlphabetType, bool is_tree_>                                                    \
  class NAME : public detail::in_out_internal {                                                      \
   public:                                                                                           \
    static constexpr bool is_parallel = PAR;                                                         \
    static constexpr bool is_tree = is_tree_;                                                        \
    static constexpr uint8_t word_width = sizeof(AlphabetType);                                      \
    static constexpr bool is_huffman_shaped = HUFF;                                                  \
    template <typename InputType>                                                                    \
    static wavelet_structure compute(const InputType& text, const uint64_t size, const uint64_t levels) { \
      return detail::compute<AlphabetType>(text, size, levels, ALGO, is_tree_);                      \
    }                                                                                                \
  }

PWM_HIP_ADAPTER(wx_hip_pc, PWMHIP_ALGO_PC, false, false);             // wx_pc.hpp
PWM_HIP_ADAPTER(wx_hip_ps, PWMHIP_ALGO_PS, false, false);             // wx_ps.hpp
PWM_HIP_ADAPTER(wx_hip_ppc, PWMHIP_ALGO_PPC, true, false);           // wx_ppc.hpp:26-58
PWM_HIP_ADAPTER(wx_hip_pps, PWMHIP_ALGO_PPS, true, false);           // wx_pps.hpp:27-68
PWM_HIP_ADAPTER(wx_hip_pc_ss, PWMHIP_ALGO_PC_SS, false, false);       // wx_pc_ss.hpp
PWM_HIP_ADAPTER(wx_hip_ppc_ss, PWMHIP_ALGO_PPC_SS, true, false);     // wx_ppc_ss.hpp
PWM_HIP_ADAPTER(wx_hip_dd_pc, PWMHIP_ALGO_DD_PC, true, false);       // wx_dd_pc.hpp:42-119
PWM_HIP_ADAPTER(wx_hip_dd_ps, PWMHIP_ALGO_DD_PS, true, false);       // wx_dd_ps.hpp
PWM_HIP_ADAPTER(wx_hip_dd_pc_ss, PWMHIP_ALGO_DD_PC_SS, true, false); // wx_dd_pc_ss.hpp
PWM_HIP_ADAPTER(wx_hip_naive, PWMHIP_ALGO_NAIVE, false, false);       // wx_naive.hpp
PWM_HIP_ADAPTER(wx_hip_huff_pc, PWMHIP_ALGO_HUFF_PC, false, true);    // wx_huff_pc.hpp:27-80
PWM_HIP_ADAPTER(wx_hip_huff_ps, PWMHIP_ALGO_HUFF_PS, false, true);    // wx_huff_ps.hpp
PWM_HIP_ADAPTER(wx_hip_huff_pc_ss, PWMHIP_ALGO_HUFF_PC_SS, false, true);  // wx_huff_pc_ss.hpp
PWM_HIP_ADAPTER(wx_hip_huff_naive, PWMHIP_ALGO_HUFF_NAIVE, false, true);  // wx_huff_naive.hpp
PWM_HIP_ADAPTER(wx_hip_huff_dd_pc, PWMHIP_ALGO_HUFF_DD_PC, true, true);  // wx_huff_dd_pc.hpp

#undef PWM_HIP_ADAPTER

// Builds every text of a batch with one launch per phase (no counterpart in
// the reference).  Results equal wx_hip_<algo>::compute of each text.
template <typename AlphabetType>
std::vector<wavelet_structure> compute_batch(const std::vector<const AlphabetType*>& texts,
                                             const std::vector<uint64_t>& sizes, uint64_t levels, uint32_t algo,
                                             bool tree, context& ctx = context::shared()) {
  std::lock_guard<std::mutex> g(ctx.lock());
  std::vector<pwmhip_ws> outs(texts.size());
  std::vector<const void*> ptrs(texts.begin(), texts.end());
  check(pwmhip_build_batch(ctx.get(), ptrs.data(), sizes.data(), texts.size(), sizeof(AlphabetType),
                           uint32_t(levels), algo, tree ? PWMHIP_FLAG_TREE : 0u, outs.data()),
        "pwmhip_build_batch");
  // outs[0] owns the device block of the whole batch (plain shapes); every
  // structure then shares it, and the last one alive frees it
  std::shared_ptr<void> block;
  if (!outs.empty() && outs[0].alloc_bytes && (outs.size() == 1 || outs[1].alloc_bytes == 0)) {
    auto* owner = new pwmhip_ws(outs[0]);
    owner->code_len = owner->code_word = nullptr;
    owner->n_codes = 0;
    context* c = &ctx;
    block = std::shared_ptr<void>(owner, [c](void* p) {
      auto* w = static_cast<pwmhip_ws*>(p);
      pwmhip_ws_free(c->get(), w);
      delete w;
    });
    outs[0].alloc_bytes = 0;
  }
  std::vector<wavelet_structure> res;
  res.reserve(texts.size());
  for (auto& w : outs) res.emplace_back(ctx, w, block);
  return res;
}

}  // namespace pwm_hip

#endif  // PWM_HIP_WX_HPP
