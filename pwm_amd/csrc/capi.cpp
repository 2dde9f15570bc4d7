// pwm_amd/csrc/capi.cpp -- the C-ABI entry points (include/pwm_hip.h).
//
// Each entry point translates exceptions into status codes, so nothing
// throws across the ABI.  The algorithm classes of the reference
// (include/wx_*.hpp) map onto pwmhip_build with an algorithm id; they differ
// only in zeros[levels-1] (see pwm_hip.h).

#include <cstdlib>
#include <cstring>
#include <new>

#include "pwm_internal.hpp"

namespace pwm {

void arena::ensure(size_t bytes) {
  if (bytes <= cap) return;
  release();
  void* p = nullptr;
  PWM_HIP_TRY(hipMalloc(&p, bytes));
  base = p;
  cap = bytes;
  used = 0;
}

void arena::release() {
  if (base) (void)hipFree(base);
  base = nullptr;
  cap = 0;
  used = 0;
}

namespace {

int run_guarded(const char* /*what*/, auto&& fn) { return guarded(fn); }

bool is_huffman_algo(uint32_t algo) { return algo >= PWMHIP_ALGO_HUFF_PC && algo <= PWMHIP_ALGO_HUFF_DD_PC; }

}  // namespace

}  // namespace pwm

extern "C" {

const char* pwmhip_status_string(int status) {
  switch (status) {
    case PWMHIP_OK: return "ok";
    case PWMHIP_ERR_INVALID: return "invalid argument";
    case PWMHIP_ERR_NO_DEVICE: return "no HIP device";
    case PWMHIP_ERR_OOM: return "out of device memory";
    case PWMHIP_ERR_HIP: return "HIP runtime error";
    case PWMHIP_ERR_SYMBOL_RANGE: return "symbol does not fit into the requested levels";
    case PWMHIP_ERR_UNSUPPORTED: return "unsupported input (undefined in the reference)";
    case PWMHIP_ERR_BUDGET: return "request exceeds the device budget";
    case PWMHIP_ERR_TIMEOUT: return "wait timed out";
    default: return "unknown status";
  }
}

int pwmhip_device_count(int* count) {
  if (!count) return PWMHIP_ERR_INVALID;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return PWMHIP_ERR_NO_DEVICE;
  }
  *count = c;
  return PWMHIP_OK;
}

int pwmhip_create(pwmhip_ctx** out, int device) {
  if (!out) return PWMHIP_ERR_INVALID;
  *out = nullptr;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || device < 0 || device >= c) {
    (void)hipGetLastError();
    return PWMHIP_ERR_NO_DEVICE;
  }
  auto* ctx = new (std::nothrow) pwmhip_ctx();
  if (!ctx) return PWMHIP_ERR_OOM;
  int st = pwm::run_guarded("create", [&] {
    ctx->device = device;
    PWM_HIP_TRY(hipSetDevice(device));
    int cus = 0;
    PWM_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    ctx->num_cus = cus > 0 ? cus : 256;
    PWM_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (auto& e : ctx->ev) PWM_HIP_TRY(hipEventCreate(&e));
    // the allocation-free build's static scratch on this device; a failure
    // here only disables that entry point (it then reports UNSUPPORTED)
    (void)pwm::guarded([&] { pwm::noalloc_setup(ctx); });
  });
  if (st != PWMHIP_OK) {
    pwmhip_destroy(ctx);
    return st;
  }
  *out = ctx;
  return PWMHIP_OK;
}

int pwmhip_destroy(pwmhip_ctx* ctx) {
  if (!ctx) return PWMHIP_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  ctx->scratch.release();
  ctx->text_scratch.release();
  ctx->rs_scratch.release();
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->d_lut) (void)hipFree(ctx->d_lut);
  if (ctx->hstage) (void)hipHostFree(ctx->hstage);
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->sev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PWMHIP_OK;
}

int pwmhip_build(pwmhip_ctx* ctx, const void* text, uint64_t size, uint32_t word_width,
                 uint32_t levels, uint32_t algo, uint32_t flags, pwmhip_ws* out) {
  if (!ctx || !out) return PWMHIP_ERR_INVALID;
  // An earlier result is taken over only when it owns its block: a view into
  // a batch's block (alloc_bytes == 0) is left to the batch's owner.
  const bool take = (flags & PWMHIP_FLAG_REUSE_OUTPUT) && out->d_words && out->alloc_bytes;
  const bool huff = pwm::is_huffman_algo(algo);
  const bool stream = (flags & PWMHIP_FLAG_STREAM) && !(flags & PWMHIP_FLAG_TEXT_ON_DEVICE) && size > 0;
  int bad = PWMHIP_OK;
  if (word_width != 1 && word_width != 2 && word_width != 4 && word_width != 8) bad = PWMHIP_ERR_INVALID;
  else if (algo > PWMHIP_ALGO_HUFF_DD_PC) bad = PWMHIP_ERR_INVALID;
  else if (!huff && (levels == 0 || levels > 8 * word_width || levels > PWMHIP_MAX_LEVELS)) bad = PWMHIP_ERR_INVALID;
  else if (size > 0 && !text) bad = PWMHIP_ERR_INVALID;
  else if (stream && huff) bad = PWMHIP_ERR_UNSUPPORTED;
  if (bad != PWMHIP_OK) {
    // rejected before anything is taken over: an earlier result stays the
    // caller's (and valid); anything else is cleared
    if (!take) std::memset(out, 0, sizeof(*out));
    return bad;
  }
  uint64_t* reuse_ptr = take ? out->d_words : nullptr;
  const uint64_t reuse_bytes = take ? out->alloc_bytes : 0;
  if (flags & PWMHIP_FLAG_REUSE_OUTPUT) {
    std::free(out->code_len);
    std::free(out->code_word);
  }
  std::memset(out, 0, sizeof(*out));
  const bool tree = (flags & PWMHIP_FLAG_TREE) != 0;
  out->is_tree = tree;
  out->is_huffman = huff;
  const int st = pwm::run_guarded("build", [&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    const void* d_text = text;
    if (size > 0 && !(flags & PWMHIP_FLAG_TEXT_ON_DEVICE) && !stream) {
      ctx->text_scratch.ensure(size * word_width + 64);
      PWM_HIP_TRY(hipMemcpyAsync(ctx->text_scratch.base, text, size * word_width, hipMemcpyHostToDevice,
                                 ctx->stream));
      d_text = ctx->text_scratch.base;
    }
    out->text_size = size;
    if (reuse_ptr && size == 0) {
      (void)hipFree(reuse_ptr);
      reuse_ptr = nullptr;
    }
    if (huff) {
      // huff_pc's border bookkeeping (huff_pc.hpp:19-83) is shared by
      // huff_ps and huff_dd_pc (the latter at one thread: with more its
      // output depends on the thread count); huff_pc_ss (huff_pc_ss.hpp:
      // 18-121) gives huff_naive's levels with zeros[levels-1] left at 0.
      const bool quirks = algo == PWMHIP_ALGO_HUFF_PC || algo == PWMHIP_ALGO_HUFF_PS ||
                          algo == PWMHIP_ALGO_HUFF_DD_PC;
      uint64_t* r = reuse_ptr;
      reuse_ptr = nullptr;  // (the Huffman build takes it over or frees it)
      pwm::build_huffman_device(ctx, d_text, size, word_width, tree, quirks, out, r, reuse_bytes);
      if (algo == PWMHIP_ALGO_HUFF_PC_SS && out->n_zeros) out->zeros[out->n_zeros - 1] = 0;
      return;
    }
    if (size == 0) {  // wx_ppc.hpp:39-45: empty structure
      out->levels = 0;
      return;
    }
    std::vector<uint64_t> bits(levels, size);
    uint64_t* r = reuse_ptr;
    reuse_ptr = nullptr;  // (alloc_structure takes it over or frees it)
    pwm::alloc_structure(out, levels, bits.data(), r, reuse_bytes);
    try {
      pwm::build_plain_device(ctx, stream ? nullptr : d_text, size, word_width, levels, tree,
                              (flags & PWMHIP_FLAG_CHECK_RANGE) != 0, out, stream ? text : nullptr);
    } catch (...) {
      (void)hipFree(out->d_words);
      out->d_words = nullptr;
      out->alloc_bytes = 0;
      throw;
    }
    if (!tree && pwm::zeros_last_is_zero(algo)) out->zeros[levels - 1] = 0;
  });
  if (reuse_ptr) (void)hipFree(reuse_ptr);  // failed before the output took it over
  if (st != PWMHIP_OK) {
    // nothing half-built is handed back
    if (out->d_words && out->alloc_bytes) (void)hipFree(out->d_words);
    std::free(out->code_len);
    std::free(out->code_word);
    std::memset(out, 0, sizeof(*out));
  }
  return st;
}

int pwmhip_build_batch(pwmhip_ctx* ctx, const void* const* texts, const uint64_t* sizes, uint64_t count,
                       uint32_t word_width, uint32_t levels, uint32_t algo, uint32_t flags, pwmhip_ws* outs) {
  if (!ctx || (count && (!texts || !sizes || !outs))) return PWMHIP_ERR_INVALID;
  // arguments first: a rejected call takes nothing over and frees nothing
  if (word_width != 1 && word_width != 2 && word_width != 4 && word_width != 8) return PWMHIP_ERR_INVALID;
  if (algo > PWMHIP_ALGO_HUFF_DD_PC) return PWMHIP_ERR_INVALID;
  const bool huff = pwm::is_huffman_algo(algo);
  const bool tree = (flags & PWMHIP_FLAG_TREE) != 0;
  if (!huff && (levels == 0 || levels > 8 * word_width || levels > PWMHIP_MAX_LEVELS)) return PWMHIP_ERR_INVALID;
  for (uint64_t t = 0; t < count; ++t)
    if (sizes[t] > 0 && !texts[t]) return PWMHIP_ERR_INVALID;
  if (count == 0) return PWMHIP_OK;
  void* reuse_ptr = nullptr;
  uint64_t reuse_bytes = 0;
  if ((flags & PWMHIP_FLAG_REUSE_OUTPUT) && outs[0].d_words && outs[0].alloc_bytes) {
    reuse_ptr = outs[0].d_words;
    reuse_bytes = outs[0].alloc_bytes;
  }
  for (uint64_t t = 0; t < count; ++t) {
    std::free(outs[t].code_len);
    std::free(outs[t].code_word);
    // the fixed header fields only: level arrays are written up to `levels`
    // by the build and zeros up to n_zeros (a full clear of 20,000 structures
    // is tens of megabytes)
    pwmhip_ws& w = outs[t];
    w.levels = 0;
    w.text_size = 0;
    w.is_tree = 0;
    w.is_huffman = 0;
    w.d_words = nullptr;
    w.n_zeros = 0;
    w.n_codes = 0;
    w.code_len = nullptr;
    w.code_word = nullptr;
    w.alloc_bytes = 0;
  }
  if (huff || levels > uint32_t(pwm::kMaxGroup)) {
    if (reuse_ptr) {
      (void)hipSetDevice(ctx->device);
      (void)hipFree(reuse_ptr);
    }
    // shapes the batched kernels do not cover: one build per text, each
    // structure owning its memory
    for (uint64_t t = 0; t < count; ++t) {
      const int st = pwmhip_build(ctx, texts[t], sizes[t], word_width, levels, algo,
                                  flags & ~uint32_t(PWMHIP_FLAG_REUSE_OUTPUT), &outs[t]);
      if (st != PWMHIP_OK) {
        for (uint64_t u = 0; u < t; ++u) pwmhip_ws_free(ctx, &outs[u]);
        return st;
      }
    }
    return PWMHIP_OK;
  }
  return pwm::run_guarded("build_batch", [&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    std::vector<const void*> d_texts(texts, texts + count);
    if (!(flags & PWMHIP_FLAG_TEXT_ON_DEVICE)) {
      // stage every host text at a 256-byte aligned offset of one buffer
      std::vector<uint64_t> off(count);
      uint64_t total = 0;
      for (uint64_t t = 0; t < count; ++t) {
        off[t] = total;
        total += (sizes[t] * word_width + 255) & ~uint64_t(255);
      }
      ctx->text_scratch.ensure(total + 64);
      for (uint64_t t = 0; t < count; ++t) {
        uint8_t* d = static_cast<uint8_t*>(ctx->text_scratch.base) + off[t];
        if (sizes[t])
          PWM_HIP_TRY(hipMemcpyAsync(d, texts[t], sizes[t] * word_width, hipMemcpyHostToDevice, ctx->stream));
        d_texts[t] = d;
      }
    }
    pwm::build_batch_device(ctx, d_texts.data(), sizes, count, word_width, levels, tree,
                            (flags & PWMHIP_FLAG_CHECK_RANGE) != 0, outs, reuse_ptr, reuse_bytes, flags);
    if (!tree && pwm::zeros_last_is_zero(algo))
      for (uint64_t t = 0; t < count; ++t)
        if (outs[t].levels) outs[t].zeros[levels - 1] = 0;
  });
}

int pwmhip_ws_free(pwmhip_ctx* ctx, pwmhip_ws* ws) {
  if (!ctx || !ws) return PWMHIP_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  // a structure with alloc_bytes == 0 is a view into a batch's memory
  if (ws->d_words && ws->alloc_bytes) (void)hipFree(ws->d_words);
  std::free(ws->code_len);
  std::free(ws->code_word);
  std::memset(ws, 0, sizeof(*ws));
  return PWMHIP_OK;
}

int pwmhip_ws_download(pwmhip_ctx* ctx, const pwmhip_ws* ws, uint64_t level, uint64_t* host_words) {
  if (!ctx || !ws || level >= ws->levels || !host_words) return PWMHIP_ERR_INVALID;
  return pwm::run_guarded("download", [&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    const uint64_t nw = pwm::words_for(ws->level_bits[level]);
    if (nw == 0) return;
    PWM_HIP_TRY(hipMemcpyAsync(host_words, ws->d_words + ws->level_offset[level], nw * 8,
                               hipMemcpyDeviceToHost, ctx->stream));
    PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  });
}

int pwmhip_last_timing(pwmhip_ctx* ctx, double* ms4) {
  if (!ctx || !ms4) return PWMHIP_ERR_INVALID;
  for (int i = 0; i < 4; ++i) ms4[i] = ctx->last_ms[i];
  return PWMHIP_OK;
}

int pwmhip_copy_bits(pwmhip_ctx* ctx, uint64_t* d_dst, const uint64_t* d_src, uint64_t dst_off,
                     uint64_t src_off, uint64_t len) {
  if (!ctx || (len && (!d_dst || !d_src))) return PWMHIP_ERR_INVALID;
  return pwm::run_guarded("copy_bits", [&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    pwm::copy_bits_device(ctx->stream, d_dst, d_src, dst_off, src_off, len);
    PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
