// Entry points whose device implementation is added in later files.
#include "pwm_internal.hpp"
namespace pwm {
void build_huffman_device(pwmhip_ctx*, const void*, uint64_t, uint32_t, bool, bool, pwmhip_ws*) {
  throw status_error{PWMHIP_ERR_UNSUPPORTED, "huffman"};
}
void wm_to_wt_device(pwmhip_ctx*, const void*, uint64_t, uint32_t, uint32_t, pwmhip_ws*) {
  throw status_error{PWMHIP_ERR_UNSUPPORTED, "tree > 8 levels"};
}
}  // namespace pwm
extern "C" {
int pwmhip_rs_build(pwmhip_ctx*, const pwmhip_ws*, pwmhip_rs**) { return PWMHIP_ERR_UNSUPPORTED; }
int pwmhip_rs_free(pwmhip_ctx*, pwmhip_rs*) { return PWMHIP_ERR_UNSUPPORTED; }
int pwmhip_bv_query(pwmhip_ctx*, const pwmhip_rs*, const uint32_t*, const uint32_t*, const uint64_t*, uint64_t,
                    uint64_t*, uint32_t) { return PWMHIP_ERR_UNSUPPORTED; }
int pwmhip_query(pwmhip_ctx*, const pwmhip_ws*, const pwmhip_rs*, const uint32_t*, const uint64_t*,
                 const uint64_t*, uint64_t, uint64_t*, uint32_t) { return PWMHIP_ERR_UNSUPPORTED; }
int pwmhip_reduce_alphabet(pwmhip_ctx*, void*, uint64_t, uint32_t, uint64_t*) { return PWMHIP_ERR_UNSUPPORTED; }
int pwmhip_build_batch(pwmhip_ctx*, const void* const*, const uint64_t*, uint64_t, uint32_t, uint32_t, uint32_t,
                       uint32_t, pwmhip_ws*) { return PWMHIP_ERR_UNSUPPORTED; }
}
