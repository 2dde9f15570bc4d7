// Entry points whose device implementation is added in later files.
#include "pwm_internal.hpp"
namespace pwm {
}  // namespace pwm
extern "C" {
int pwmhip_reduce_alphabet(pwmhip_ctx*, void*, uint64_t, uint32_t, uint64_t*) { return PWMHIP_ERR_UNSUPPORTED; }
}
