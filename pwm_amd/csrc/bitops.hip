// pwm_amd/csrc/bitops.hip -- device bit-range copy (the copy_bits analogue,
// reference construction/merge.hpp:19-190) and the matrix -> tree bucket
// reordering used for trees with more levels than one construction pass.

#include "pwm_internal.hpp"

namespace pwm {

namespace {

// 64 MSB-first bits starting at bit s (s > -64), reading only words that
// hold bits of [s, s + need).
__device__ __forceinline__ uint64_t window64(const uint64_t* a, int64_t s, uint32_t need) {
  const int64_t w = s >> 6;
  const uint32_t sh = uint32_t(s & 63);
  const uint64_t w0 = w >= 0 ? a[w] : 0ull;
  if (sh == 0 || sh + need <= 64) return w0 << sh;
  const uint64_t w1 = a[w + 1];
  return (w0 << sh) | (w1 >> (64 - sh));
}

__device__ __forceinline__ uint64_t range_mask(uint32_t a, uint32_t b) {
  const uint64_t hi = a == 0 ? ~0ull : (~0ull >> a);
  const uint64_t lo = b >= 64 ? ~0ull : ~(~0ull >> b);
  return hi & lo;
}

// One thread per destination word of [dst_off, dst_off + len).  Partially
// covered words are OR-ed (as copy_bits does), covered words overwritten.
__global__ void k_copy_bits(uint64_t* dst, const uint64_t* src, uint64_t dst_off, uint64_t src_off,
                            uint64_t len, uint64_t nwords) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  const uint64_t w = (dst_off >> 6) + i;
  const uint64_t ws = w << 6;
  const uint64_t a = ws > dst_off ? ws : dst_off;
  const uint64_t e = dst_off + len;
  const uint64_t b = (ws + 64) < e ? (ws + 64) : e;
  // source bit that lands on destination bit ws
  const int64_t s = int64_t(src_off) - int64_t(dst_off) + int64_t(ws);
  // read only source bits of the range: positions before src_off are masked
  uint64_t v = window64(src, s, uint32_t(b - ws)) & range_mask(uint32_t(a - ws), uint32_t(b - ws));
  if (a == ws && b == ws + 64) dst[w] = v;
  else dst[w] |= v;  // at most the first and the last word; distinct threads
}

}  // namespace

void copy_bits_device(hipStream_t st, uint64_t* dst, const uint64_t* src, uint64_t dst_off,
                      uint64_t src_off, uint64_t len) {
  if (len == 0) return;
  const uint64_t nwords = ((dst_off + len - 1) >> 6) - (dst_off >> 6) + 1;
  const uint32_t blocks = uint32_t((nwords + 255) / 256);
  hipLaunchKernelGGL(k_copy_bits, dim3(blocks), dim3(256), 0, st, dst, src, dst_off, src_off, len, nwords);
  PWM_HIP_TRY(hipGetLastError());
}

}  // namespace pwm
