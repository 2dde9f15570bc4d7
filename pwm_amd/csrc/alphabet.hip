// Alphabet reduction on the device: every symbol is replaced by the rank of
// its first occurrence (0, 1, 2, ...), and the number of levels of the
// reduced text is returned.  Reference: util/alphabet_util.hpp:18-45
// (reduce_alphabet) and :99-106 (levels_for_max_char); unlike the
// reference's 255-entry table for bytes, byte value 255 is handled.
//
// Widths 1 and 2: a direct table of first positions (atomicMin).  Widths 4
// and 8: an open-addressing hash table (key + 1 in a 64-bit slot, linear
// probing), compacted on the device; the distinct symbols are ordered by
// first position on the host and their codes written back into the table.
#include <algorithm>
#include <vector>

#include "pwm_internal.hpp"

namespace pwm {
namespace {

template <typename T>
__global__ void k_first_direct(const T* text, uint64_t n, unsigned long long* first) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    atomicMin(&first[text[i]], (unsigned long long)i);
}

template <typename T>
__global__ void k_map_direct(T* text, uint64_t n, const uint64_t* code) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    text[i] = T(code[text[i]]);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Slot of symbol x (inserted if absent).  Keys are stored as x + 1 (0 =
// empty); the all-ones 64-bit symbol uses the extra slot `cap`.
template <typename T>
__device__ __forceinline__ uint64_t hash_slot(unsigned long long* keys, uint64_t mask, uint64_t x, bool insert) {
  if (sizeof(T) == 8 && x == ~0ull) return mask + 1;
  const unsigned long long k = x + 1;
  uint64_t h = mix64(x) & mask;
  for (;;) {
    const unsigned long long cur = keys[h];
    if (cur == k) return h;
    if (cur == 0) {
      if (!insert) return ~0ull;
      const unsigned long long old = atomicCAS(&keys[h], 0ull, k);
      if (old == 0 || old == k) return h;
    }
    h = (h + 1) & mask;
  }
}

template <typename T>
__global__ void k_first_hash(const T* text, uint64_t n, unsigned long long* keys, unsigned long long* first,
                             uint64_t mask) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t h = hash_slot<T>(keys, mask, uint64_t(text[i]), true);
    atomicMin(&first[h], (unsigned long long)i);
  }
}

// Occupied slots -> (first position, slot) pairs.
__global__ void k_compact(const unsigned long long* first, uint64_t slots, unsigned long long* count,
                          uint64_t* pairs) {
  for (uint64_t h = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; h < slots; h += uint64_t(gridDim.x) * blockDim.x)
    if (first[h] != ~0ull) {
      const unsigned long long k = atomicAdd(count, 1ull);
      pairs[2 * k] = first[h];
      pairs[2 * k + 1] = h;
    }
}

__global__ void k_set_codes(const uint64_t* slot_code, uint64_t d, uint64_t* code) {
  const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < d) code[slot_code[2 * k]] = slot_code[2 * k + 1];
}

template <typename T>
__global__ void k_map_hash(T* text, uint64_t n, unsigned long long* keys, const uint64_t* code, uint64_t mask) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    text[i] = T(code[hash_slot<T>(keys, mask, uint64_t(text[i]), false)]);
}

uint64_t levels_for(uint64_t distinct) {
  uint64_t max_char = distinct ? distinct - 1 : 0, levels = 0;
  while (max_char) {
    max_char >>= 1;
    ++levels;
  }
  return levels;
}

template <typename T>
uint64_t reduce_direct(pwmhip_ctx* ctx, T* d_text, uint64_t n) {
  hipStream_t s = ctx->stream;
  const uint64_t K = uint64_t(1) << (8 * sizeof(T));
  ctx->scratch.reset();
  ctx->scratch.ensure(2 * K * 8 + 1024);
  auto* first = ctx->scratch.take<unsigned long long>(K);
  auto* code = ctx->scratch.take<uint64_t>(K);
  PWM_HIP_TRY(hipMemsetAsync(first, 0xff, K * 8, s));
  const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, uint64_t(ctx->num_cus) * 16));
  hipLaunchKernelGGL(k_first_direct<T>, dim3(grid), dim3(256), 0, s, d_text, n, first);
  std::vector<uint64_t> f(K);
  PWM_HIP_TRY(hipMemcpyAsync(f.data(), first, K * 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipStreamSynchronize(s));
  std::vector<uint64_t> sym;
  for (uint64_t c = 0; c < K; ++c)
    if (f[c] != ~0ull) sym.push_back(c);
  std::sort(sym.begin(), sym.end(), [&](uint64_t a, uint64_t b) { return f[a] < f[b]; });
  std::vector<uint64_t> h_code(K, 0);
  for (uint64_t r = 0; r < sym.size(); ++r) h_code[sym[r]] = r;
  PWM_HIP_TRY(hipMemcpyAsync(code, h_code.data(), K * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_map_direct<T>, dim3(grid), dim3(256), 0, s, d_text, n, code);
  PWM_HIP_TRY(hipGetLastError());
  PWM_HIP_TRY(hipStreamSynchronize(s));
  return levels_for(sym.size());
}

template <typename T>
uint64_t reduce_hash(pwmhip_ctx* ctx, T* d_text, uint64_t n) {
  hipStream_t s = ctx->stream;
  uint64_t cap = 1024;
  const uint64_t want = std::min<uint64_t>(2 * n, uint64_t(1) << (sizeof(T) * 8 > 40 ? 40 : sizeof(T) * 8 + 1));
  while (cap < want) cap <<= 1;
  const uint64_t slots = cap + 1;  // + the slot of the all-ones symbol
  ctx->scratch.reset();
  ctx->scratch.ensure(slots * 8 * 3 + 2 * n * 8 + 4096);
  auto* keys = ctx->scratch.take<unsigned long long>(slots);
  auto* first = ctx->scratch.take<unsigned long long>(slots);
  auto* code = ctx->scratch.take<uint64_t>(slots);
  auto* count = ctx->scratch.take<unsigned long long>(1);
  auto* pairs = ctx->scratch.take<uint64_t>(2 * std::min<uint64_t>(n, slots));
  PWM_HIP_TRY(hipMemsetAsync(keys, 0, slots * 8, s));
  PWM_HIP_TRY(hipMemsetAsync(first, 0xff, slots * 8, s));
  PWM_HIP_TRY(hipMemsetAsync(count, 0, 8, s));
  const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, uint64_t(ctx->num_cus) * 16));
  hipLaunchKernelGGL(k_first_hash<T>, dim3(grid), dim3(256), 0, s, d_text, n, keys, first, cap - 1);
  const uint32_t cgrid = uint32_t(std::min<uint64_t>((slots + 255) / 256, uint64_t(ctx->num_cus) * 16));
  hipLaunchKernelGGL(k_compact, dim3(cgrid), dim3(256), 0, s, first, slots, count, pairs);
  uint64_t d = 0;
  PWM_HIP_TRY(hipMemcpyAsync(&d, count, 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipStreamSynchronize(s));
  std::vector<uint64_t> pr(2 * d);
  if (d) PWM_HIP_TRY(hipMemcpy(pr.data(), pairs, 2 * d * 8, hipMemcpyDeviceToHost));
  // order by first position; rewrite as (slot, code)
  std::vector<uint64_t> idx(d);
  for (uint64_t k = 0; k < d; ++k) idx[k] = k;
  std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return pr[2 * a] < pr[2 * b]; });
  std::vector<uint64_t> sc(2 * d);
  for (uint64_t r = 0; r < d; ++r) {
    sc[2 * r] = pr[2 * idx[r] + 1];
    sc[2 * r + 1] = r;
  }
  if (d) {
    PWM_HIP_TRY(hipMemcpyAsync(pairs, sc.data(), 2 * d * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_set_codes, dim3(uint32_t((d + 255) / 256)), dim3(256), 0, s, pairs, d, code);
    hipLaunchKernelGGL(k_map_hash<T>, dim3(grid), dim3(256), 0, s, d_text, n, keys, code, cap - 1);
  }
  PWM_HIP_TRY(hipGetLastError());
  PWM_HIP_TRY(hipStreamSynchronize(s));
  return levels_for(d);
}

}  // namespace
}  // namespace pwm

extern "C" int pwmhip_reduce_alphabet(pwmhip_ctx* ctx, void* d_text, uint64_t size, uint32_t word_width,
                                      uint64_t* levels_out) {
  if (!ctx || !levels_out || (size && !d_text)) return PWMHIP_ERR_INVALID;
  if (word_width != 1 && word_width != 2 && word_width != 4 && word_width != 8) return PWMHIP_ERR_INVALID;
  *levels_out = 0;
  if (size == 0) return PWMHIP_OK;
  try {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    switch (word_width) {
      case 1: *levels_out = pwm::reduce_direct(ctx, static_cast<uint8_t*>(d_text), size); break;
      case 2: *levels_out = pwm::reduce_direct(ctx, static_cast<uint16_t*>(d_text), size); break;
      case 4: *levels_out = pwm::reduce_hash(ctx, static_cast<uint32_t*>(d_text), size); break;
      default: *levels_out = pwm::reduce_hash(ctx, static_cast<uint64_t*>(d_text), size); break;
    }
  } catch (const pwm::hip_failure& f) {
    (void)hipGetLastError();
    return f.status;
  } catch (const pwm::status_error& e) {
    return e.status;
  } catch (...) {
    return PWMHIP_ERR_OOM;
  }
  return PWMHIP_OK;
}
