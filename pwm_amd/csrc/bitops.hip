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

// Leaf histogram.  Up to 2^13 leaves the block counts in LDS first (global
// atomics on few addresses serialise); above that it adds in global memory.
constexpr uint32_t kLdsLeaves = 1u << 13;

template <typename T>
__global__ __launch_bounds__(256) void k_leaf_hist(const T* text, uint64_t n, uint32_t L,
                                                   unsigned long long* hist) {
  __shared__ uint32_t lh[kLdsLeaves];
  const bool local = L <= 13;
  const uint32_t K = local ? (1u << L) : 0u;
  for (uint32_t i = threadIdx.x; i < K; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  auto add = [&](uint64_t c) {
    if (L >= 64 || (c >> L) == 0) {
      if (local) atomicAdd(&lh[c], 1u);
      else atomicAdd(&hist[c], 1ull);
    }
  };
  // 16-byte loads (16 / sizeof(T) symbols each) when the text is aligned
  constexpr uint32_t PER = 16 / sizeof(T);
  const uint64_t gid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t nv = (reinterpret_cast<uintptr_t>(text) & 15) == 0 ? n / PER : 0;
  const uint4* tv = reinterpret_cast<const uint4*>(text);
  // four 16-byte loads issued together per round (one round trip, not four)
  constexpr int U = 4;
  for (uint64_t i = gid; i < nv; i += U * step) {
    union {
      uint4 v;
      T x[PER];
    } u[U];
#pragma unroll
    for (int r = 0; r < U; ++r)
      if (i + r * step < nv) u[r].v = tv[i + r * step];
#pragma unroll
    for (int r = 0; r < U; ++r)
      if (i + r * step < nv)
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) add(u[r].x[k]);
  }
  for (uint64_t i = nv * PER + gid; i < n; i += step) add(text[i]);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < K; i += blockDim.x)
    if (lh[i]) atomicAdd(&hist[i], (unsigned long long)lh[i]);
}

// Every tree level l >= 1 in one launch (blockIdx.y = l - 1), one thread per
// destination word; the word gathers the bits of every bucket run that
// overlaps it (runs sorted by destination).  tab[3 l], tab[3 l + 1],
// tab[3 l + 2]: where the level's run starts are in tab (destination starts,
// nruns + 1 of them, then as many source starts), its number of runs, and
// the matrix level's first word.
__global__ void k_gather_runs(uint64_t* lv, uint64_t stride, const uint64_t* words, const uint64_t* tab,
                              uint64_t nbits) {
  const uint32_t l = blockIdx.y + 1;
  const uint64_t nruns = tab[3 * l + 1];
  const uint64_t* dstart = tab + tab[3 * l];
  const uint64_t* sstart = dstart + nruns + 1;
  const uint64_t* src = words + tab[3 * l + 2];
  uint64_t* dst = lv + uint64_t(l - 1) * stride;
  const uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t nw = (nbits + 63) >> 6;
  if (w >= nw) return;
  const uint64_t ws = w << 6, we = min(ws + 64, nbits);
  // last run with dstart <= ws
  uint64_t lo = 0, hi = nruns - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (dstart[mid] <= ws) lo = mid;
    else hi = mid - 1;
  }
  uint64_t v = 0;
  for (uint64_t r = lo; r < nruns && dstart[r] < we; ++r) {
    const uint64_t d0 = dstart[r], d1 = dstart[r + 1];  // dstart[nruns] = nbits
    const uint64_t a = max(d0, ws), b = min(d1, we);
    if (a >= b) continue;
    const int64_t s = int64_t(sstart[r]) - int64_t(d0) + int64_t(ws);
    v |= window64(src, s, uint32_t(b - ws)) & range_mask(uint32_t(a - ws), uint32_t(b - ws));
  }
  dst[w] = v;
}

// The wide stitch's run tables for trees of up to 2^8 leaves, from the leaf
// histogram, in one wave (no host round trip).  Level l's runs are its
// non-empty l-bit prefixes p in numeric (tree) order; run r of level l is
// entry (1 << l) - 2 + r, with its tree start (the symbols whose prefix is
// below p) and its matrix start (the symbols whose prefix, read with its l
// bits reversed, is below p's reversal).  tab: [0, kMaxGroup) runs per
// level, then K tree starts, then K matrix starts.
constexpr uint32_t kTableLeaves = 256;

__device__ __forceinline__ uint64_t wave_excl_scan64(uint64_t v) {
  const uint32_t lane = threadIdx.x & 63;
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, 64);
    if (lane >= uint32_t(d)) x += y;
  }
  return x - v;
}

__global__ __launch_bounds__(64) void k_wt_tables(const unsigned long long* hist, uint32_t L, uint64_t* tab) {
  // prefix counts as a heap: level l's prefix p at pc[(1 << l) + p], leaves at level L
  __shared__ uint64_t pc[2 * kTableLeaves], ms[kTableLeaves];
  const uint32_t lane = threadIdx.x;
  const uint32_t K = 1u << L;
  for (uint32_t i = lane; i < K; i += 64) pc[K + i] = hist[i];
  for (uint32_t i = lane; i < kMaxGroup + 2 * K; i += 64) tab[i] = 0;
  __syncthreads();
  for (uint32_t l = L - 1; l >= 1; --l) {
    const uint32_t nb = 1u << l;
    for (uint32_t q = lane; q < nb; q += 64) pc[nb + q] = pc[2 * nb + 2 * q] + pc[2 * nb + 2 * q + 1];
    __syncthreads();
  }
  uint64_t* ne = tab;
  uint64_t* td = tab + kMaxGroup;
  uint64_t* ts = td + K;
  // each lane holds elements 2 lane and 2 lane + 1 of a level (at most 128)
  const uint32_t e0 = 2 * lane, e1 = 2 * lane + 1;
  for (uint32_t l = 1; l < L; ++l) {
    const uint32_t nb = 1u << l;
    const uint64_t* c = pc + nb;
    auto rev = [&](uint32_t q) { return __builtin_bitreverse32(q) >> (32 - l); };
    // matrix order: element q holds the prefix whose l bits reversed are q
    const uint64_t a0 = e0 < nb ? c[rev(e0)] : 0ull, a1 = e1 < nb ? c[rev(e1)] : 0ull;
    const uint64_t mx = wave_excl_scan64(a0 + a1);
    if (e0 < nb) ms[rev(e0)] = mx;
    if (e1 < nb) ms[rev(e1)] = mx + a0;
    // tree order, and the run index among the non-empty prefixes
    const uint64_t b0 = e0 < nb ? c[e0] : 0ull, b1 = e1 < nb ? c[e1] : 0ull;
    const uint64_t tx = wave_excl_scan64(b0 + b1);
    const uint32_t n0 = b0 ? 1u : 0u, n1 = b1 ? 1u : 0u;
    const uint64_t rx = wave_excl_scan64(n0 + n1);
    __syncthreads();  // ms is complete
    if (b0) {
      td[nb - 2 + rx] = tx;
      ts[nb - 2 + rx] = ms[e0];
    }
    if (b1) {
      td[nb - 2 + rx + n0] = tx + b0;
      ts[nb - 2 + rx + n0] = ms[e1];
    }
    if (lane == 63) ne[l] = rx + n0 + n1;
    __syncthreads();  // ms is read
  }
}

void host_tables_gather(pwmhip_ctx* ctx, const unsigned long long* d_hist, uint64_t n, uint32_t L, pwmhip_ws* ws,
                        uint64_t* d_tab, uint64_t* d_lv, uint64_t stride, uint64_t tab);

}  // namespace

void wm_to_wt_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width, uint32_t L,
                     pwmhip_ws* ws) {
  if (L > 24) throw status_error{PWMHIP_ERR_UNSUPPORTED, "tree above 24 levels"};
  if (L < 2 || n == 0) return;
  hipStream_t st = ctx->stream;
  const uint64_t K = 1ull << L;
  const uint64_t stride = level_stride(n);
  // scratch: leaf histogram, every level's run table, the reordered levels
  const uint64_t tab = 3 * uint64_t(L) + 2 * (K + 2 * uint64_t(L));
  // levels 1.. of the matrix (the sources) and of the output contiguous at
  // one stride, and at most kMaxGroup - 1 of them: the plain builds' wide
  // stitch moves the runs (tables in its layout: level l's entries from
  // 2^l - 2); else one thread per destination word (k_gather_runs)
  bool contiguous = true;
  for (uint32_t l = 2; l < L; ++l) contiguous &= ws->level_offset[l] == ws->level_offset[1] + (l - 1) * stride;
  const bool wide = contiguous && L - 1 < uint32_t(kMaxGroup);
  const uint64_t fw = wide ? wide_stitch_first_words(n, L - 1) : 0;
  ctx->scratch.reset();
  ctx->scratch.ensure((K + tab + uint64_t(L - 1) * stride) * 8 + fw * 4 + 4096);
  unsigned long long* d_hist = ctx->scratch.take<unsigned long long>(K);
  uint64_t* d_tab = ctx->scratch.take<uint64_t>(tab);
  uint64_t* d_lv = ctx->scratch.take<uint64_t>(uint64_t(L - 1) * stride);
  uint32_t* d_first = wide ? ctx->scratch.take<uint32_t>(fw) : nullptr;
  PWM_HIP_TRY(hipMemsetAsync(d_hist, 0, K * 8, st));
  // (a grid-stride grid that is resident at once: the kernel's registers
  // leave room for five waves a SIMD)
  const uint32_t grid = uint32_t(std::min<uint64_t>((n + 4095) / 4096, uint64_t(ctx->num_cus) * 4));
  switch (width) {
    case 1: hipLaunchKernelGGL(k_leaf_hist<uint8_t>, dim3(grid), dim3(256), 0, st, (const uint8_t*)d_text, n, L, d_hist); break;
    case 2: hipLaunchKernelGGL(k_leaf_hist<uint16_t>, dim3(grid), dim3(256), 0, st, (const uint16_t*)d_text, n, L, d_hist); break;
    case 4: hipLaunchKernelGGL(k_leaf_hist<uint32_t>, dim3(grid), dim3(256), 0, st, (const uint32_t*)d_text, n, L, d_hist); break;
    default: hipLaunchKernelGGL(k_leaf_hist<uint64_t>, dim3(grid), dim3(256), 0, st, (const uint64_t*)d_text, n, L, d_hist); break;
  }
  PWM_HIP_TRY(hipGetLastError());
  const uint64_t nw = words_for(n);
  if (wide) {
    // (wide: L <= kMaxGroup, so K <= kTableLeaves) the run tables on the device
    hipLaunchKernelGGL(k_wt_tables, dim3(1), dim3(64), 0, st, d_hist, L, d_tab);
    PWM_HIP_TRY(hipGetLastError());
    WideStitch w{};
    w.level_ne = d_tab;
    w.tab_dst = d_tab + kMaxGroup;
    w.tab_src = d_tab + kMaxGroup + K;
    w.first = d_first;
    w.src = ws->d_words + ws->level_offset[1];
    w.src_stride = stride;
    for (uint32_t j = 1; j < L; ++j) {
      w.dst[j] = d_lv + uint64_t(j - 1) * stride;
      w.nbits[j] = n;
    }
    w.units = 1;
    w.nsub = L - 1;
    launch_wide_stitch(ctx, w);
  } else {
    host_tables_gather(ctx, d_hist, n, L, ws, d_tab, d_lv, stride, tab);
  }
  // the tree levels replace the matrix levels in place (one copy when
  // levels 1.. are contiguous at one stride)
  if (contiguous) {
    PWM_HIP_TRY(hipMemcpyAsync(ws->d_words + ws->level_offset[1], d_lv, uint64_t(L - 1) * stride * 8,
                               hipMemcpyDeviceToDevice, st));
  } else {
    for (uint32_t l = 1; l < L; ++l)
      PWM_HIP_TRY(hipMemcpyAsync(ws->d_words + ws->level_offset[l], d_lv + uint64_t(l - 1) * stride, nw * 8,
                                 hipMemcpyDeviceToDevice, st));
  }
  PWM_HIP_TRY(hipStreamSynchronize(st));  // (the scratch is free again on return)
}

namespace {

// Trees whose levels are not one contiguous block, or of more than
// kMaxGroup levels: the run tables on the host from the leaf histogram,
// then one thread per destination word (k_gather_runs).
void host_tables_gather(pwmhip_ctx* ctx, const unsigned long long* d_hist, uint64_t n, uint32_t L, pwmhip_ws* ws,
                        uint64_t* d_tab, uint64_t* d_lv, uint64_t stride, uint64_t tab) {
  hipStream_t st = ctx->stream;
  const uint64_t K = 1ull << L;
  std::vector<uint64_t> pc(K);
  PWM_HIP_TRY(hipMemcpyAsync(pc.data(), d_hist, K * 8, hipMemcpyDeviceToHost, st));
  PWM_HIP_TRY(hipStreamSynchronize(st));

  // run tables of every level l >= 1, in destination (tree) order: runs of
  // the non-empty l-bit prefixes p, from matrix position (prefixes in
  // bit-reversed order) to tree position (numeric order)
  std::vector<uint64_t> all;  // a header of 3 words per level, then per level: ds[0..nruns], ss[0..nruns]
  std::vector<uint64_t> off(L), nr(L);
  all.reserve(tab);
  all.assign(3 * uint64_t(L), 0);
  for (uint32_t l = L - 1; l >= 1; --l) {
    std::vector<uint64_t> up(1ull << l);
    for (uint64_t q = 0; q < up.size(); ++q) up[q] = pc[2 * q] + pc[2 * q + 1];
    pc.swap(up);
    const uint64_t nb = 1ull << l;
    std::vector<uint64_t> wm_start(nb);
    uint64_t acc = 0;
    for (uint64_t q = 0; q < nb; ++q) {  // matrix position order: prefix reverse_l(q)
      uint64_t p = 0;
      for (uint32_t i = 0; i < l; ++i) p |= ((q >> i) & 1) << (l - 1 - i);
      wm_start[p] = acc;
      acc += pc[p];
    }
    std::vector<uint64_t> ds, ss;
    acc = 0;
    for (uint64_t p = 0; p < nb; ++p) {
      if (pc[p]) {
        ds.push_back(acc);
        ss.push_back(wm_start[p]);
      }
      acc += pc[p];
    }
    ds.push_back(n);
    ss.push_back(0);
    off[l] = all.size();
    nr[l] = ds.size() - 1;
    all.insert(all.end(), ds.begin(), ds.end());
    all.insert(all.end(), ss.begin(), ss.end());
  }
  for (uint32_t l = 1; l < L; ++l) {
    all[3 * l] = off[l];
    all[3 * l + 1] = nr[l];
    all[3 * l + 2] = ws->level_offset[l];
  }
  PWM_HIP_TRY(hipMemcpyAsync(d_tab, all.data(), all.size() * 8, hipMemcpyHostToDevice, st));
  const uint64_t nw = words_for(n);
  hipLaunchKernelGGL(k_gather_runs, dim3(uint32_t((nw + 255) / 256), L - 1), dim3(256), 0, st, d_lv, stride,
                     ws->d_words, d_tab, n);
  PWM_HIP_TRY(hipGetLastError());
  PWM_HIP_TRY(hipStreamSynchronize(st));  // (the host run table)
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
