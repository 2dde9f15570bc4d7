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
  for (uint64_t i = gid; i < nv; i += step) {
    union {
      uint4 v;
      T x[PER];
    } u;
    u.v = tv[i];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) add(u.x[k]);
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
  // the wide stitch's layout: entries per level, destination starts, source starts
  std::vector<uint64_t> wtab(wide ? kMaxGroup + 2 * K : 0, 0);
  uint64_t* w_ne = wtab.data();
  uint64_t* w_dst = w_ne + kMaxGroup;
  uint64_t* w_src = w_dst + K;
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
    if (wide) {
      w_ne[l] = nr[l];
      for (uint64_t r = 0; r < nr[l]; ++r) {
        w_dst[(1ull << l) - 2 + r] = ds[r];
        w_src[(1ull << l) - 2 + r] = ss[r];
      }
      continue;
    }
    all.insert(all.end(), ds.begin(), ds.end());
    all.insert(all.end(), ss.begin(), ss.end());
  }
  const uint64_t nw = words_for(n);
  if (wide) {
    PWM_HIP_TRY(hipMemcpyAsync(d_tab, wtab.data(), wtab.size() * 8, hipMemcpyHostToDevice, st));
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
    for (uint32_t l = 1; l < L; ++l) {
      all[3 * l] = off[l];
      all[3 * l + 1] = nr[l];
      all[3 * l + 2] = ws->level_offset[l];
    }
    PWM_HIP_TRY(hipMemcpyAsync(d_tab, all.data(), all.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gather_runs, dim3(uint32_t((nw + 255) / 256), L - 1), dim3(256), 0, st, d_lv, stride,
                       ws->d_words, d_tab, n);
    PWM_HIP_TRY(hipGetLastError());
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
  PWM_HIP_TRY(hipStreamSynchronize(st));  // (the host run table)
}

void copy_bits_device(hipStream_t st, uint64_t* dst, const uint64_t* src, uint64_t dst_off,
                      uint64_t src_off, uint64_t len) {
  if (len == 0) return;
  const uint64_t nwords = ((dst_off + len - 1) >> 6) - (dst_off >> 6) + 1;
  const uint32_t blocks = uint32_t((nwords + 255) / 256);
  hipLaunchKernelGGL(k_copy_bits, dim3(blocks), dim3(256), 0, st, dst, src, dst_off, src_off, len, nwords);
  PWM_HIP_TRY(hipGetLastError());
}

}  // namespace pwm
