// pwm_amd/csrc/huffman.hip -- Huffman-shaped wavelet matrices and trees
// (the wx_huff_* algorithms, reference include/wx_huff_pc.hpp:45-80).
//
// Level l of a Huffman-shaped structure holds the occurrences whose code is
// longer than l.  Every occurrence with code length > l belongs to exactly one
// internal node (its l-bit code prefix); its bit at level l lands at
//     start_l[prefix] + (earlier occurrences in text order with that prefix)
// where start_l comes from one of two host-side plans:
//   * reference huff_pc bookkeeping (huffman/huff_pc.hpp:19-83,
//     huff_building_blocks.hpp:186-222): borders assigned by key position,
//     kept across levels, zeros[l-1] = borders[1];
//   * naive semantics (huffman/huff_naive.hpp:18-126): matrix buckets in
//     bit-reversed prefix order, tree nodes in numeric order, true zeros.
// The code table follows canonical_huff_codes (huffman/huff_codes.hpp:
// 141-285) with the same standard-library algorithms on the same input order,
// so ties resolve exactly as in the reference.
//
// Device work: histogram; per chunk (one wave, 4096 symbols) counts per
// internal node and chunk-local level bits ordered by node; scan of node
// counts over chunks; stitch of every (chunk, node) run to its global place.

#include "pwm_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <queue>
#include <unordered_map>

namespace pwm {

namespace {

constexpr uint32_t kHChunk = 4096;       // symbols per chunk (one wave)
constexpr uint32_t kHMaxSym = 4096;      // supported alphabet (max symbol + 1)
constexpr uint32_t kHMaxLevels = 48;

template <typename T>
__global__ void k_hist(const T* text, uint64_t n, uint32_t nsym, unsigned long long* hist,
                       unsigned long long* first, unsigned long long* maxsym) {
  __shared__ unsigned int lh[256];
  const bool small = nsym <= 256;
  if (small)
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  unsigned long long mx = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t c = text[i];
    mx = c > mx ? c : mx;
    if (c < nsym) {
      if (small) atomicAdd(&lh[c], 1u);
      else atomicAdd(&hist[c], 1ull);
      if (first) atomicMin(&first[c], (unsigned long long)i);
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(mx, d, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(maxsym, mx);
  __syncthreads();
  if (small)
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
      if (lh[i]) atomicAdd(&hist[i], (unsigned long long)lh[i]);
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// Byte texts: counts of 8 candidate values (the most frequent in a sample)
// in registers, 4 bytes at a time: byte k of d = x ^ (v * 0x01010101) is 0
// exactly when bit 7 of byte k of ~(((d & 0x7f..) + 0x7f..) | d | 0x7f..)
// is set.  Bytes equal to no candidate go to LDS atomics.  16 bytes per
// lane per step.
struct HistCand {
  uint32_t rep[8];  // candidate value repeated in 4 bytes
  uint32_t val[8];
  uint32_t used;    // number of candidates
};

__device__ __forceinline__ uint32_t zero_bytes(uint32_t d) {
  return ~(((d & 0x7f7f7f7fu) + 0x7f7f7f7fu) | d | 0x7f7f7f7fu);
}

__global__ __launch_bounds__(256) void k_hist_bytes(const uint8_t* text, uint64_t n, HistCand hc,
                                                    unsigned long long* hist) {
  __shared__ unsigned int lh[256];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const bool aligned = (reinterpret_cast<uintptr_t>(text) & 15) == 0;
  const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i0 = tid * 16; i0 < n; i0 += nthr * 16) {
    uint32_t w[4] = {0, 0, 0, 0}, vmask[4] = {0, 0, 0, 0};
    if (i0 + 16 <= n && aligned) {
      const uint4 v = *reinterpret_cast<const uint4*>(text + i0);
      w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
      vmask[0] = vmask[1] = vmask[2] = vmask[3] = 0x80808080u;
    } else {
      const uint32_t nv = uint32_t(min<uint64_t>(16, n - i0));
      for (uint32_t k = 0; k < nv; ++k) {
        w[k >> 2] |= uint32_t(text[i0 + k]) << (8 * (k & 3));
        vmask[k >> 2] |= 0x80u << (8 * (k & 3));
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t hit = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (uint32_t(k) < hc.used) {
          const uint32_t z = zero_bytes(w[q] ^ hc.rep[k]) & vmask[q];
          cnt[k] += __builtin_popcount(z);
          hit |= z;
        }
      }
      uint32_t miss = vmask[q] & ~hit;
      while (miss) {
        const uint32_t bit = __builtin_ctz(miss);
        atomicAdd(&lh[(w[q] >> (bit - 7)) & 0xffu], 1u);
        miss &= miss - 1;
      }
    }
  }
  // candidate counts: wave sums, one atomic per wave and candidate
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t v = cnt[k];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0 && uint32_t(k) < hc.used && v) atomicAdd(&lh[hc.val[k]], v);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
    if (lh[i]) atomicAdd(&hist[i], (unsigned long long)lh[i]);
}

// Byte histogram for texts whose values are mostly below 8 (DNA-like).
// 16 bytes all below 4 pack into one word of 2-bit fields (below 8: two
// words of 3-bit fields, 6 bits a byte), and popcounts under masks give, per
// lane, the number of symbols with every bit of each set of value bits set;
// inclusion-exclusion turns these into exact counts at the end (about one
// instruction a byte for 2-bit data, against some 10 for the candidate
// tests).  16 bytes with a value of 8 or more go word by word into a 64-bit
// register of eight 8-bit counters (emptied into 32-bit totals every 240
// bytes), and bytes of 8 or more into LDS.  The bytes after the last whole
// grid step, and all of an unaligned text, go the 64-bit register way.
__global__ __launch_bounds__(256) void k_hist_small(const uint8_t* text, uint64_t n, unsigned long long* hist) {
  __shared__ unsigned int lh[256];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  uint32_t tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t acc = 0;
  uint32_t since = 0;  // bytes added to acc since it was emptied
  auto empty = [&]() {
#pragma unroll
    for (int v = 0; v < 8; ++v) tot[v] += uint32_t(acc >> (8 * v)) & 0xffu;
    acc = 0;
    since = 0;
  };
  const bool aligned = (reinterpret_cast<uintptr_t>(text) & 15) == 0;
  const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  auto add4 = [&](uint32_t x) {  // four bytes below 8
    acc += 1ull << (8 * (x & 7u));
    acc += 1ull << (8 * ((x >> 8) & 7u));
    acc += 1ull << (8 * ((x >> 16) & 7u));
    acc += 1ull << (8 * (x >> 24));
  };
  // The bulk: 64 bytes a lane per step, four 16-byte loads issued together.
  // 16 bytes whose values are all below 4 (or below 8) are packed into one
  // word (two words) of 2-bit (3-bit) fields, and the lane counts the
  // symbols with each set of value bits set: S0, S2, S01, S02, S12, S012
  // and the set bits overall; the exact counts follow at the end by
  // inclusion-exclusion.
  uint32_t nb_sym = 0, s_all = 0, s0 = 0, s2 = 0, s01 = 0, s02 = 0, s12 = 0, s012 = 0;
  auto count16 = [&](const uint4 v) {
    const uint32_t o = v.x | v.y | v.z | v.w;
    if ((o & 0xfcfcfcfcu) == 0) {
      const uint32_t x = v.x | (v.y << 2) | (v.z << 4) | (v.w << 6);
      s_all += __popc(x);
      s0 += __popc(x & 0x55555555u);
      s01 += __popc(x & (x >> 1) & 0x55555555u);
      nb_sym += 16;
    } else if ((o & 0xf8f8f8f8u) == 0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t x = h ? (v.z | (v.w << 3)) : (v.x | (v.y << 3));
        const uint32_t a = x & (x >> 1), b = x & (x >> 2);
        s_all += __popc(x);
        s0 += __popc(x & 0x09090909u);
        s2 += __popc(x & 0x24242424u);
        s01 += __popc(a & 0x09090909u);
        s12 += __popc(a & 0x12121212u);
        s02 += __popc(b & 0x09090909u);
        s012 += __popc(a & b & 0x09090909u);
      }
      nb_sym += 16;
    } else {
      const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if ((w4[q] & 0xf8f8f8f8u) == 0) {
          add4(w4[q]);
        } else {
          for (int e = 0; e < 4; ++e) {
            const uint32_t c = (w4[q] >> (8 * e)) & 0xffu;
            if (c < 8) acc += 1ull << (8 * c);
            else atomicAdd(&lh[c], 1u);
          }
        }
      }
      since += 16;
      if (since >= 240) empty();  // (no counter above 255)
    }
  };
  uint64_t done = 0;  // bytes [0, done) handled here
  if (aligned) {
    const uint64_t nb = n / (nthr * 64) * (nthr * 64);
    const uint4* tv = reinterpret_cast<const uint4*>(text);
    // (each wave reads 4 KiB in a row per step)
    const uint64_t wave = tid >> 6, nwaves = nthr >> 6, lane = tid & 63;
    for (uint64_t blk = wave; blk * 4096 < nb; blk += nwaves) {
      uint4 v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = tv[blk * 256 + k * 64 + lane];
#pragma unroll
      for (int k = 0; k < 4; ++k) count16(v[k]);
    }
    done = nb;
  }
  {
    const uint32_t s1 = s_all - s0 - s2;
    tot[7] += s012;
    tot[3] += s01 - s012;
    tot[5] += s02 - s012;
    tot[6] += s12 - s012;
    tot[1] += s0 - s01 - s02 + s012;
    tot[2] += s1 - s01 - s12 + s012;
    tot[4] += s2 - s02 - s12 + s012;
    tot[0] += nb_sym - s0 - s1 - s2 + s01 + s02 + s12 - s012;
  }
  for (uint64_t i0 = done + tid * 16; i0 < n; i0 += nthr * 16) {
    uint32_t w[4] = {0, 0, 0, 0}, vmask[4] = {0, 0, 0, 0};
    if (i0 + 16 <= n && aligned) {
      const uint4 v = *reinterpret_cast<const uint4*>(text + i0);
      w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
      vmask[0] = vmask[1] = vmask[2] = vmask[3] = 0x80808080u;
    } else {
      const uint32_t nv = uint32_t(min<uint64_t>(16, n - i0));
      for (uint32_t k = 0; k < nv; ++k) {
        w[k >> 2] |= uint32_t(text[i0 + k]) << (8 * (k & 3));
        vmask[k >> 2] |= 0x80u << (8 * (k & 3));
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (vmask[q] == 0x80808080u && (w[q] & 0xf8f8f8f8u) == 0) {
        add4(w[q]);
      } else {
        for (int k = 0; k < 4; ++k) {
          if (!((vmask[q] >> (8 * k + 7)) & 1)) continue;
          const uint32_t c = (w[q] >> (8 * k)) & 0xffu;
          if (c < 8) acc += 1ull << (8 * c);
          else atomicAdd(&lh[c], 1u);
        }
      }
    }
    since += 16;
    if (since >= 240) empty();  // (no counter above 255)
  }
  empty();
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    uint32_t t = tot[v];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(&lh[v], t);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
    if (lh[i]) atomicAdd(&hist[i], (unsigned long long)lh[i]);
}

struct HParams {
  const void* text;
  uint32_t width;
  uint64_t n;
  uint32_t levels;
  uint32_t nslot;
  const uint32_t* code_len;    // per symbol
  const uint64_t* code_word;   // per symbol
  const uint32_t* slot_of;     // [sym * levels + l] internal node slot (l >= 1)
  const uint32_t* slot_level;  // [slot] level
  const uint32_t* level_slot0; // [l] first slot of level l (slots grouped by level)
  uint64_t* level0;            // final level 0
  uint64_t* temp;              // chunk-local level l at temp + l*temp_stride
  uint64_t temp_stride;
  uint32_t* counts;            // [c * nslot + slot]
  uint64_t nchunks;
};

__device__ __forceinline__ uint64_t load_any(const void* t, uint32_t w, uint64_t i) {
  switch (w) {
    case 1: return static_cast<const uint8_t*>(t)[i];
    case 2: return static_cast<const uint16_t*>(t)[i];
    case 4: return static_cast<const uint32_t*>(t)[i];
    default: return static_cast<const uint64_t*>(t)[i];
  }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// One wave per chunk.  Pass A counts occurrences per internal node; pass B
// places every occurrence's level bit at chunk-local position
// local_start[node] + rank (text order inside the node).
__global__ __launch_bounds__(64) void k_huff_local(HParams p) {
  __shared__ uint32_t cnt[kHMaxSym];
  __shared__ uint32_t run[kHMaxSym];
  __shared__ unsigned long long bits[kHMaxLevels][kHChunk / 64];
  const uint64_t c = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const uint64_t base = c * kHChunk;
  const uint32_t Ec = uint32_t(min<uint64_t>(kHChunk, p.n - base));
  for (uint32_t s = lane; s < p.nslot; s += 64) cnt[s] = 0;
  for (uint32_t l = 0; l < p.levels; ++l)
    for (uint32_t w = lane; w < kHChunk / 64; w += 64) bits[l][w] = 0;
  __syncthreads();
  // pass A: counts
  for (uint32_t g = 0; g < Ec; g += 64) {
    const bool valid = g + lane < Ec;
    const uint64_t sym = valid ? load_any(p.text, p.width, base + g + lane) : 0;
    const uint32_t len = valid ? p.code_len[sym] : 0;
    const uint64_t word = valid ? p.code_word[sym] : 0;
    // level 0 bit, text order -> final level 0
    const uint64_t b0 = __ballot(valid && ((word >> (len - 1)) & 1));
    if (lane == 0) p.level0[(base + g) >> 6] = __builtin_bitreverse64(b0);
    for (uint32_t l = 1; l < p.levels; ++l) {
      const bool act = valid && len > l;
      uint64_t pending = __ballot(act);
      const uint32_t slot = act ? p.slot_of[sym * p.levels + l] : 0xffffffffu;
      while (pending) {
        const uint32_t leader = __ffsll((long long)pending) - 1;
        const uint32_t ls = __shfl(slot, leader, 64);
        const uint64_t m = __ballot(slot == ls);
        if (lane == leader) cnt[ls] += __popcll(m);
        pending &= ~m;
      }
    }
  }
  __syncthreads();
  // chunk-local starts per node (slots of one level are contiguous)
  if (lane == 0) {
    for (uint32_t l = 1; l < p.levels; ++l) {
      uint32_t acc = 0;
      for (uint32_t s = p.level_slot0[l]; s < p.level_slot0[l + 1]; ++s) {
        run[s] = acc;
        acc += cnt[s];
      }
    }
  }
  for (uint32_t s = lane; s < p.nslot; s += 64) p.counts[c * p.nslot + s] = cnt[s];
  __syncthreads();
  // pass B: place bits
  for (uint32_t g = 0; g < Ec; g += 64) {
    const bool valid = g + lane < Ec;
    const uint64_t sym = valid ? load_any(p.text, p.width, base + g + lane) : 0;
    const uint32_t len = valid ? p.code_len[sym] : 0;
    const uint64_t word = valid ? p.code_word[sym] : 0;
    for (uint32_t l = 1; l < p.levels; ++l) {
      const bool act = valid && len > l;
      uint64_t pending = __ballot(act);
      const uint32_t slot = act ? p.slot_of[sym * p.levels + l] : 0xffffffffu;
      uint32_t pos = 0;
      while (pending) {
        const uint32_t leader = __ffsll((long long)pending) - 1;
        const uint32_t ls = __shfl(slot, leader, 64);
        const uint64_t m = __ballot(slot == ls);
        const uint32_t r0 = run[ls];  // read by all lanes before the leader updates it
        if (slot == ls) pos = r0 + __popcll(m & lanemask_lt());
        __syncthreads();
        if (lane == leader) run[ls] = r0 + __popcll(m);
        __syncthreads();
        pending &= ~m;
      }
      if (act && ((word >> (len - 1 - l)) & 1))
        atomicOr(&bits[l][pos >> 6], 1ull << (pos & 63));
    }
  }
  __syncthreads();
  for (uint32_t l = 1; l < p.levels; ++l)
    for (uint32_t w = lane; w < kHChunk / 64; w += 64)
      p.temp[l * p.temp_stride + c * (kHChunk / 64) + w] = __builtin_bitreverse64(bits[l][w]);
}

__device__ __forceinline__ uint64_t window64(const uint64_t* a, int64_t s) {
  const int64_t w = s >> 6;
  const uint32_t sh = uint32_t(s & 63);
  const uint64_t w0 = w >= 0 ? a[w] : 0ull;
  if (sh == 0) return w0;
  return (w0 << sh) | (a[w + 1] >> (64 - sh));
}

__device__ __forceinline__ uint64_t range_mask(uint32_t a, uint32_t b) {
  const uint64_t hi = a == 0 ? ~0ull : (~0ull >> a);
  const uint64_t lo = b >= 64 ? ~0ull : ~(~0ull >> b);
  return hi & lo;
}

struct HStitch {
  const uint32_t* counts;
  const uint64_t* pref;
  const uint64_t* gstart;      // [slot]
  const uint32_t* slot_level;
  const uint32_t* level_slot0;
  const uint64_t* temp;
  uint64_t temp_stride;
  uint64_t* const* dst;        // level base pointers (device array)
  const uint64_t* level_bits;  // [l]
  uint32_t nslot, levels;
  int all_atomic;
  uint64_t unit_syms;          // symbols per run unit (chunk or super-chunk)
  uint32_t temp_l0;            // level l of temp at temp + (l - temp_l0) * temp_stride
};

// PWMHIP_HUFF_RUN_STITCH=1: the run-major stitch for huff_naive layouts too.
bool wide_gather_disabled_huff() {
  const char* e = std::getenv("PWMHIP_HUFF_RUN_STITCH");  // read at every call
  return e && e[0] == '1';
}

// Words of levels >= 1 that hold a run's first or last bit (shared with the
// neighbouring run, or a level's last word): cleared before the run stitch,
// which writes every other word whole and ORs these.  Enough when the nodes
// of every level tile it without gaps (huff_naive's layout).  One thread per
// (unit, slot).
__global__ __launch_bounds__(256) void k_huff_clear_edges(HStitch p, uint64_t units) {
  const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= units * p.nslot) return;
  const uint64_t c = i / p.nslot;
  const uint32_t s = uint32_t(i % p.nslot);
  const uint64_t len = p.counts[c * p.nslot + s];
  if (!len) return;
  const uint32_t l = p.slot_level[s];
  const uint64_t d0 = p.gstart[s] + p.pref[c * p.nslot + s], d1 = d0 + len;
  uint64_t* dst = p.dst[l];
  if (d0 & 63) dst[d0 >> 6] = 0;
  if (d1 & 63) dst[d1 >> 6] = 0;
}

// Run tables of a Huffman-shaped super-chunk build for the wide stitch
// (huff_naive's layout: a level's nodes, numbered in destination order, tile
// it).  Per-level bases: node s's first entry among its level's entries, and
// the entries of every level.
__global__ void k_huff_run_bases(const uint64_t* nztot, const uint32_t* slot_level, uint32_t nslot,
                                 uint64_t* nebase, uint64_t* level_ne) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint32_t j = 0; j < kMaxGroup; ++j) level_ne[j] = 0;
  for (uint32_t s = 0; s < nslot; ++s) {
    const uint32_t l = slot_level[s];
    nebase[s] = level_ne[l];
    level_ne[l] += nztot[s];
  }
}

// One thread per (unit, node) with symbols: its entry (destination bit,
// source bit in the unit-local level) at level l's base + node base + its
// rank among the node's non-empty units.
__global__ __launch_bounds__(256) void k_huff_runs(HStitch p, uint64_t units, const uint64_t* nzpref,
                                                   const uint64_t* nebase, uint64_t* tab_dst, uint64_t* tab_src) {
  const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= units * p.nslot) return;
  const uint64_t c = i / p.nslot;
  const uint32_t s = uint32_t(i % p.nslot);
  const uint32_t len = p.counts[c * p.nslot + s];
  if (!len) return;
  const uint32_t l = p.slot_level[s];
  uint64_t loc = 0;  // the node's start in the unit's level-l bits: earlier nodes of the level
  for (uint32_t t = p.level_slot0[l]; t < s; ++t) loc += p.counts[c * p.nslot + t];
  const uint64_t e = uint64_t((1u << l) - 2) * units + nebase[s] + nzpref[c * p.nslot + s];
  tab_dst[e] = p.gstart[s] + p.pref[c * p.nslot + s];
  tab_src[e] = c * p.unit_syms + loc;
}

// One block per unit, one thread per (run, destination word), runs looped.
__global__ __launch_bounds__(256) void k_huff_stitch(HStitch p) {
  const uint64_t c = blockIdx.x;
  for (uint32_t l = 1; l < p.levels; ++l) {
    uint64_t loc = 0;
    for (uint32_t s = p.level_slot0[l]; s < p.level_slot0[l + 1]; ++s) {
      const uint64_t len = p.counts[c * p.nslot + s];
      if (!len) continue;
      const uint64_t src = c * p.unit_syms + loc;
      loc += len;
      const uint64_t dbit = p.gstart[s] + p.pref[c * p.nslot + s];
      const uint64_t w0 = dbit >> 6, w1 = (dbit + len - 1) >> 6;
      const uint64_t nbits = p.level_bits[l];
      uint64_t* dst = p.dst[l];
      const uint64_t* sa = p.temp + (l - p.temp_l0) * p.temp_stride;
      for (uint64_t w = w0 + threadIdx.x; w <= w1; w += blockDim.x) {
        const uint64_t ws = w << 6;
        const uint64_t a = max(ws, dbit), b = min(ws + 64, dbit + len);
        const uint64_t v = window64(sa, int64_t(src) - int64_t(dbit - ws)) &
                           range_mask(uint32_t(a - ws), uint32_t(b - ws));
        const uint64_t wend = min(ws + 64, nbits);
        if (!p.all_atomic && a == ws && b == wend) dst[w] = v;
        else if (v) atomicOr(reinterpret_cast<unsigned long long*>(&dst[w]), (unsigned long long)v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host planning
// ---------------------------------------------------------------------------
struct code_table {
  std::vector<uint64_t> len, word, freq;
  uint64_t levels = 0;
  std::vector<uint64_t> level_sizes;
};

struct heap_item {
  uint64_t occ;
  std::vector<uint64_t> syms;
  bool operator>(const heap_item& o) const { return occ > o.occ; }
};

// canonical_huff_codes<A, is_tree>::compute_codes (huff_codes.hpp:141-285)
// for the histogram `entries` given in the reference's histogram order.
code_table make_codes(const std::vector<std::pair<uint64_t, uint64_t>>& entries, uint64_t max_sym, bool tree) {
  code_table ct;
  if (entries.empty()) return ct;
  ct.len.assign(max_sym + 1, 0);
  ct.word.assign(max_sym + 1, 0);
  ct.freq.assign(max_sym + 1, 0);
  for (auto& e : entries) ct.freq[e.first] = e.second;
  std::priority_queue<heap_item, std::vector<heap_item>, std::greater<heap_item>> heap;
  for (auto& e : entries) heap.emplace(heap_item{e.second, std::vector<uint64_t>{e.first}});
  if (heap.size() == 1) ++ct.len[heap.top().syms.front()];
  while (heap.size() > 1) {
    heap_item x = heap.top();
    heap.pop();
    heap_item y = heap.top();
    heap.pop();
    x.syms.insert(x.syms.end(), y.syms.begin(), y.syms.end());
    for (uint64_t s : x.syms) ++ct.len[s];
    heap.emplace(heap_item{x.occ + y.occ, x.syms});
  }
  std::vector<uint64_t> ord(ct.len.size());
  std::iota(ord.begin(), ord.end(), uint64_t(0));
  std::sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) { return ct.len[a] < ct.len[b]; });
  ct.levels = ct.len[ord.back()];
  ct.level_sizes.assign(ct.levels, 0);
  size_t k = 0;
  while (k < ord.size() && ct.len[ord[k]] == 0) ++k;
  if (!tree) {
    // code words handed out from the back of the bit-reversal-ordered list
    uint64_t cur_len = 1;
    std::vector<uint64_t> pool{0, 1};
    for (; k < ord.size(); ++k) {
      const uint64_t want = ct.len[ord[k]];
      while (cur_len < want) {
        std::vector<uint64_t> grown;
        grown.reserve(pool.size() * 2);
        for (uint64_t v : pool) grown.push_back(v << 1);
        for (uint64_t v : pool) grown.push_back((v << 1) | 1);
        pool.swap(grown);
        ++cur_len;
      }
      ct.word[ord[k]] = pool.back();
      pool.pop_back();
    }
  } else {
    // complemented canonical code words
    uint64_t cw = 0;
    for (size_t first = k; k < ord.size(); ++k) {
      if (k > first) cw = (cw + 1) << (ct.len[ord[k]] - ct.len[ord[k - 1]]);
      const uint64_t l = ct.len[ord[k]];
      ct.word[ord[k]] = (~cw) & (l >= 64 ? ~0ull : ((1ull << l) - 1));
    }
  }
  for (uint64_t s = 0; s < ct.len.size(); ++s)
    if (ct.len[s]) ct.level_sizes[ct.len[s] - 1] += ct.freq[s];
  for (uint64_t i = ct.levels; i-- > 1;) ct.level_sizes[i - 1] += ct.level_sizes[i];
  return ct;
}

template <typename A>
std::vector<std::pair<uint64_t, uint64_t>> map_order(const std::vector<uint64_t>& hist,
                                                     const std::vector<uint64_t>& first) {
  // Insert the distinct symbols in first-occurrence order into the same
  // unordered_map type the reference fills while scanning the text
  // (util/histogram.hpp:60-75); its iteration order is the entry order.
  std::vector<uint64_t> syms;
  for (uint64_t s = 0; s < hist.size(); ++s)
    if (hist[s]) syms.push_back(s);
  std::sort(syms.begin(), syms.end(), [&](uint64_t a, uint64_t b) { return first[a] < first[b]; });
  std::unordered_map<A, uint64_t> m;
  for (uint64_t s : syms) m.emplace(A(s), hist[s]);
  std::vector<std::pair<uint64_t, uint64_t>> out;
  for (const auto& e : m) out.emplace_back(uint64_t(e.first), e.second);
  return out;
}

inline uint64_t prefix_of(uint64_t len, uint64_t word, uint64_t l) { return word >> (len - l); }
inline uint32_t rev(uint64_t x, uint64_t bits) {
  uint64_t r = 0;
  for (uint64_t i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return uint32_t(r);
}

}  // namespace

// PWMHIP_HUFF_CHUNKS=1: the general per-chunk Huffman kernels for every input.
bool huffman_super_disabled() {
  const char* e = std::getenv("PWMHIP_HUFF_CHUNKS");  // read at every call
  return e && e[0] == '1';
}

// Byte texts, codes of 2..4 bits, node ranges that do not overlap: local
// build over super-chunks (level 0 in place), scan of the node counts over
// super-chunks, run stitch into the cleared levels 1..L-1.
void build_huffman_super(pwmhip_ctx* ctx, const uint8_t* d_text, uint64_t n, const code_table& ct, uint64_t L,
                         uint64_t maxsym, const std::vector<uint64_t>& slot_prefix,
                         const std::vector<uint32_t>& slot_level, const std::vector<uint32_t>& level_slot0,
                         const std::vector<uint64_t>& start, uint32_t nslot, const std::vector<uint64_t>& bits,
                         pwmhip_ws* ws) {
  hipStream_t st = ctx->stream;
  const uint64_t nchunks = (n + kChunkSyms - 1) / kChunkSyms;
  const uint64_t units = (nchunks + 15) / 16;
  const uint64_t tstride = units * (kSuperChunkSyms / 64) + 64;
  const uint32_t ns = std::max<uint32_t>(nslot, 1);
  size_t need = 0;
  auto add = [&](size_t b) { need += (b + 255) & ~size_t(255); };
  add(256); add((L - 1) * tstride * 8); add(units * ns * 4); add(units * ns * 8); add(ns * 8 + 8);
  add(ns * 8 + 8); add(L * 8); add(L * 8); add(scan_scratch_words(units, nslot) * 8); add((L + 1) * 4);
  add(ns * 4 + 4);
  // the wide stitch (huff_naive's layout): non-empty ranks and counts, node
  // bases, entries per level, tables, every wave's first entry
  const bool wide = ws->is_huffman == 2 && !wide_gather_disabled_huff();
  uint64_t maxbits = 0;
  for (uint64_t l = 1; l < L; ++l) maxbits = std::max(maxbits, bits[l]);
  const uint64_t ntab = ((uint64_t(1) << L) - 2) * units;
  if (wide) {
    add(units * ns * 8); add(ns * 8 + 8); add(ns * 8 + 8); add(kMaxGroup * 8); add(ntab * 8); add(ntab * 8);
    add(wide_stitch_first_words(maxbits, uint32_t(L - 1)) * 4); add(scan_scratch_words(units, nslot) * 8);
  }
  ctx->scratch.reset();
  ctx->scratch.ensure(need);
  arena& A = ctx->scratch;
  uint8_t* d_code = A.take<uint8_t>(256);
  uint64_t* d_temp = A.take<uint64_t>((L - 1) * tstride);
  uint32_t* d_counts = A.take<uint32_t>(units * ns);
  uint64_t* d_pref = A.take<uint64_t>(units * ns);
  uint64_t* d_tot = A.take<uint64_t>(ns + 1);
  uint64_t* d_start = A.take<uint64_t>(ns + 1);
  uint64_t** d_dst = A.take<uint64_t*>(L);
  uint64_t* d_lbits = A.take<uint64_t>(L);
  uint64_t* d_tiles = A.take<uint64_t>(scan_scratch_words(units, nslot));
  uint32_t* d_level_slot0 = A.take<uint32_t>(L + 1);
  uint32_t* d_slot_level = A.take<uint32_t>(ns + 1);
  uint64_t *d_nzpref = nullptr, *d_nztot = nullptr, *d_nebase = nullptr, *d_lne = nullptr;
  uint64_t *d_tdst = nullptr, *d_tsrc = nullptr, *d_tiles2 = nullptr;
  uint32_t* d_first = nullptr;
  if (wide) {
    d_nzpref = A.take<uint64_t>(units * ns);
    d_nztot = A.take<uint64_t>(ns + 1);
    d_nebase = A.take<uint64_t>(ns + 1);
    d_lne = A.take<uint64_t>(kMaxGroup);
    d_tdst = A.take<uint64_t>(ntab);
    d_tsrc = A.take<uint64_t>(ntab);
    d_first = A.take<uint32_t>(wide_stitch_first_words(maxbits, uint32_t(L - 1)));
    d_tiles2 = A.take<uint64_t>(scan_scratch_words(units, nslot));
  }

  std::vector<uint8_t> code(256, 0);
  for (uint64_t s = 0; s < ct.len.size() && s < 256; ++s)
    if (ct.len[s]) code[s] = uint8_t(ct.word[s] << (L - ct.len[s]));
  HuffSuperParams hp{};
  hp.text = d_text;
  hp.n = n;
  hp.nchunks = nchunks;
  hp.levels = uint32_t(L);
  hp.code = d_code;
  hp.small_alpha = maxsym < 8 ? 1u : 0u;
  hp.level0 = ws->d_words + ws->level_offset[0];
  hp.temp = d_temp;
  hp.temp_stride = tstride;
  hp.counts = d_counts;
  hp.nslot = nslot;
  std::memset(hp.rank, 0xff, sizeof(hp.rank));
  for (uint64_t j = 1; j < L; ++j) {
    hp.slot0[j] = level_slot0[j];
    for (uint32_t sl = level_slot0[j]; sl < level_slot0[j + 1]; ++sl)
      hp.rank[j][rev(slot_prefix[sl], j)] = uint8_t(sl - level_slot0[j]);
  }
  std::vector<uint64_t*> dst(L);
  for (uint64_t l = 0; l < L; ++l) dst[l] = ws->d_words + ws->level_offset[l];
  PWM_HIP_TRY(hipMemcpyAsync(d_code, code.data(), 256, hipMemcpyHostToDevice, st));
  if (nslot) {
    PWM_HIP_TRY(hipMemcpyAsync(d_start, start.data(), nslot * 8, hipMemcpyHostToDevice, st));
    PWM_HIP_TRY(hipMemcpyAsync(d_slot_level, slot_level.data(), nslot * 4, hipMemcpyHostToDevice, st));
  }
  PWM_HIP_TRY(hipMemcpyAsync(d_level_slot0, level_slot0.data(), (L + 1) * 4, hipMemcpyHostToDevice, st));
  PWM_HIP_TRY(hipMemcpyAsync(d_dst, dst.data(), L * sizeof(uint64_t*), hipMemcpyHostToDevice, st));
  PWM_HIP_TRY(hipMemcpyAsync(d_lbits, bits.data(), L * 8, hipMemcpyHostToDevice, st));
  // runs of levels >= 1 meet inside words: with huff_naive's layout the
  // words holding run ends are cleared after the scan (k_huff_clear_edges);
  // huff_pc's node borders can leave gaps, so there the levels start cleared
  const bool edges_only = ws->is_huffman == 2;  // (with `wide`: nothing to clear)
  if (!edges_only)
    for (uint64_t l = 1; l < L; ++l)
      PWM_HIP_TRY(hipMemsetAsync(dst[l], 0, words_for(bits[l]) * 8, st));

  PWM_HIP_TRY(hipEventRecord(ctx->ev[0], st));
  launch_huff_super(ctx, hp);
  PWM_HIP_TRY(hipEventRecord(ctx->ev[1], st));
  if (nslot) {
    launch_scan_slots(st, d_counts, d_pref, d_tot, units, nslot, d_tiles);
    PWM_HIP_TRY(hipEventRecord(ctx->ev[2], st));
    HStitch hs{};
    hs.counts = d_counts;
    hs.pref = d_pref;
    hs.gstart = d_start;
    hs.slot_level = d_slot_level;
    hs.level_slot0 = d_level_slot0;
    hs.temp = d_temp;
    hs.temp_stride = tstride;
    hs.dst = d_dst;
    hs.level_bits = d_lbits;
    hs.nslot = nslot;
    hs.levels = uint32_t(L);
    hs.all_atomic = 0;
    hs.unit_syms = kSuperChunkSyms;
    hs.temp_l0 = 1;
    if (wide) {
      // destination-major: every word of every level written once
      launch_scan_slots(st, d_counts, d_nzpref, d_nztot, units, nslot, d_tiles2, true);
      hipLaunchKernelGGL(k_huff_run_bases, dim3(1), dim3(64), 0, st, d_nztot, d_slot_level, nslot, d_nebase, d_lne);
      hipLaunchKernelGGL(k_huff_runs, dim3(uint32_t((units * nslot + 255) / 256)), dim3(256), 0, st, hs, units,
                         d_nzpref, d_nebase, d_tdst, d_tsrc);
      WideStitch w{};
      w.tab_dst = d_tdst;
      w.tab_src = d_tsrc;
      w.level_ne = d_lne;
      w.first = d_first;
      w.src = d_temp;
      w.src_stride = tstride;
      for (uint64_t l = 1; l < L; ++l) {
        w.dst[l] = dst[l];
        w.nbits[l] = bits[l];
      }
      w.units = units;
      w.nsub = uint32_t(L - 1);
      launch_wide_stitch(ctx, w);
    } else {
      if (edges_only)
        hipLaunchKernelGGL(k_huff_clear_edges, dim3(uint32_t((units * nslot + 255) / 256)), dim3(256), 0, st, hs,
                           units);
      hipLaunchKernelGGL(k_huff_stitch, dim3(uint32_t(units)), dim3(256), 0, st, hs);
    }
    PWM_HIP_TRY(hipGetLastError());
  } else {
    PWM_HIP_TRY(hipEventRecord(ctx->ev[2], st));
  }
  PWM_HIP_TRY(hipEventRecord(ctx->ev[3], st));
  PWM_HIP_TRY(hipStreamSynchronize(st));
  float ms = 0;
  for (int i = 0; i < 3; ++i) {
    PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
    ctx->last_ms[i] = ms;
  }
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
  ctx->last_ms[3] = ms;
}

void build_huffman_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width, bool tree,
                          bool pc_quirks, pwmhip_ws* ws, uint64_t* reuse, uint64_t reuse_bytes) {
  struct ReuseGuard {  // an earlier output not taken over is freed on every exit
    uint64_t*& p;
    ~ReuseGuard() {
      if (p) (void)hipFree(p);
    }
  } reuse_guard{reuse};
  hipStream_t st = ctx->stream;
  ws->is_tree = tree;
  ws->is_huffman = 2;  // (an empty structure has the naive layout)
  // 1. histogram, max symbol (and first occurrences for map ordering)
  const uint32_t nsym_cap = kHMaxSym;
  ctx->scratch.reset();
  ctx->scratch.ensure(size_t(nsym_cap) * 16 + 4096);
  unsigned long long* d_hist = ctx->scratch.take<unsigned long long>(nsym_cap);
  unsigned long long* d_first = ctx->scratch.take<unsigned long long>(nsym_cap);
  unsigned long long* d_max = ctx->scratch.take<unsigned long long>(1);
  PWM_HIP_TRY(hipMemsetAsync(d_hist, 0, nsym_cap * 8, st));
  PWM_HIP_TRY(hipMemsetAsync(d_first, 0xff, nsym_cap * 8, st));
  PWM_HIP_TRY(hipMemsetAsync(d_max, 0, 8, st));
  const uint32_t nsym = width == 1 ? 256 : nsym_cap;
  const uint32_t grid = uint32_t(std::min<uint64_t>(std::max<uint64_t>((n + 1023) / 1024, 1), uint64_t(ctx->num_cus) * 8));
  unsigned long long* first_arg = width == 1 ? nullptr : d_first;
  if (n) {
    switch (width) {
      case 1: {
        // candidates: the 8 most frequent values of a 1 MiB sample
        const uint64_t ns = std::min<uint64_t>(n, 1 << 20);
        const uint32_t gs = uint32_t(std::min<uint64_t>((ns + 1023) / 1024, 256));
        hipLaunchKernelGGL(k_hist<uint8_t>, dim3(gs), dim3(256), 0, st, (const uint8_t*)d_text, ns, 256u, d_hist,
                           (unsigned long long*)nullptr, d_max);
        std::vector<uint64_t> sh(256);
        PWM_HIP_TRY(hipMemcpyAsync(sh.data(), d_hist, 256 * 8, hipMemcpyDeviceToHost, st));
        PWM_HIP_TRY(hipMemsetAsync(d_hist, 0, 256 * 8, st));
        PWM_HIP_TRY(hipStreamSynchronize(st));
        std::vector<uint32_t> ord(256);
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return sh[x] > sh[y]; });
        HistCand hc{};
        for (uint32_t k = 0; k < 8; ++k) {
          if (!sh[ord[k]]) break;
          hc.val[k] = ord[k];
          hc.rep[k] = ord[k] * 0x01010101u;
          hc.used = k + 1;
        }
        uint64_t below8 = 0, all = 0;  // (the sample decides the kernel; both count exactly)
        for (uint32_t c = 0; c < 256; ++c) {
          all += sh[c];
          if (c < 8) below8 += sh[c];
        }
        if (below8 * 1000 >= all * 999)
          // (four workgroups a CU: the kernel's registers leave room for
          // seven waves a SIMD, and a grid-stride grid must be resident at once)
          hipLaunchKernelGGL(k_hist_small, dim3(std::min<uint32_t>(grid, uint32_t(ctx->num_cus) * 4)), dim3(256), 0, st,
                             (const uint8_t*)d_text, n, d_hist);
        else
          hipLaunchKernelGGL(k_hist_bytes, dim3(grid), dim3(256), 0, st, (const uint8_t*)d_text, n, hc, d_hist);
        break;
      }
      case 2: hipLaunchKernelGGL(k_hist<uint16_t>, dim3(grid), dim3(256), 0, st, (const uint16_t*)d_text, n, nsym, d_hist, first_arg, d_max); break;
      case 4: hipLaunchKernelGGL(k_hist<uint32_t>, dim3(grid), dim3(256), 0, st, (const uint32_t*)d_text, n, nsym, d_hist, first_arg, d_max); break;
      default: hipLaunchKernelGGL(k_hist<uint64_t>, dim3(grid), dim3(256), 0, st, (const uint64_t*)d_text, n, nsym, d_hist, first_arg, d_max); break;
    }
    PWM_HIP_TRY(hipGetLastError());
  }
  std::vector<uint64_t> hist(nsym), first(nsym, ~0ull);
  uint64_t maxsym = 0;
  PWM_HIP_TRY(hipMemcpyAsync(hist.data(), d_hist, nsym * 8, hipMemcpyDeviceToHost, st));
  if (first_arg) PWM_HIP_TRY(hipMemcpyAsync(first.data(), d_first, nsym * 8, hipMemcpyDeviceToHost, st));
  PWM_HIP_TRY(hipMemcpyAsync(&maxsym, d_max, 8, hipMemcpyDeviceToHost, st));
  PWM_HIP_TRY(hipStreamSynchronize(st));
  if (width == 1) {  // the byte histogram does not track the maximum
    maxsym = 0;
    for (uint64_t c = 0; c < 256; ++c)
      if (hist[c]) maxsym = c;
  }
  if (n && maxsym >= nsym) throw status_error{PWMHIP_ERR_UNSUPPORTED, "Huffman alphabet above 4096 symbols"};

  // 2. codes (histogram entry order as in util/histogram.hpp:39-76)
  std::vector<std::pair<uint64_t, uint64_t>> entries;
  if (width == 1) {
    for (uint64_t s = 0; s < 256; ++s)
      if (hist[s]) entries.emplace_back(s, hist[s]);
  } else if (width == 2) {
    entries = map_order<uint16_t>(hist, first);
  } else if (width == 4) {
    entries = map_order<uint32_t>(hist, first);
  } else {
    entries = map_order<uint64_t>(hist, first);
  }
  code_table ct = make_codes(entries, maxsym, tree);
  ws->n_codes = ct.len.size();
  if (ws->n_codes) {
    ws->code_len = static_cast<uint64_t*>(std::malloc(ws->n_codes * 8));
    ws->code_word = static_cast<uint64_t*>(std::malloc(ws->n_codes * 8));
    if (!ws->code_len || !ws->code_word) throw status_error{PWMHIP_ERR_OOM, "codes"};
    std::memcpy(ws->code_len, ct.len.data(), ws->n_codes * 8);
    std::memcpy(ws->code_word, ct.word.data(), ws->n_codes * 8);
  }
  const uint64_t L = ct.levels;
  if (n == 0 || L == 0) {
    ws->levels = 0;
    return;
  }
  if (L > kHMaxLevels) throw status_error{PWMHIP_ERR_UNSUPPORTED, "Huffman code longer than 48 bits"};

  // 3. internal nodes per level and their global starts
  const uint64_t S = ct.len.size();
  std::vector<uint32_t> slot_of(S * L, 0xffffffffu);
  std::vector<uint32_t> slot_level, level_slot0(L + 1, 0);
  std::vector<uint64_t> slot_prefix, slot_count;
  std::vector<uint64_t> start;
  uint64_t zeros[PWMHIP_MAX_LEVELS] = {};
  // B state of the reference's border arrays (levels < 8 / >= 8)
  std::unordered_map<uint64_t, uint64_t> bstate[2];
  bool overlap = false;
  bool naive_layout = true;  // every node starts where huff_naive puts it
  for (uint64_t l = 1; l < L; ++l) {
    level_slot0[l] = uint32_t(slot_level.size());
    // counts per l-bit prefix of occurrences with code length > l (continuing)
    // and with code length >= l (the reference's level histogram)
    std::unordered_map<uint64_t, uint64_t> cont, ge, ends1;
    std::vector<uint64_t> seeded;  // pre-seeded keys (ctx_huffman.hpp:95-101)
    std::unordered_map<uint64_t, bool> seen;
    for (uint64_t s = 0; s < S; ++s) {
      if (ct.len[s] < l) continue;
      const uint64_t p = prefix_of(ct.len[s], ct.word[s], l);
      if (!seen.count(p)) {
        seen[p] = true;
        seeded.push_back(p);
      }
      ge[p] += ct.freq[s];
      if (ct.len[s] > l) cont[p] += ct.freq[s];
    }
    std::vector<uint64_t> keys;
    for (auto& e : cont)
      if (e.second) keys.push_back(e.first);
    // node order: numeric (tree) or bit-reversed (matrix) -- naive semantics
    if (tree) std::sort(keys.begin(), keys.end());
    else std::sort(keys.begin(), keys.end(), [&](uint64_t a, uint64_t b) { return rev(a, l) < rev(b, l); });
    std::unordered_map<uint64_t, uint64_t> naive_start;
    uint64_t acc = 0;
    for (uint64_t p : keys) {
      naive_start[p] = acc;
      acc += cont[p];
    }
    if (pc_quirks) {
      auto& B = bstate[l < 8 ? 0 : 1];
      const uint64_t nk = seeded.size();
      B[0] = 0;
      for (uint64_t i = 1; i < nk; ++i) {
        auto it = ge.find(i - 1);
        B[i] = B[i - 1] + (it == ge.end() ? 0 : it->second);
      }
      if (!tree) zeros[l - 1] = B[1];
    }
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    for (uint64_t p : keys) {
      const uint32_t slot = uint32_t(slot_level.size());
      slot_level.push_back(uint32_t(l));
      slot_prefix.push_back(p);
      slot_count.push_back(cont[p]);
      const uint64_t s0 = pc_quirks ? bstate[l < 8 ? 0 : 1][p] : naive_start[p];
      if (s0 != naive_start[p]) naive_layout = false;
      if (s0 + cont[p] > ct.level_sizes[l])
        throw status_error{PWMHIP_ERR_UNSUPPORTED, "reference huff_pc writes outside the level here"};
      start.push_back(s0);
      ranges.emplace_back(s0, s0 + cont[p]);
      for (uint64_t s = 0; s < S; ++s)
        if (ct.len[s] > l && prefix_of(ct.len[s], ct.word[s], l) == p) slot_of[s * L + l] = slot;
    }
    std::sort(ranges.begin(), ranges.end());
    for (size_t i = 1; i < ranges.size(); ++i)
      if (ranges[i].first < ranges[i - 1].second) overlap = true;
    if (pc_quirks) {
      // borders advance for every processed occurrence (huff_pc.hpp:46-58),
      // at level 1 including those of code length 1
      auto& B = bstate[l < 8 ? 0 : 1];
      for (auto& e : cont) B[e.first] += e.second;
      if (l == 1)
        for (uint64_t s = 0; s < S; ++s)
          if (ct.len[s] == 1) B[ct.word[s]] += ct.freq[s];
    }
  }
  level_slot0[L] = uint32_t(slot_level.size());
  if (L >= 1) level_slot0[0] = 0;
  const uint32_t nslot = uint32_t(slot_level.size());
  if (nslot > kHMaxSym) throw status_error{PWMHIP_ERR_UNSUPPORTED, "too many Huffman nodes"};

  // true zeros (naive): zeros of every level from the codes
  if (!pc_quirks && !tree) {
    for (uint64_t s = 0; s < S; ++s)
      for (uint64_t l = 0; l < ct.len[s]; ++l)
        if (!((ct.word[s] >> (ct.len[s] - 1 - l)) & 1)) zeros[l] += ct.freq[s];
  }

  // 1 = Huffman-shaped, the reference's huff_pc node borders; 2 = the
  // huff_naive layout (symbol queries can walk it)
  ws->is_huffman = naive_layout ? 2 : 1;

  // 4. output allocation
  std::vector<uint64_t> bits(ct.level_sizes.begin(), ct.level_sizes.end());
  uint64_t* earlier = reuse;
  reuse = nullptr;  // alloc_structure takes the earlier output over or frees it
  alloc_structure(ws, L, bits.data(), earlier, reuse_bytes);
  ws->n_zeros = tree ? 0 : L;
  for (uint64_t l = 0; l < L; ++l) ws->zeros[l] = tree ? 0 : zeros[l];

  // 5a. byte texts with codes of 2..4 bits: the plain construction's
  // super-chunk build with code-mapped symbols, then the run stitch
  if (width == 1 && L >= 2 && L <= 4 && !overlap && !huffman_super_disabled()) {
    build_huffman_super(ctx, static_cast<const uint8_t*>(d_text), n, ct, L, maxsym, slot_prefix, slot_level,
                        level_slot0, start, nslot, bits, ws);
    return;
  }

  // 5. device tables and scratch
  const uint64_t nchunks = (n + kHChunk - 1) / kHChunk;
  const uint64_t tstride = level_stride(nchunks * kHChunk);
  size_t need = 0;
  auto add = [&](size_t b) { need += (b + 255) & ~size_t(255); };
  add(S * 4); add(S * 8); add(S * L * 4); add(nslot * 4 + 4); add((L + 1) * 4); add(nslot * 8 + 8);
  add(L * tstride * 8); add(nchunks * std::max<uint32_t>(nslot, 1) * 4); add(nchunks * std::max<uint32_t>(nslot, 1) * 8);
  add(nslot * 8 + 8); add(L * 8); add(L * 8); add(scan_scratch_words(nchunks, nslot) * 8);
  ctx->scratch.reset();
  ctx->scratch.ensure(need);
  arena& A = ctx->scratch;
  uint32_t* d_len = A.take<uint32_t>(S);
  uint64_t* d_word = A.take<uint64_t>(S);
  uint32_t* d_slot_of = A.take<uint32_t>(S * L);
  uint32_t* d_slot_level = A.take<uint32_t>(nslot + 1);
  uint32_t* d_level_slot0 = A.take<uint32_t>(L + 1);
  uint64_t* d_start = A.take<uint64_t>(nslot + 1);
  uint64_t* d_temp = A.take<uint64_t>(L * tstride);
  uint32_t* d_counts = A.take<uint32_t>(nchunks * std::max<uint32_t>(nslot, 1));
  uint64_t* d_pref = A.take<uint64_t>(nchunks * std::max<uint32_t>(nslot, 1));
  uint64_t* d_tot = A.take<uint64_t>(nslot + 1);
  uint64_t** d_dst = A.take<uint64_t*>(L);
  uint64_t* d_lbits = A.take<uint64_t>(L);
  uint64_t* d_tiles = A.take<uint64_t>(scan_scratch_words(nchunks, nslot));
  std::vector<uint32_t> len32(S);
  for (uint64_t s = 0; s < S; ++s) len32[s] = uint32_t(ct.len[s]);
  std::vector<uint64_t*> dst(L);
  for (uint64_t l = 0; l < L; ++l) dst[l] = ws->d_words + ws->level_offset[l];
  PWM_HIP_TRY(hipMemcpyAsync(d_len, len32.data(), S * 4, hipMemcpyHostToDevice, st));
  PWM_HIP_TRY(hipMemcpyAsync(d_word, ct.word.data(), S * 8, hipMemcpyHostToDevice, st));
  PWM_HIP_TRY(hipMemcpyAsync(d_slot_of, slot_of.data(), S * L * 4, hipMemcpyHostToDevice, st));
  if (nslot) {
    PWM_HIP_TRY(hipMemcpyAsync(d_slot_level, slot_level.data(), nslot * 4, hipMemcpyHostToDevice, st));
    PWM_HIP_TRY(hipMemcpyAsync(d_start, start.data(), nslot * 8, hipMemcpyHostToDevice, st));
  }
  PWM_HIP_TRY(hipMemcpyAsync(d_level_slot0, level_slot0.data(), (L + 1) * 4, hipMemcpyHostToDevice, st));
  PWM_HIP_TRY(hipMemcpyAsync(d_dst, dst.data(), L * sizeof(uint64_t*), hipMemcpyHostToDevice, st));
  PWM_HIP_TRY(hipMemcpyAsync(d_lbits, bits.data(), L * 8, hipMemcpyHostToDevice, st));
  // levels >= 1 are assembled from runs; clear them first
  for (uint64_t l = 1; l < L; ++l)
    PWM_HIP_TRY(hipMemsetAsync(dst[l], 0, words_for(bits[l]) * 8, st));

  PWM_HIP_TRY(hipEventRecord(ctx->ev[0], st));
  HParams hp{};
  hp.text = d_text;
  hp.width = width;
  hp.n = n;
  hp.levels = uint32_t(L);
  hp.nslot = nslot;
  hp.code_len = d_len;
  hp.code_word = d_word;
  hp.slot_of = d_slot_of;
  hp.slot_level = d_slot_level;
  hp.level_slot0 = d_level_slot0;
  hp.level0 = dst[0];
  hp.temp = d_temp;
  hp.temp_stride = tstride;
  hp.counts = d_counts;
  hp.nchunks = nchunks;
  hipLaunchKernelGGL(k_huff_local, dim3(uint32_t(nchunks)), dim3(64), 0, st, hp);
  PWM_HIP_TRY(hipGetLastError());
  PWM_HIP_TRY(hipEventRecord(ctx->ev[1], st));
  if (nslot) {
    launch_scan_slots(st, d_counts, d_pref, d_tot, nchunks, nslot, d_tiles);
    PWM_HIP_TRY(hipEventRecord(ctx->ev[2], st));
    HStitch hs{};
    hs.counts = d_counts;
    hs.pref = d_pref;
    hs.gstart = d_start;
    hs.slot_level = d_slot_level;
    hs.level_slot0 = d_level_slot0;
    hs.temp = d_temp;
    hs.temp_stride = tstride;
    hs.dst = d_dst;
    hs.level_bits = d_lbits;
    hs.nslot = nslot;
    hs.levels = uint32_t(L);
    hs.all_atomic = overlap ? 1 : 0;
    hs.unit_syms = kHChunk;
    hs.temp_l0 = 0;
    hipLaunchKernelGGL(k_huff_stitch, dim3(uint32_t(nchunks)), dim3(256), 0, st, hs);
    PWM_HIP_TRY(hipGetLastError());
  } else {
    PWM_HIP_TRY(hipEventRecord(ctx->ev[2], st));
  }
  PWM_HIP_TRY(hipEventRecord(ctx->ev[3], st));
  PWM_HIP_TRY(hipStreamSynchronize(st));
  float ms = 0;
  for (int i = 0; i < 3; ++i) {
    PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
    ctx->last_ms[i] = ms;
  }
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
  ctx->last_ms[3] = ms;
}

}  // namespace pwm
