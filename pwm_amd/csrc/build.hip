// pwm_amd/csrc/build.hip -- plain (non-Huffman) wavelet matrix / tree
// construction for gfx950.
//
// The reference builds each level with a scan of the text that scatters one
// bit per symbol into per-prefix cursors (construction/ppc.hpp:74-86,
// building_blocks.hpp:139-153), or with a per-level counting sort
// (construction/pps.hpp:60-151).  Its domain-decomposition variant
// (wx_dd_pc.hpp:63-118) builds one structure per thread slice and then
// stitches the per-slice level blocks together in bucket order
// (construction/merge.hpp:192-465).  This file is that last schedule,
// re-designed for the GPU:
//
//  K1  k_local_build   one workgroup per 16 Ki-symbol chunk (persistent).
//      The chunk's symbols are turned into bit planes in LDS (plane k =
//      level l0+k bit of every symbol, 64 symbols per word).  Level l0 is the
//      input order, so it is written straight to its final place.  Levels
//      l0+1.. are produced by the matrix recursion "next sequence = stable
//      partition of the current one by the current bit" (wx_naive.hpp:108-151)
//      applied to whole 64-bit words: every remaining plane is split by the
//      mask plane with a byte-wise compress table (zeros-part | ones-part),
//      and the parts are concatenated in LDS at offsets from a block scan.
//      The chunk-local level bits, the chunk's bucket counts and, for the
//      next pass, the partitioned remaining planes go to scratch.
//  K2  k_tile_* / k_bucket_starts / k_make_runs: per (level, bucket)
//      exclusive scan over the chunks, global bucket starts (matrix: bucket
//      index order = bit-reversed prefix, permutation.hpp:36-51; tree:
//      numeric prefix order), and a compact table of the non-empty
//      (chunk, bucket) runs of every level in destination order.
//  K3  k_gather: every destination word is assembled once from the runs that
//      overlap it (the copy_bits analogue, merge.hpp:19-190), with no atomics
//      and no pre-clearing.
//
// More than kMaxGroup levels are built in passes of up to kMaxGroup levels;
// the partitioned remaining planes of one pass are the next pass's input.
// Trees with more than kMaxGroup levels are built as matrices and then have
// each level's buckets moved into numeric order (bitops.hip).

#include "pwm_internal.hpp"

#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <mutex>

namespace pwm {

namespace {

constexpr int P = kChunkWords;  // words per plane per chunk == threads per block

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t brev64(uint64_t x) { return __builtin_bitreverse64(x); }

__host__ __device__ inline uint32_t compress8(uint32_t x, uint32_t m) {
  uint32_t r = 0, s = 0;
  for (int i = 0; i < 8; ++i) {
    if ((m >> i) & 1) {
      r |= ((x >> i) & 1) << s;
      ++s;
    }
  }
  return r;
}

// Bits of x at the set positions of m, packed towards bit 0 (LSB-first),
// through a 64 Ki-entry byte table in LDS: lut[(m_byte << 8) | x_byte].
__device__ __forceinline__ uint64_t compress64(uint64_t x, uint64_t m, const uint8_t* lut) {
  if (m == ~0ull) return x;
  if (m == 0) return 0;
  uint64_t r = 0;
  uint32_t s = 0;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint32_t mb = uint32_t(m >> (8 * b)) & 0xffu;
    const uint32_t xb = uint32_t(x >> (8 * b)) & 0xffu;
    r |= uint64_t(lut[(mb << 8) | xb]) << s;
    s += __popc(mb);
  }
  return r;
}

// OR `len` low bits of `bits` into an LDS bit array at bit position pos.
__device__ __forceinline__ void or_bits(uint64_t* arr, uint32_t pos, uint64_t bits, uint32_t len) {
  if (len == 0) return;
  const uint32_t w = pos >> 6, sh = pos & 63;
  atomicOr(reinterpret_cast<unsigned long long*>(&arr[w]), (unsigned long long)(bits << sh));
  if (sh && sh + len > 64)
    atomicOr(reinterpret_cast<unsigned long long*>(&arr[w + 1]), (unsigned long long)(bits >> (64 - sh)));
}

// Workgroup barrier that orders LDS only: outstanding global loads (the next
// chunk's prefetch) stay in flight across it.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive block scan of two counters; totals returned to every thread.
// sm needs 2 * NT/64 words.  Contains two barriers.
template <int NT>
__device__ __forceinline__ void block_exscan2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb,
                                              uint32_t& ta, uint32_t& tb, uint32_t* sm) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t ia = a, ib = b;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t xa = __shfl_up(ia, d, 64);
    const uint32_t xb = __shfl_up(ib, d, 64);
    if (lane >= d) {
      ia += xa;
      ib += xb;
    }
  }
  if (lane == 63) {
    sm[wid] = ia;
    sm[NT / 64 + wid] = ib;
  }
  lds_sync();
  uint32_t oa = 0, ob = 0, sa = 0, sb = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const uint32_t va = sm[w], vb = sm[NT / 64 + w];
    if (w < wid) {
      oa += va;
      ob += vb;
    }
    sa += va;
    sb += vb;
  }
  ea = oa + ia - a;
  eb = ob + ib - b;
  ta = sa;
  tb = sb;
  lds_sync();
}

// rank1 in an LDS plane with per-word exclusive one-counts `ooff`.
__device__ __forceinline__ uint32_t rank1_lds(const uint64_t* plane, const uint32_t* ooff, uint32_t pos) {
  const uint32_t w = pos >> 6, r = pos & 63;
  if (r == 0) return ooff[w];
  return ooff[w] + __popcll(plane[w] & ((1ull << r) - 1));
}

// Device-memory view of a pointer: global (not flat) loads and stores, so
// waits for LDS traffic never wait on them.
template <typename T>
__device__ __forceinline__ auto gmem(T* p) {
#if __HIP_DEVICE_COMPILE__
  return (__attribute__((address_space(1))) T*)p;
#else
  return p;  // host pass: never executed
#endif
}

// Streaming accesses.  Stores of data written once (levels, chunk-local
// rows) carry the non-temporal hint; measured on the 4 GiB DNA build, the
// hint helps stores (-1.5%) and hurts the text loads (+40%), so loads stay
// ordinary.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return *gmem(p);
}

template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
#if __HIP_DEVICE_COMPILE__
  __builtin_nontemporal_store(v, gmem(p));
#else
  *p = v;
#endif
}

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2_u __attribute__((ext_vector_type(2), aligned(8)));

// Two consecutive words at an 8-byte aligned address.
#if __HIP_DEVICE_COMPILE__
typedef __attribute__((address_space(1))) const u64x2_u gc_u64x2;
typedef __attribute__((address_space(1))) u64x2_u g_u64x2;
#else
typedef const u64x2_u gc_u64x2;
typedef u64x2_u g_u64x2;
#endif

__device__ __forceinline__ u64x2_u ld_stream2(const uint64_t* p) {
  return *(gc_u64x2*)p;
}

__device__ __forceinline__ void st_stream2(uint64_t* p, u64x2_u v) {
#if __HIP_DEVICE_COMPILE__
  __builtin_nontemporal_store(v, (g_u64x2*)p);
#else
  *(g_u64x2*)p = v;
#endif
}

template <typename T>
__device__ __forceinline__ uint64_t ld_sym(const void* text, uint64_t i) {
  return static_cast<const T*>(text)[i];
}

struct K1Params {
  const void* text;          // text mode if planes_in == nullptr
  uint32_t width;
  const uint64_t* planes_in; // plane mode: plane k at planes_in + k*in_stride
  uint64_t in_stride;
  uint64_t n;
  uint32_t L, l0, g, np;     // np = L - l0 planes carried, first g are built
  uint64_t* out_level0;      // final level l0
  uint64_t* temp_levels;     // sub-level j (1..g-1) at temp_levels + (j-1)*temp_stride
  uint64_t* temp_payload;    // payload plane k-g at temp_payload + (k-g)*temp_stride
  uint64_t temp_stride;
  uint32_t* counts;          // counts[c*nslot + slot]
  uint32_t nslot;
  unsigned long long* zeros_out;  // zeros of sub-level j accumulate into zeros_out[j]
  unsigned long long* sym_or;
  uint64_t nchunks;
  // batch mode (b_chunk_base != nullptr): text t owns chunks
  // [b_chunk_base[t], b_chunk_base[t+1]); its level 0 starts at word
  // b_out0[t] of out_level0 and its zero counts at b_zeros[t*kMaxGroup]
  const uint64_t* b_chunk_base;
  const uint64_t* b_text;
  const uint64_t* b_n;
  const uint64_t* b_out0;
  unsigned long long* b_zeros;
  uint32_t ntexts;
  unsigned int* work;        // dynamic chunk queue (atomic counter), or nullptr
  uint32_t grab;             // chunks taken from the queue at a time
  uint32_t* b_chunk_text;    // batch mode: text of every chunk (written here)
  uint32_t* chunk_sc;        // super-chunk passes followed by a reorder: every
                             // chunk's counts of the 2^G buckets of the order
                             // after the pass, [c * 2^G + q]
  uint8_t* lane_sc;          // or nullptr: every lane's counts of the G-bit field
                             // values of its 64 symbols, byte (c * 64 + lane) * 16 + f
  const uint8_t* lut;        // 64 KiB compress table in device memory
};

// Where a chunk's symbols come from and where its level-l0 word goes.
struct ChunkView {
  const void* text;          // the chunk's text
  uint64_t base;             // first symbol of the chunk inside its text
  uint64_t n;                // the text's length
  uint64_t* lvl0;            // level-l0 word of the chunk's first 64 symbols
  unsigned long long* zout;  // zero counters of the chunk's text
  uint32_t t;                // the chunk's text (batch mode)
};

// `hint` is the text of the wave's previous chunk: chunks arrive in
// increasing order within a grab, so a few forward steps usually find the
// text; otherwise a binary search over the chunk bases.
__device__ __forceinline__ ChunkView chunk_view(const K1Params& p, uint64_t c, uint32_t& hint) {
  ChunkView v;
  if (!p.b_chunk_base) {
    v.text = p.text;
    v.base = c * kChunkSyms;
    v.n = p.n;
    v.lvl0 = p.out_level0 + c * kChunkWords;
    v.zout = p.zeros_out;
    v.t = 0;
    return v;
  }
  uint32_t lo = hint;
  if (!(p.b_chunk_base[lo] <= c && c < p.b_chunk_base[lo + 1])) {
    int steps = 0;
    while (steps < 4 && lo + 1 < p.ntexts && p.b_chunk_base[lo + 1] <= c) {
      ++lo;
      ++steps;
    }
    if (!(p.b_chunk_base[lo] <= c && c < p.b_chunk_base[lo + 1])) {
      lo = 0;
      uint32_t hi = p.ntexts;  // chunk_base[lo] <= c < chunk_base[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p.b_chunk_base[mid] <= c) lo = mid;
        else hi = mid;
      }
    }
  }
  hint = lo;
  const uint64_t lc = c - p.b_chunk_base[lo];
  v.text = reinterpret_cast<const void*>(p.b_text[lo]);
  v.base = lc * kChunkSyms;
  v.n = p.b_n[lo];
  v.lvl0 = p.out_level0 + p.b_out0[lo] + lc * kChunkWords;
  v.zout = p.b_zeros + uint64_t(lo) * kMaxGroup;
  v.t = lo;
  return v;
}

// Gather bit s (0 = least significant) of 4 packed bytes into 4 bits.
// (v_dot4_u32_u8 multiplies the masked bytes by their place values 1,2,4,8.)
__device__ __forceinline__ uint32_t gather4(uint32_t x4, uint32_t s) {
  return __builtin_amdgcn_udot4((x4 >> s) & 0x01010101u, 0x08040201u, 0u, false);
}

template <int NP>
__device__ __forceinline__ void planes_from_bytes(const K1Params& p, const uint4* d, uint64_t* acc,
                                                  uint64_t& my_or);

template <int NP, typename T>
__device__ __forceinline__ void load_planes_text(const K1Params& p, const void* text, uint64_t e0, int cnt,
                                                 uint64_t* acc, uint64_t& my_or) {
  // symbol bit of plane k: s_k = L-1-l0-k ; byte index s_k>>3, bit s_k&7
  const auto tx = gmem(static_cast<const T*>(text));
  auto planes4 = [&](const uint64_t* v, int q) {
    my_or |= v[0] | v[1] | v[2] | v[3];
    uint32_t cur_bi = 0xffffffffu, x4 = 0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      if (k < int(p.np)) {
        const uint32_t s = p.L - 1 - p.l0 - k;
        const uint32_t bi = s >> 3;
        if (bi != cur_bi) {
          cur_bi = bi;
          x4 = uint32_t((v[0] >> (8 * bi)) & 0xff) | (uint32_t((v[1] >> (8 * bi)) & 0xff) << 8) |
               (uint32_t((v[2] >> (8 * bi)) & 0xff) << 16) | (uint32_t((v[3] >> (8 * bi)) & 0xff) << 24);
        }
        acc[k] |= uint64_t(gather4(x4, s & 7)) << (4 * q);
      }
    }
  };
  if constexpr (sizeof(T) == 2 || sizeof(T) == 4) {
    // a full lane at a 16-byte aligned address: its 64 symbols in batches of
    // 64 bytes, each batch's four 16-byte loads issued together; a plane's
    // byte of four symbols is picked from the loaded dwords with v_perm
    // (one for 2-byte symbols, three for 4-byte ones)
    const T* lp = static_cast<const T*>(text) + e0;
    if (cnt == 64 && (reinterpret_cast<uintptr_t>(lp) & 15) == 0) {
      constexpr int SB = 64 / int(sizeof(T));  // symbols per batch
      const auto g4 = gmem(reinterpret_cast<const uint4*>(lp));
      // every plane's bit in one byte of the symbol (a pass's 4 levels
      // usually are): that byte of four symbols with one v_perm, then the
      // byte planes' chained dot products (8 symbols per plane byte)
      const uint32_t s0 = p.L - 1 - p.l0;  // plane 0's bit
      const bool one_byte = (s0 >> 3) == ((s0 + 1 - p.np) >> 3);
#pragma unroll 1
      for (int b = 0; b < 64 / SB; ++b) {
        uint32_t u[16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint4 w = g4[4 * b + i];
          u[4 * i] = w.x;
          u[4 * i + 1] = w.y;
          u[4 * i + 2] = w.z;
          u[4 * i + 3] = w.w;
        }
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) o |= u[i];
        my_or |= sizeof(T) == 2 ? ((o | (o >> 16)) & 0xffffu) : o;
        if (one_byte) {
          const uint32_t bi = s0 >> 3;
          uint32_t xq[SB / 4];  // byte bi of symbols 4g..4g+3 of the batch
#pragma unroll
          for (int g = 0; g < SB / 4; ++g) {
            if constexpr (sizeof(T) == 2) {
              xq[g] = __builtin_amdgcn_perm(u[2 * g + 1], u[2 * g], 0x06040200u + bi * 0x01010101u);
            } else {
              const uint32_t lo = __builtin_amdgcn_perm(u[4 * g + 1], u[4 * g], 0x0c0c0400u + bi * 0x0101u);
              const uint32_t hi = __builtin_amdgcn_perm(u[4 * g + 3], u[4 * g + 2], 0x0c0c0400u + bi * 0x0101u);
              xq[g] = lo | (hi << 16);
            }
          }
#pragma unroll
          for (int k = 0; k < NP; ++k) {
            if (k < int(p.np)) {
              const uint32_t sft = (s0 - k) & 7;
              const uint32_t m = 0x01010101u << sft;
              uint32_t part = 0;  // the plane's SB bits of this batch
#pragma unroll
              for (int h = 0; h < SB / 8; ++h)
                part |= (__builtin_amdgcn_udot4(xq[2 * h + 1] & m, 0x80402010u,
                                                __builtin_amdgcn_udot4(xq[2 * h] & m, 0x08040201u, 0u, false),
                                                false) >>
                         sft)
                        << (8 * h);
              acc[k] |= uint64_t(part) << (SB * b);
            }
          }
          continue;
        }
#pragma unroll
        for (int g = 0; g < SB / 4; ++g) {
          uint32_t cur_bi = 0xffffffffu, x4 = 0;
#pragma unroll
          for (int k = 0; k < NP; ++k) {
            if (k < int(p.np)) {
              const uint32_t sb = p.L - 1 - p.l0 - k;
              const uint32_t bi = sb >> 3;
              if (bi != cur_bi) {
                cur_bi = bi;
                if constexpr (sizeof(T) == 2) {
                  // symbols 4g..4g+3 are the halves of u[2g], u[2g + 1]
                  x4 = __builtin_amdgcn_perm(u[2 * g + 1], u[2 * g], 0x06040200u + bi * 0x01010101u);
                } else {
                  const uint32_t lo = __builtin_amdgcn_perm(u[4 * g + 1], u[4 * g], 0x0c0c0400u + bi * 0x0101u);
                  const uint32_t hi = __builtin_amdgcn_perm(u[4 * g + 3], u[4 * g + 2], 0x0c0c0400u + bi * 0x0101u);
                  x4 = lo | (hi << 16);
                }
              }
              acc[k] |= uint64_t(gather4(x4, sb & 7)) << (4 * (b * (SB / 4) + g));
            }
          }
        }
      }
      return;
    }
  }
  if constexpr (sizeof(T) == 8) {
    // a full lane at a 16-byte aligned address: its 64 symbols in batches of
    // 64 bytes, each batch's four 16-byte loads issued together
    const T* lp = static_cast<const T*>(text) + e0;
    if (cnt == 64 && (reinterpret_cast<uintptr_t>(lp) & 15) == 0) {
      constexpr int SB = 64 / int(sizeof(T));  // symbols per batch
      const auto g4 = gmem(reinterpret_cast<const uint4*>(lp));
#pragma unroll 1
      for (int b = 0; b < 64 / SB; ++b) {
        uint32_t u[16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint4 w = g4[4 * b + i];
          u[4 * i] = w.x;
          u[4 * i + 1] = w.y;
          u[4 * i + 2] = w.z;
          u[4 * i + 3] = w.w;
        }
#pragma unroll
        for (int g = 0; g < SB / 4; ++g) {
          uint64_t v[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int i = 4 * g + e;  // symbol of the batch
            v[e] = uint64_t(u[2 * i]) | (uint64_t(u[2 * i + 1]) << 32);
          }
          planes4(v, b * (SB / 4) + g);
        }
      }
      return;
    }
  }
  if constexpr (sizeof(T) == 1) {
    const uint8_t* lp = static_cast<const uint8_t*>(text) + e0;
    const bool a16 = (reinterpret_cast<uintptr_t>(lp) & 15) == 0;
    if (p.L - 1 - p.l0 < 8) {
      // every plane's bit lies in the byte: the lane's 64 bytes (zero past
      // the chunk's end) with all loads issued together, then the planes as
      // for whole chunks
      uint4 d[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        if (cnt >= 16 * g + 16 && a16) {
          d[g] = *gmem(reinterpret_cast<const uint4*>(lp + 16 * g));
        } else {
          uint32_t w[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int q = 4 * g + i;
            if (cnt >= 4 * q + 4) {
              w[i] = *gmem(reinterpret_cast<const uint32_t*>(lp + 4 * q));
            } else {
              uint32_t x = 0;
              for (int e = 0; e < 4; ++e)
                if (4 * q + e < cnt) x |= uint32_t(tx[e0 + 4 * q + e]) << (8 * e);
              w[i] = x;
            }
          }
          d[g] = make_uint4(w[0], w[1], w[2], w[3]);
        }
      }
      planes_from_bytes<NP>(p, d, acc, my_or);
      return;
    }
    // (symbol bits above the byte: planes of zeros among them) four groups
    // of 16 bytes, each group's loads issued together
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
      uint32_t w[4];
      if (cnt >= 16 * g + 16) {
        if (a16) {
          const uint4 x = *gmem(reinterpret_cast<const uint4*>(lp + 16 * g));
          w[0] = x.x; w[1] = x.y; w[2] = x.z; w[3] = x.w;
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) w[i] = *gmem(reinterpret_cast<const uint32_t*>(lp + 16 * g + 4 * i));
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int q = 4 * g + i;
          if (cnt >= 4 * q + 4) {
            w[i] = *gmem(reinterpret_cast<const uint32_t*>(lp + 4 * q));
          } else {
            uint32_t x = 0;
            for (int e = 0; e < 4; ++e)
              if (4 * q + e < cnt) x |= uint32_t(tx[e0 + 4 * q + e]) << (8 * e);
            w[i] = x;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint64_t v[4] = {w[i] & 0xffu, (w[i] >> 8) & 0xffu, (w[i] >> 16) & 0xffu, w[i] >> 24};
        planes4(v, 4 * g + i);
      }
    }
    return;
  }
#pragma unroll 1
  for (int q = 0; q < 16; ++q) {
    uint64_t v[4];
    if (cnt >= 4 * q + 4) {
      if constexpr (sizeof(T) == 1) {
        const uint32_t x = *gmem(reinterpret_cast<const uint32_t*>(static_cast<const T*>(text) + e0 + 4 * q));
        v[0] = x & 0xff; v[1] = (x >> 8) & 0xff; v[2] = (x >> 16) & 0xff; v[3] = x >> 24;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = tx[e0 + 4 * q + e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = (4 * q + e < cnt) ? uint64_t(tx[e0 + 4 * q + e]) : 0;
    }
    planes4(v, q);
  }
}

// 64 bytes (one thread's symbols) -> bit planes, symbol bit L-1-l0-k -> plane k.
template <int NP>
__device__ __forceinline__ void planes_from_bytes(const K1Params& p, const uint4* d, uint64_t* acc,
                                                  uint64_t& my_or) {
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    x[4 * i] = d[i].x;
    x[4 * i + 1] = d[i].y;
    x[4 * i + 2] = d[i].z;
    x[4 * i + 3] = d[i].w;
  }
  uint32_t o = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) o |= x[q];
  my_or |= (o | (o >> 8) | (o >> 16) | (o >> 24)) & 0xffu;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const uint32_t sft = p.L - 1 - p.l0 - k;
    const uint32_t m = 0x01010101u << sft;
    // a plane whose bit none of the lane's 64 bytes has is zero (small
    // alphabets: a rare symbol's bit is clear in most lanes)
    if (k < int(p.np)) {
      if (!(o & m)) {
        acc[k] = 0;
      } else {
        // 8 symbols -> one byte of the plane: masked bit s of each byte
        // times its place value, two 4-byte dot products chained through
        // the accumulator; the result is scaled by 2^s.
        uint32_t b[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          b[q] = __builtin_amdgcn_udot4(x[2 * q + 1] & m, 0x80402010u,
                                        __builtin_amdgcn_udot4(x[2 * q] & m, 0x08040201u, 0u, false), false) >>
                 sft;
        const uint32_t lo = b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
        const uint32_t hi = b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24);
        acc[k] = (uint64_t(hi) << 32) | lo;
      }
    }
  }
}

// Inclusive wave64 prefix sum with DPP row shifts and row broadcasts.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Sums over each row of 16 lanes, in the row's last lane (DPP row shifts,
// no LDS traffic); other lanes hold partial sums.
__device__ __forceinline__ uint32_t row_sum16(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  return v;
}

// Orders this wave's LDS accesses (lanes exchange data through LDS).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Exclusive wave scan of two counters plus their totals.
__device__ __forceinline__ void wave_exscan2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb,
                                             uint32_t& ta, uint32_t& tb) {
  const uint32_t ia = wave_incl_scan(a), ib = wave_incl_scan(b);
  ta = __builtin_amdgcn_readlane(ia, 63);
  tb = __builtin_amdgcn_readlane(ib, 63);
  ea = ia - a;
  eb = ib - b;
}

// Copies the device compress table into LDS (16-byte loads), then a barrier.
template <int WAVES>
__device__ __forceinline__ void load_lut(uint8_t* lut, const uint8_t* src) {
  const auto g = gmem(reinterpret_cast<const uint4*>(src));
  uint4* l = reinterpret_cast<uint4*>(lut);
  for (int i = threadIdx.x; i < 65536 / 16; i += 64 * WAVES) l[i] = g[i];
  __syncthreads();
}

__global__ void k_make_lut(uint8_t* lut) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 65536) lut[i] = uint8_t(compress8(i & 255u, i >> 8));
}

template <int NP>
struct WaveLds {
  static constexpr int kB = (1 << (NP < kMaxGroup ? NP : kMaxGroup)) + 1;
  uint64_t buf[2][NP][64];
  uint32_t ooff[65];
  uint32_t bst[2][kB];
};

template <int NP>
constexpr int waves_for() {
  return NP <= 4 ? 16 : (NP <= 8 ? 8 : (NP <= 12 ? 6 : 4));
}

// One wave builds one chunk (64 words per plane, one word per lane); the
// workgroup's waves share the compress table and never synchronise with
// each other after it is filled.
template <int NP>
__global__ __launch_bounds__(64 * waves_for<NP>()) void k_local_build(K1Params p) {
  constexpr int WAVES = waves_for<NP>();
  __shared__ __align__(16) uint8_t lut[65536];
  __shared__ WaveLds<NP> wl[WAVES];
  __shared__ unsigned long long red[kMaxGroup + 1];  // block sums: zeros per sub-level, symbol OR
  if (threadIdx.x <= kMaxGroup) red[threadIdx.x] = 0;

  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  load_lut<WAVES>(lut, p.lut);
  WaveLds<NP>& L = wl[wid];
  uint64_t my_or = 0;
  // zeros per sub-level over this wave's recent chunks (flushed before
  // 2^19 chunks could overflow 32 bits)
  constexpr int kZ = NP < kMaxGroup ? NP : kMaxGroup;
  uint32_t zsum[kZ];
#pragma unroll
  for (int j = 0; j < kZ; ++j) zsum[j] = 0;
  uint32_t zchunks = 0;

  const uint64_t gw = uint64_t(blockIdx.x) * WAVES + wid;
  const uint64_t tw = uint64_t(gridDim.x) * WAVES;
  // next chunk of this wave: static round robin, or grabs of kGrab
  // consecutive chunks from the device work queue
  uint64_t grab_end = 0;
  auto next_chunk = [&](uint64_t prev, bool first) -> uint64_t {
    if (!p.work) return first ? gw : prev + tw;
    if (!first && prev + 1 < grab_end) return prev + 1;
    uint32_t q = 0;
    if (lane == 0) q = atomicAdd(p.work, 1u);
    q = __builtin_amdgcn_readfirstlane(q);
    const uint64_t c0 = uint64_t(q) * p.grab;
    grab_end = min<uint64_t>(c0 + p.grab, p.nchunks);
    return c0;
  };
  uint32_t hint = 0;
  // zero counts accumulate per wave while the chunks belong to one text
  unsigned long long* zdst = nullptr;
  auto flush_zeros = [&]() {
    if (zdst && lane == 0) {
#pragma unroll
      for (int j = 0; j < kZ; ++j)
        if (j < int(p.g) && zsum[j]) atomicAdd(zdst + j, (unsigned long long)zsum[j]);
    }
#pragma unroll
    for (int j = 0; j < kZ; ++j) zsum[j] = 0;
    zchunks = 0;
  };
  // byte texts at 16-byte aligned addresses take 16-byte loads and prefetch
  const bool fast_mode = !p.planes_in && p.width == 1;
  auto full_chunk = [&](const ChunkView& v) {
    return v.base + kChunkSyms <= v.n && (reinterpret_cast<uintptr_t>(v.text) & 15) == 0;
  };
  auto load_chunk = [&](const ChunkView& v, uint4* d) {
    const auto src =
        gmem(reinterpret_cast<const uint4*>(static_cast<const uint8_t*>(v.text) + v.base + 64ull * lane));
    d[0] = src[0];
    d[1] = src[1];
    d[2] = src[2];
    d[3] = src[3];
  };
  uint4 pf[4];
  uint64_t c = next_chunk(0, true);
  ChunkView cv;
  if (c < p.nchunks) {
    cv = chunk_view(p, c, hint);
    if (fast_mode && full_chunk(cv)) load_chunk(cv, pf);
  }

  while (c < p.nchunks) {
    wave_sync();
    const uint64_t cn = next_chunk(c, false);
    ChunkView nv;
    if (cn < p.nchunks) nv = chunk_view(p, cn, hint);
    if (p.b_chunk_text && lane == 0) p.b_chunk_text[c] = cv.t;
    if (cv.zout != zdst || zchunks == (1u << 19)) {
      flush_zeros();
      zdst = cv.zout;
    }
    ++zchunks;
    const bool have_cur = fast_mode && full_chunk(cv);
    const uint32_t Ec = uint32_t(min<uint64_t>(kChunkSyms, cv.n - cv.base));
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * lane));

    uint64_t acc[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) acc[k] = 0;
    if (p.planes_in) {
      if (lane < int(nw)) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
          if (k < int(p.np)) acc[k] = brev64(p.planes_in[k * p.in_stride + c * P + lane]);
      }
    } else if (have_cur) {
      planes_from_bytes<NP>(p, pf, acc, my_or);
      if (cn < p.nchunks && fast_mode && full_chunk(nv)) load_chunk(nv, pf);  // prefetch
    } else if (cnt > 0) {
      const uint64_t e0 = cv.base + 64ull * lane;
      switch (p.width) {
        case 1: load_planes_text<NP, uint8_t>(p, cv.text, e0, cnt, acc, my_or); break;
        case 2: load_planes_text<NP, uint16_t>(p, cv.text, e0, cnt, acc, my_or); break;
        case 4: load_planes_text<NP, uint32_t>(p, cv.text, e0, cnt, acc, my_or); break;
        default: load_planes_text<NP, uint64_t>(p, cv.text, e0, cnt, acc, my_or); break;
      }
      if (cn < p.nchunks && fast_mode && full_chunk(nv)) load_chunk(nv, pf);
    } else if (cn < p.nchunks && fast_mode && full_chunk(nv)) {
      load_chunk(nv, pf);
    }
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (k < int(p.np)) L.buf[0][k][lane] = acc[k];
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (lane == 0) {
      L.bst[0][0] = 0;
      L.bst[0][1] = Ec;
    }
    wave_sync();

    int cur = 0;
    uint32_t nb = 1;
    for (uint32_t j = 0; j < p.g; ++j) {
      const uint64_t m = L.buf[cur][j][lane];
      if (lane < int(nw)) {
        if (j == 0) gmem(cv.lvl0)[lane] = brev64(m);
        else p.temp_levels[(j - 1) * p.temp_stride + c * P + lane] = brev64(m);
      }
      const uint64_t m1 = m & vm, m0 = ~m & vm;
      const uint32_t o = __popcll(m1), z = __popcll(m0);
      const uint32_t zinc = wave_incl_scan(z);
      const uint32_t zoff = zinc - z, Zt = __builtin_amdgcn_readlane(zinc, 63);
      const uint32_t oof = min(64u * lane, Ec) - zoff;  // every earlier word is full
      const uint32_t Ot = Ec - Zt;
      zsum[j] += Zt;
      const bool more = (j + 1 < p.g) || (p.np > p.g);
      if (!more) break;
      L.ooff[lane] = oof;
      if (lane == 63) L.ooff[64] = Ot;
      const int nxt = cur ^ 1;
#pragma unroll
      for (int k = 1; k < NP; ++k)
        if (k > int(j) && k < int(p.np)) L.buf[nxt][k][lane] = 0;
      wave_sync();
#pragma unroll
      for (int k = 1; k < NP; ++k) {
        if (k > int(j) && k < int(p.np)) {
          const uint64_t x = L.buf[cur][k][lane];
          or_bits(L.buf[nxt][k], zoff, compress64(x, m0, lut), z);
          or_bits(L.buf[nxt][k], Zt + oof, compress64(x, m1, lut), o);
        }
      }
      // bucket counts of sub-level j+1: zeros and ones of plane j inside
      // every current bucket; new order = zeros of all buckets, then ones.
      const uint32_t slot = (2u << j) - 2u;
      uint32_t zcar = 0, ocar = 0;
      for (uint32_t b0 = 0; b0 < nb; b0 += 64) {
        const uint32_t bi = b0 + lane;
        uint32_t zc = 0, oc = 0;
        if (bi < nb) {
          const uint32_t s0 = L.bst[cur][bi], e0 = L.bst[cur][bi + 1];
          oc = rank1_lds(L.buf[cur][j], L.ooff, e0) - rank1_lds(L.buf[cur][j], L.ooff, s0);
          zc = (e0 - s0) - oc;
          p.counts[c * p.nslot + slot + bi] = zc;
          p.counts[c * p.nslot + slot + nb + bi] = oc;
        }
        uint32_t ez, eo, tz, to;
        wave_exscan2(zc, oc, ez, eo, tz, to);
        if (bi < nb) {
          L.bst[nxt][bi] = zcar + ez;
          L.bst[nxt][nb + bi] = Zt + ocar + eo;
        }
        zcar += tz;
        ocar += to;
      }
      if (lane == 0) L.bst[nxt][2 * nb] = Ec;
      wave_sync();
      cur = nxt;
      nb <<= 1;
    }
    if (lane < int(nw)) {
#pragma unroll
      for (int k = 0; k < NP; ++k)
        if (k >= int(p.g) && k < int(p.np))
          p.temp_payload[(k - p.g) * p.temp_stride + c * P + lane] = brev64(L.buf[cur][k][lane]);
    }
    c = cn;
    cv = nv;
  }
  // one text: the block's waves add up in LDS, then one global atomic per
  // counter per block (thousands of same-address atomics serialise)
  if (p.b_chunk_base) flush_zeros();
  else if (lane == 0) {
#pragma unroll
    for (int j = 0; j < kZ; ++j)
      if (j < int(p.g) && zsum[j]) atomicAdd(&red[j], (unsigned long long)zsum[j]);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) my_or |= __shfl_xor(my_or, d, 64);
  if (lane == 0 && my_or) atomicOr(&red[kMaxGroup], (unsigned long long)my_or);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (!p.b_chunk_base)
      for (int j = 0; j < kZ; ++j)
        if (red[j]) atomicAdd(p.zeros_out + j, red[j]);
    if (!p.planes_in && red[kMaxGroup]) atomicOr(p.sym_or, red[kMaxGroup]);
  }
}

// ---- K1 for few levels -----------------------------------------------------
// All G levels of the pass come from the text (no payload planes), G <= 4.
// Every plane stays in registers.  Sub-level j of a chunk is the chunk's
// symbols ordered by bucket q = (bit j-1 .. bit 0 of the symbol's top
// levels), stable inside a bucket -- the order the level-by-level stable
// partition produces -- so it is assembled directly: for every bucket q the
// lane's bits of plane j under the bucket mask are compressed and OR-ed into
// the wave's LDS word array at the bucket's start plus the lane's exclusive
// prefix.  Bucket sizes and prefixes come from packed 16-bit wave scans.

// pext of a 32-bit word through the byte table (lut at LDS address 0).
__device__ __forceinline__ uint32_t compress32(uint32_t x, uint32_t m, const uint8_t* lut) {
  uint32_t r = lut[__builtin_amdgcn_perm(m, x, 0x0c0c0400u)];
  uint32_t sft = __builtin_popcount(m & 0xffu);
  r |= uint32_t(lut[__builtin_amdgcn_perm(m, x, 0x0c0c0501u)]) << sft;
  sft = __builtin_popcount(m & 0xffffu);
  r |= uint32_t(lut[__builtin_amdgcn_perm(m, x, 0x0c0c0602u)]) << sft;
  sft = __builtin_popcount(m & 0xffffffu);
  r |= uint32_t(lut[__builtin_amdgcn_perm(m, x, 0x0c0c0703u)]) << sft;
  return r;
}

__device__ __forceinline__ uint64_t compress64h(uint64_t x, uint64_t m, const uint8_t* lut) {
  const uint32_t lo = compress32(uint32_t(x), uint32_t(m), lut);
  const uint32_t hi = compress32(uint32_t(x >> 32), uint32_t(m >> 32), lut);
  return uint64_t(lo) | (uint64_t(hi) << __builtin_popcount(uint32_t(m)));
}

// A lane's bits of plane x under bucket mask mq, packed.  When all of the
// wave's symbols fall in the bucket (in_bucket == wave_syms, both
// wave-uniform) every lane's mask is its valid-symbol mask, a prefix of the
// word, and x holds no bits past it: the bits are x as it is, and the wave
// skips the table compress.
__device__ __forceinline__ uint64_t bucket_bits(uint64_t x, uint64_t mq, uint32_t in_bucket, uint32_t wave_syms,
                                                const uint8_t* lut) {
  return in_bucket == wave_syms ? x : compress64h(x, mq, lut);
}

// With SC (a pass followed by a symbol reorder) the bucket counts of the
// order after all G levels (2^G buckets, the next pass's level order) are
// produced as well, as "sub-level G".
// WIDE: symbols of 2, 4 or 8 bytes; otherwise bytes.
template <int G, int SC, bool WIDE>
__global__ __launch_bounds__(64 * 16) void k_local_small(K1Params p) {
  constexpr int WAVES = 16;
  constexpr int NSUB = G > 1 ? G - 1 : 1;  // sub-levels 1..G-1 (bits assembled)
  constexpr int JC = G - 1 + SC;           // last sub-level with bucket counts
  constexpr int NQ = 1 << JC;              // buckets of that sub-level
  constexpr int NF = NQ + (SC ? 0 : 1);    // scan fields: bucket counts (+ last level's zeros)
  constexpr int NS = (NF + 1) / 2;         // packed scans
  __shared__ __align__(16) uint8_t lut[65536];
  __shared__ uint64_t sbuf[WAVES][NSUB][64];
  __shared__ unsigned long long red[G + 1];  // block sums: zeros per level, symbol OR
  if (threadIdx.x <= G) red[threadIdx.x] = 0;

  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  load_lut<WAVES>(lut, p.lut);
  uint64_t(&buf)[NSUB][64] = sbuf[wid];

  uint32_t zsum[G];
#pragma unroll
  for (int j = 0; j < G; ++j) zsum[j] = 0;
  uint32_t zchunks = 0;
  uint64_t my_or = 0;
  const uint64_t gw = uint64_t(blockIdx.x) * WAVES + wid;
  const uint64_t tw = uint64_t(gridDim.x) * WAVES;
  uint64_t grab_end = 0;
  auto next_chunk = [&](uint64_t prev, bool first) -> uint64_t {
    if (!p.work) return first ? gw : prev + tw;
    if (!first && prev + 1 < grab_end) return prev + 1;
    uint32_t q = 0;
    if (lane == 0) q = atomicAdd(p.work, 1u);
    q = __builtin_amdgcn_readfirstlane(q);
    const uint64_t c0 = uint64_t(q) * p.grab;
    grab_end = min<uint64_t>(c0 + p.grab, p.nchunks);
    return c0;
  };
  uint32_t hint = 0;
  unsigned long long* zdst = nullptr;
  auto flush_zeros = [&]() {
    if (zdst && lane == 0) {
#pragma unroll
      for (int j = 0; j < G; ++j)
        if (zsum[j]) atomicAdd(zdst + j, (unsigned long long)zsum[j]);
    }
#pragma unroll
    for (int j = 0; j < G; ++j) zsum[j] = 0;
    zchunks = 0;
  };
  const bool fast_mode = !WIDE;
  auto full_chunk = [&](const ChunkView& v) {
    return v.base + kChunkSyms <= v.n && (reinterpret_cast<uintptr_t>(v.text) & 15) == 0;
  };
  auto load_chunk = [&](const ChunkView& v, uint4* d) {
    const auto src =
        gmem(reinterpret_cast<const uint4*>(static_cast<const uint8_t*>(v.text) + v.base + 64ull * lane));
    d[0] = src[0];
    d[1] = src[1];
    d[2] = src[2];
    d[3] = src[3];
  };
  uint4 pf[4];
  uint64_t c = next_chunk(0, true);
  ChunkView cv;
  if (c < p.nchunks) {
    cv = chunk_view(p, c, hint);
    if (fast_mode && full_chunk(cv)) load_chunk(cv, pf);
  }

  while (c < p.nchunks) {
    wave_sync();  // the previous chunk's LDS reads are done
    const uint64_t cn = next_chunk(c, false);
    ChunkView nv;
    if (cn < p.nchunks) nv = chunk_view(p, cn, hint);
    if (p.b_chunk_text && lane == 0) p.b_chunk_text[c] = cv.t;
    if (cv.zout != zdst || zchunks == (1u << 19)) {
      flush_zeros();
      zdst = cv.zout;
    }
    ++zchunks;
    const bool have_cur = fast_mode && full_chunk(cv);
    const uint32_t Ec = uint32_t(min<uint64_t>(kChunkSyms, cv.n - cv.base));
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * lane));

    uint64_t P0[G];
#pragma unroll
    for (int k = 0; k < G; ++k) P0[k] = 0;
    if (have_cur) {
      planes_from_bytes<G>(p, pf, P0, my_or);
      if (cn < p.nchunks && full_chunk(nv)) load_chunk(nv, pf);  // prefetch
    } else {
      if (cnt > 0) {
        const uint64_t e0 = cv.base + 64ull * lane;
        if constexpr (!WIDE) {
          load_planes_text<G, uint8_t>(p, cv.text, e0, cnt, P0, my_or);
        } else {
          switch (p.width) {
            case 2: load_planes_text<G, uint16_t>(p, cv.text, e0, cnt, P0, my_or); break;
            case 4: load_planes_text<G, uint32_t>(p, cv.text, e0, cnt, P0, my_or); break;
            default: load_planes_text<G, uint64_t>(p, cv.text, e0, cnt, P0, my_or); break;
          }
        }
      }
      if (cn < p.nchunks && fast_mode && full_chunk(nv)) load_chunk(nv, pf);
    }
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (lane < int(nw)) st_stream(cv.lvl0 + lane, brev64(P0[0]));
#pragma unroll
    for (int j = 0; j < NSUB; ++j) buf[j][lane] = 0;

    // bucket masks of sub-level JC, M[q] = prod_i (bit i of q ? P_i : ~P_i)
    uint64_t M[NQ];
    M[0] = vm;
#pragma unroll
    for (int i = 0; i < JC; ++i) {
#pragma unroll
      for (int q = (1 << i) - 1; q >= 0; --q) {
        M[q + (1 << i)] = M[q] & P0[i];
        M[q] = M[q] & ~P0[i];
      }
    }
    uint32_t f[2 * NS];
#pragma unroll
    for (int q = 0; q < NQ; ++q) f[q] = __popcll(M[q]);
    if constexpr (!SC) f[NQ] = __popcll(~P0[G - 1] & vm);
    if (NF < 2 * NS) f[2 * NS - 1] = 0;
    uint32_t ex[2 * NS], tot[2 * NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const uint32_t v = f[2 * s2] | (f[2 * s2 + 1] << 16);
      const uint32_t inc = wave_incl_scan(v);
      const uint32_t t = __builtin_amdgcn_readlane(inc, 63);
      const uint32_t e = inc - v;
      ex[2 * s2] = e & 0xffffu;
      ex[2 * s2 + 1] = e >> 16;
      tot[2 * s2] = t & 0xffffu;
      tot[2 * s2 + 1] = t >> 16;
    }
    // zeros of every level: level j < G-1 has bit j == 0 in buckets q < 2^j
    // of sub-level j+1, i.e. q with bit j clear among the last sub-level's
    {
#pragma unroll
      for (int j = 0; j < JC; ++j) {
        uint32_t z = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (!((q >> j) & 1)) z += tot[q];
        zsum[j] += z;
      }
      if constexpr (!SC) zsum[G - 1] += tot[NQ];
    }

    // sub-level j (1..G-1): buckets q < 2^j; counts/prefixes are sums over
    // the last sub-level's buckets that extend q
    uint32_t cslot[NQ > 1 ? 2 * NQ - 2 : 1];
#pragma unroll
    for (int j = 1; j <= JC; ++j) {
      const int nb = 1 << j;
      uint32_t start = 0;
#pragma unroll
      for (int q = 0; q < nb; ++q) {
        uint32_t T = 0, E = 0;
#pragma unroll
        for (int r = q; r < NQ; r += nb) {
          T += tot[r];
          E += ex[r];
        }
        if (j < G && T) {  // (T: the wave's symbols of the bucket, wave-uniform)
          uint64_t mq = vm;
#pragma unroll
          for (int i = 0; i < j; ++i) mq &= ((q >> i) & 1) ? P0[i] : ~P0[i];
          const uint32_t len = __popcll(mq);
          or_bits(buf[j - 1], start + E, bucket_bits(P0[j], mq, T, Ec, lut), len);
        }
        cslot[nb - 2 + q] = T;
        start += T;
      }
    }
    wave_sync();
    if (lane < int(nw)) {
#pragma unroll
      for (int j = 1; j < G; ++j)
        st_stream(p.temp_levels + (j - 1) * p.temp_stride + c * P + lane, brev64(buf[j - 1][lane]));
    }
    if (G > 1) {
      uint32_t v = 0;
#pragma unroll
      for (int i = 0; i < 2 * NQ - 2; ++i) v = lane == i ? cslot[i] : v;
      if (lane < 2 * NQ - 2) gmem(p.counts)[c * p.nslot + lane] = v;
    }
    c = cn;
    cv = nv;
  }
  // one text: the block's waves add up in LDS, then one global atomic per
  // counter per block (thousands of same-address atomics serialise)
  if (p.b_chunk_base) flush_zeros();
  else if (lane == 0) {
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (zsum[j]) atomicAdd(&red[j], (unsigned long long)zsum[j]);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) my_or |= __shfl_xor(my_or, d, 64);
  if (lane == 0 && my_or) atomicOr(&red[G], (unsigned long long)my_or);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (!p.b_chunk_base)
      for (int j = 0; j < G; ++j)
        if (red[j]) atomicAdd(p.zeros_out + j, red[j]);
    if (red[G]) atomicOr(p.sym_or, red[G]);
  }
}

// ---- K1 over super-chunks ----------------------------------------------------
// One text, all G (2..4) levels from the text.  The block's 16 waves take 16
// consecutive chunks (a super-chunk of 64 Ki symbols) and the sub-levels are
// assembled for the whole super-chunk in LDS: a symbol of bucket q in wave w
// lands at (start of q in the super-chunk) + (q's count in waves < w) + (its
// rank among the wave's q symbols).  The super-chunk's image is then written
// out coalesced.  Runs of the stitch are 16 times longer than with one
// chunk per run: fewer table entries and fewer cache lines shared by two
// runs (each such line is fetched once per run).
constexpr int kSuper = 16;                                  // chunks per super-chunk
constexpr uint64_t kSuperSyms = uint64_t(kSuper) * kChunkSyms;
constexpr uint32_t kSuperWords = kSuper * kChunkWords;      // words per sub-level

// With SC (a pass followed by a symbol reorder) every wave also counts its
// chunk's 2^G buckets of the order after the pass (chunk_sc), and the unit's
// sums of them follow the sub-level counts in its count row.
template <int G, bool WIDE, int WAVES, bool SC = false>
__global__ __launch_bounds__(64 * WAVES) void k_local_super(K1Params p) {
  static_assert(G >= 2 && G <= 4, "two to four levels");
  constexpr uint64_t kUnitSyms = uint64_t(WAVES) * kChunkSyms;
  constexpr uint32_t kUnitWords = WAVES * kChunkWords;
  constexpr int NSUB = G - 1;      // sub-levels 1..G-1
  constexpr int JC = G - 1;        // last sub-level
  constexpr int NQ = 1 << JC;      // its buckets
  constexpr int NF = NQ + 1;       // bucket counts + the last level's zeros
  constexpr int NS = (NF + 1) / 2; // packed scans
  __shared__ __align__(16) uint8_t lut[65536];
  __shared__ uint64_t img[NSUB][kUnitWords];
  __shared__ uint32_t wt[WAVES][NQ];
  __shared__ uint32_t wsc[SC ? WAVES : 1][SC ? 2 * NQ : 1];  // per wave: next-order bucket counts
  __shared__ unsigned long long red[G + 1];  // block sums: zeros per level, symbol OR
  if (threadIdx.x <= G) red[threadIdx.x] = 0;
  for (uint32_t i = threadIdx.x; i < NSUB * kUnitWords; i += 64 * WAVES) (&img[0][0])[i] = 0;

  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int uwid = __builtin_amdgcn_readfirstlane(wid);
  load_lut<WAVES>(lut, p.lut);  // (ends with a barrier)

  uint32_t zsum[G];
#pragma unroll
  for (int j = 0; j < G; ++j) zsum[j] = 0;
  uint64_t my_or = 0;
  const uint64_t nsuper = (p.nchunks + WAVES - 1) / WAVES;
  const bool fast_mode = !WIDE && (reinterpret_cast<uintptr_t>(p.text) & 15) == 0;
  auto full_chunk = [&](uint64_t c) { return fast_mode && c < p.nchunks && (c + 1) * kChunkSyms <= p.n; };
  auto load_chunk = [&](uint64_t c, uint4* d) {
    const u32x4_t* src = reinterpret_cast<const u32x4_t*>(static_cast<const uint8_t*>(p.text) + c * kChunkSyms +
                                                          64ull * lane);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const u32x4_t v = ld_stream(src + i);
      d[i] = make_uint4(v.x, v.y, v.z, v.w);
    }
  };
  uint4 pf[4];
  uint64_t sc = blockIdx.x;
  if (sc < nsuper && full_chunk(sc * WAVES + wid)) load_chunk(sc * WAVES + wid, pf);

  for (; sc < nsuper; sc += gridDim.x) {
    const uint64_t c = sc * WAVES + wid;
    const uint64_t cn = (sc + gridDim.x) * WAVES + wid;
    const bool have = c < p.nchunks;
    const uint32_t Ec = have ? uint32_t(min<uint64_t>(kChunkSyms, p.n - c * kChunkSyms)) : 0u;
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * lane));

    uint64_t P0[G];
#pragma unroll
    for (int k = 0; k < G; ++k) P0[k] = 0;
    if (full_chunk(c)) {
      planes_from_bytes<G>(p, pf, P0, my_or);
      if (full_chunk(cn)) load_chunk(cn, pf);  // prefetch
    } else {
      if (cnt > 0) {
        const uint64_t e0 = c * kChunkSyms + 64ull * lane;
        if constexpr (!WIDE) {
          load_planes_text<G, uint8_t>(p, p.text, e0, cnt, P0, my_or);
        } else {
          switch (p.width) {
            case 2: load_planes_text<G, uint16_t>(p, p.text, e0, cnt, P0, my_or); break;
            case 4: load_planes_text<G, uint32_t>(p, p.text, e0, cnt, P0, my_or); break;
            default: load_planes_text<G, uint64_t>(p, p.text, e0, cnt, P0, my_or); break;
          }
        }
      }
      if (full_chunk(cn)) load_chunk(cn, pf);
    }
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (lane < int(nw)) st_stream(p.out_level0 + c * kChunkWords + lane, brev64(P0[0]));

    // bucket masks of the last sub-level, M[q] = prod_i (bit i of q ? P_i : ~P_i)
    uint64_t M[NQ];
    M[0] = vm;
#pragma unroll
    for (int i = 0; i < JC; ++i) {
#pragma unroll
      for (int q = (1 << i) - 1; q >= 0; --q) {
        M[q + (1 << i)] = M[q] & P0[i];
        M[q] = M[q] & ~P0[i];
      }
    }
    uint32_t f[2 * NS];
#pragma unroll
    for (int q = 0; q < NQ; ++q) f[q] = __popcll(M[q]);
    f[NQ] = __popcll(~P0[G - 1] & vm);
    if (NF < 2 * NS) f[2 * NS - 1] = 0;
    uint32_t ex[2 * NS], tot[2 * NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const uint32_t v = f[2 * s2] | (f[2 * s2 + 1] << 16);
      const uint32_t inc = wave_incl_scan(v);
      const uint32_t t = __builtin_amdgcn_readlane(inc, 63);
      const uint32_t e = inc - v;
      ex[2 * s2] = e & 0xffffu;
      ex[2 * s2 + 1] = e >> 16;
      tot[2 * s2] = t & 0xffffu;
      tot[2 * s2 + 1] = t >> 16;
    }
#pragma unroll
    for (int j = 0; j < JC; ++j) {
      uint32_t z = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (!((q >> j) & 1)) z += tot[q];
      zsum[j] += z;
    }
    zsum[G - 1] += tot[NQ];
    if (lane < NQ) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v = lane == q ? tot[q] : v;
      wt[wid][lane] = v;
    }
    if constexpr (SC) {
      // bucket q of the last sub-level splits by the last level's bit into
      // next-order buckets q and q + NQ (counted in pairs, 16 bits each)
      uint32_t lo = 0, hi = 0;
      uint32_t lw[4] = {0, 0, 0, 0};  // this lane's counts, byte f: field value f
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const uint32_t c1 = uint32_t(__popcll(M[q] & P0[G - 1]));
        const uint32_t v = (f[q] - c1) | (c1 << 16);  // (f[q]: all of bucket q)
        const uint32_t t = __builtin_amdgcn_readlane(wave_incl_scan(v), 63);
        lo = lane == q ? (t & 0xffffu) : lo;
        hi = lane == q ? (t >> 16) : hi;
        // next-order bucket r holds the symbols whose G-bit field is r reversed
        const int f0 = int(__builtin_bitreverse32(uint32_t(q)) >> (32 - G));
        const int f1 = int(__builtin_bitreverse32(uint32_t(q + NQ)) >> (32 - G));
        lw[f0 >> 2] |= (f[q] - c1) << (8 * (f0 & 3));
        lw[f1 >> 2] |= c1 << (8 * (f1 & 3));
      }
      if (p.lane_sc && c < p.nchunks)
        *gmem(reinterpret_cast<uint4*>(p.lane_sc + (c * 64 + lane) * 16)) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
      if (lane < NQ) {
        wsc[wid][lane] = lo;
        wsc[wid][lane + NQ] = hi;
        if (c < p.nchunks) {
          gmem(p.chunk_sc)[c * (2 * NQ) + lane] = lo;
          gmem(p.chunk_sc)[c * (2 * NQ) + lane + NQ] = hi;
        }
      }
    }
    lds_sync();  // every wave's totals are there; the image was cleared
    uint32_t unit_sc = 0;  // wave 0, lane q < 2^G: the unit's next-order bucket q
    if constexpr (SC) {
      if (uwid == 0 && lane < 2 * NQ) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) unit_sc += wsc[w][lane];
      }
    }

    // the super-chunk's totals and this wave's prefix, per last-level bucket
    uint32_t stot[NQ], spre[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const uint32_t v = lane < WAVES ? wt[lane][q] : 0u;
      static_assert(WAVES <= 16, "per-wave totals in one row of lanes");
      const uint32_t inc = row_sum16(v);  // (values in lanes < WAVES: row 0's prefix)
      stot[q] = __builtin_amdgcn_readlane(inc, WAVES - 1);
      spre[q] = __builtin_amdgcn_readlane(inc - v, uwid);
    }
    uint32_t cslot[2 * NQ - 2];
#pragma unroll
    for (int j = 1; j <= JC; ++j) {
      const int nb = 1 << j;
      uint32_t start = 0;
#pragma unroll
      for (int q = 0; q < nb; ++q) {
        uint32_t T = 0, E = 0, Wq = 0;
#pragma unroll
        for (int r = q; r < NQ; r += nb) {
          T += stot[r];
          E += spre[r] + ex[r];
          Wq += tot[r];  // this wave's symbols of the bucket (wave-uniform)
        }
        if (Wq) {  // a bucket the wave does not hold has no bits to place
          uint64_t mq = vm;
#pragma unroll
          for (int i = 0; i < j; ++i) mq &= ((q >> i) & 1) ? P0[i] : ~P0[i];
          or_bits(img[j - 1], start + E, bucket_bits(P0[j], mq, Wq, Ec, lut), __popcll(mq));
        }
        cslot[nb - 2 + q] = T;
        start += T;
      }
    }
    lds_sync();  // the image is complete

    const uint64_t s0 = sc * kUnitSyms;
    const uint32_t nws = uint32_t((min<uint64_t>(kUnitSyms, p.n - s0) + 63) >> 6);
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      if (threadIdx.x < nws)
        st_stream(p.temp_levels + j * p.temp_stride + sc * kUnitWords + threadIdx.x, brev64(img[j][threadIdx.x]));
      img[j][threadIdx.x] = 0;  // cleared before the next super-chunk's first barrier
    }
    if (uwid == 0) {
      uint32_t v = 0;
#pragma unroll
      for (int i = 0; i < 2 * NQ - 2; ++i) v = lane == i ? cslot[i] : v;
      if (lane < 2 * NQ - 2) gmem(p.counts)[sc * p.nslot + lane] = v;
      if constexpr (SC) {
        if (lane < 2 * NQ) gmem(p.counts)[sc * p.nslot + 2 * NQ - 2 + lane] = unit_sc;
      }
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (zsum[j]) atomicAdd(&red[j], (unsigned long long)zsum[j]);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) my_or |= __shfl_xor(my_or, d, 64);
  if (lane == 0 && my_or) atomicOr(&red[G], (unsigned long long)my_or);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < G; ++j)
      if (red[j]) atomicAdd(p.zeros_out + j, red[j]);
    if (red[G]) atomicOr(p.sym_or, red[G]);
  }
}

// ---- K1 with bucket-0 runs written in place ----------------------------------
// Bucket 0 of a sub-level (the symbols whose bits of the levels above are all
// zero) comes first in its level, so a unit's bucket-0 run starts at the sum
// of the earlier units' bucket-0 counts: none of the other buckets' totals is
// needed.  Every unit publishes its bucket-0 counts as one word with a
// "published" bit in each half.  The launch is cooperative (every workgroup
// resident) and workgroup b takes units b, b + G, b + 2G, ...; its running
// prefix goes from one of its units to the next by adding the counts of the
// G - 1 units in between (the gap), which the other workgroups publish at
// about the same time.  Waves 0..NSUB-1 keep one sub-level's prefix each:
// a unit's gap entries are loaded one step after it was built, right after
// the text prefetch, and summed at the next step's start, where the step
// waits for every load anyway; the unit's in-place words go out in that
// step, so every unit's image stays in LDS for three steps (three buffers).
// The words that lie wholly inside a unit's bucket-0 run go from the image
// straight to the level, shifted to the run's bit offset; the stitch builds
// every other word (the other buckets' runs and the words a run shares with
// its neighbours) from the image words written to scratch: word 0 and every
// word from the run's last two on.
struct DirectParams {
  unsigned long long* agg;    // [unit]: bucket-0 count of sub-level j in the low 31
                              // bits of dword j - 1, bit 31 of both dwords = published
  unsigned long long* fault;  // nonzero: a wait for a unit's counts gave up
  uint64_t* dst[kMaxGroup];   // final sub-level j at dst[j]
  uint32_t spin_limit;        // reloads of one entry before its wait gives up
  uint32_t publish;           // 0 (a test of the give-up path): no unit publishes
};

constexpr unsigned long long kAggFlag = 0x8000000080000000ull;  // in both dwords: a dword read alone is whole
constexpr uint32_t kGapEntries = 256;      // gap entries loaded a step ahead (grid <= 257)
constexpr uint32_t kSpinLimit = 1u << 16;  // reloads of one entry before its wait gives up (default)

__device__ __forceinline__ unsigned long long agg_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

template <int G>
__global__ __launch_bounds__(64 * kSuper) void k_local_direct(K1Params p, DirectParams d) {
  static_assert(G >= 2 && G <= 3, "two or three levels (three LDS images of each sub-level)");
  constexpr int WAVES = kSuper;
  constexpr int NSUB = G - 1;      // sub-levels 1..G-1
  constexpr int JC = G - 1;        // last sub-level
  constexpr int NQ = 1 << JC;      // its buckets
  constexpr int NF = NQ + 1;       // bucket counts + the last level's zeros
  constexpr int NS = (NF + 1) / 2; // packed scans
  constexpr uint32_t kFieldMask = 0x7fffffffu;
  __shared__ __align__(16) uint8_t lut[65536];
  __shared__ uint64_t img[3][NSUB][kSuperWords];
  __shared__ uint32_t wt[WAVES][NQ];
  __shared__ unsigned long long red[G + 1];  // block sums: zeros per level, symbol OR
  __shared__ uint32_t part[WAVES][NSUB];     // per wave: its share of a gap's bucket-0 counts
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int uwid = __builtin_amdgcn_readfirstlane(wid);
  const uint64_t nsuper = (p.nchunks + WAVES - 1) / WAVES;
  const uint64_t G_ = gridDim.x;
  if (threadIdx.x <= G) red[threadIdx.x] = 0;
  for (uint32_t i = threadIdx.x; i < 3 * NSUB * kSuperWords; i += 64 * WAVES) (&img[0][0][0])[i] = 0;
  load_lut<WAVES>(lut, p.lut);  // (ends with a barrier)

  uint32_t zsum[G];
#pragma unroll
  for (int j = 0; j < G; ++j) zsum[j] = 0;
  uint64_t my_or = 0;
  const bool fast_mode = (reinterpret_cast<uintptr_t>(p.text) & 15) == 0;
  auto full_chunk = [&](uint64_t c) { return fast_mode && c < p.nchunks && (c + 1) * kChunkSyms <= p.n; };
  auto load_chunk = [&](uint64_t c, uint4* dd) {
    const u32x4_t* src = reinterpret_cast<const u32x4_t*>(static_cast<const uint8_t*>(p.text) + c * kChunkSyms) +
                         4 * lane;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const u32x4_t v = ld_stream(src + i);
      dd[i] = make_uint4(v.x, v.y, v.z, v.w);
    }
  };
  // words wholly inside the bucket-0 runs [D[j], D[j] + cl[j]) of sub-levels
  // j + 1, from LDS image b, one per thread (to dstp[j]: d.dst[j + 1] held in
  // vector registers, set below)
  uint64_t dstp[NSUB];
  auto write_in_place = [&](int b, const uint64_t* D, const uint32_t* cl) {
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const uint64_t wf = (D[j] + 63) >> 6, wl = (D[j] + cl[j]) >> 6;
      const uint64_t w0 = D[j] >> 6;
      const uint32_t s = uint32_t(D[j] & 63);
      const uint64_t w = wf + threadIdx.x;  // (a run holds at most 1024 whole words)
      if (w < wl) {
        const uint32_t m = uint32_t(w - w0);  // 1 <= m when s > 0, m <= 1023
        // (no branch on s: with s = 0 the second word is the first again
        // and the shifts leave it as it is)
        const uint64_t v0 = img[b][j][m], v1 = img[b][j][m - (s ? 1u : 0u)];
        const uint64_t v = (v0 << s) | (v1 >> ((64 - s) & 63));
        st_stream(reinterpret_cast<uint64_t*>(dstp[j]) + w, brev64(v));
      }
    }
  };

  // Running prefix, the same in every wave: incl[j] = sub-level j + 1's
  // bucket-0 count of every unit up to the last one resolved; c0a / c0b =
  // the counts of the units built one and two steps ago.  The gap of unit u
  // is [u - G + 1, u) (from 0 for the first round); lane l < 16 NSUB of wave
  // w holds its entry 16 w + (l & 15) of sub-level (l >> 4) + 1 (G <= 257).
  uint64_t incl[NSUB];
  uint32_t c0a[NSUB], c0b[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; ++j) incl[j] = c0a[j] = c0b[j] = 0;
  // (values touched once a step live in vector registers: the scalar file
  // is full, and scalar spills cost vector instructions every step)
  auto in_vgpr = [](auto& x) { asm volatile("" : "+v"(x)); };
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    dstp[j] = reinterpret_cast<uint64_t>(d.dst[j + 1]);
    in_vgpr(dstp[j]);
  }
  // (unit indices fit 32 bits: the count array has one entry per unit)
  const uint32_t G32 = uint32_t(G_);
  const uint32_t gi = 16 * uint32_t(uwid) + uint32_t(lane & 15);
  const uint32_t gj = uint32_t(lane) >> 4;
  const bool glane = gj < uint32_t(NSUB);
  const uint32_t* const agg_l = reinterpret_cast<const uint32_t*>(d.agg) + 2 * gi + (glane ? gj : 0);
  uint32_t gvv = 0;  // the entry loaded a step ago
  bool gave_up = false;
  auto gap_lo = [&](uint32_t u) { return u >= G32 ? u - G32 + 1 : 0u; };
  auto issue_gap = [&](uint32_t u) {  // after the prefetch: waited for at the next step's start
    const uint32_t lo = gap_lo(u);
    const bool v = glane && lo + gi < u;  // (otherwise entry lo, which exists, is loaded and not used)
    gvv = __hip_atomic_load(agg_l + 2 * lo - (v ? 0 : 2 * gi), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  // this wave's share of unit u's gap (the loads have landed) into part[]
  auto share = [&](uint32_t u) {
    const uint32_t lo = gap_lo(u);
    const bool v = glane && lo + gi < u;
    uint32_t x = v ? gvv : 0x80000000u;
    if (!(x & 0x80000000u)) {  // not published yet: wait for it
      uint32_t spins = 0;
#pragma clang loop unroll(disable)
      while (!(x & 0x80000000u)) {
        if (gave_up || ++spins > d.spin_limit) {
          gave_up = true;
          atomicOr(d.fault, 1ull);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
        x = __hip_atomic_load(agg_l + 2 * lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    x &= kFieldMask;
    x = row_sum16(x);  // lanes 15 and 31: the two sub-levels' shares
    // (read out in uniform control flow: inside a branch taken by lane 0
    // alone, the compiler may finish the sum for the active lane only)
    uint32_t sh[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) sh[j] = __builtin_amdgcn_readlane(x, 16 * j + 15);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < NSUB; ++j) part[uwid][j] = sh[j];
    }
  };
  // after a barrier: the run starts of the unit whose shares are in part[]
  auto finish = [&](uint64_t* D, const uint32_t* cl) {
    // lane l < 16 NSUB reads wave (l & 15)'s share of sub-level (l >> 4) + 1
    // (each share is below 2^24: 16 entries of at most 2^16)
    const uint32_t x = row_sum16(glane ? part[lane & 15][gj] : 0u);
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      D[j] = incl[j] + __builtin_amdgcn_readlane(x, 16 * j + 15);
      incl[j] = D[j] + cl[j];
    }
  };

  uint4 pf[4];
  uint64_t cur = blockIdx.x;  // units blockIdx.x + k * gridDim.x (every workgroup resident)
  if (cur < nsuper && full_chunk(cur * WAVES + uwid)) load_chunk(cur * WAVES + uwid, pf);

  uint32_t step = 0;
  for (; cur < nsuper; ++step, cur += G_) {
    const int buf = int(step % 3), dbuf = int((step + 1) % 3);
    const uint64_t sc = cur;
    const uint64_t nxt = sc + G_;
    const uint64_t c = sc * WAVES + uwid;  // (wave-uniform: chunk addresses stay scalar)
    const uint64_t cn = nxt * WAVES + uwid;
    const bool have = c < p.nchunks;
    const uint32_t Ec = have ? uint32_t(min<uint64_t>(kChunkSyms, p.n - c * kChunkSyms)) : 0u;
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * lane));

    uint64_t P0[G];
#pragma unroll
    for (int k = 0; k < G; ++k) P0[k] = 0;
    const bool fc = full_chunk(c);
    if (fc) {
      planes_from_bytes<G>(p, pf, P0, my_or);
    } else if (cnt > 0) {
      load_planes_text<G, uint8_t>(p, p.text, c * kChunkSyms + 64ull * lane, cnt, P0, my_or);
    }
    // this wave's share of the gap of the unit built two steps ago (its
    // entries were loaded a step ago, after that step's prefetch: the wait
    // for them falls where the step waits for its stores before the next
    // prefetch anyway)
    if (step >= 2) share(uint32_t(sc - 2 * G_));
    if (nxt < nsuper && full_chunk(cn)) load_chunk(cn, pf);  // prefetch
    if (step >= 1) issue_gap(uint32_t(sc - G_));              // the gap of the unit built a step ago
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (lane < int(nw)) st_stream(p.out_level0 + c * kChunkWords + lane, brev64(P0[0]));

    // bucket masks of the last sub-level, M[q] = prod_i (bit i of q ? P_i : ~P_i)
    uint64_t M[NQ];
    M[0] = vm;
#pragma unroll
    for (int i = 0; i < JC; ++i) {
#pragma unroll
      for (int q = (1 << i) - 1; q >= 0; --q) {
        M[q + (1 << i)] = M[q] & P0[i];
        M[q] = M[q] & ~P0[i];
      }
    }
    uint32_t f[2 * NS];
#pragma unroll
    for (int q = 0; q < NQ; ++q) f[q] = __popcll(M[q]);
    f[NQ] = __popcll(~P0[G - 1] & vm);
    if (NF < 2 * NS) f[2 * NS - 1] = 0;
    uint32_t ex[2 * NS], tot[2 * NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const uint32_t v = f[2 * s2] | (f[2 * s2 + 1] << 16);
      const uint32_t inc2 = wave_incl_scan(v);
      const uint32_t t = __builtin_amdgcn_readlane(inc2, 63);
      const uint32_t e = inc2 - v;
      ex[2 * s2] = e & 0xffffu;
      ex[2 * s2 + 1] = e >> 16;
      tot[2 * s2] = t & 0xffffu;
      tot[2 * s2 + 1] = t >> 16;
    }
#pragma unroll
    for (int j = 0; j < JC; ++j) {
      uint32_t z = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (!((q >> j) & 1)) z += tot[q];
      zsum[j] += z;
    }
    zsum[G - 1] += tot[NQ];
    if (lane < NQ) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v = lane == q ? tot[q] : v;
      wt[wid][lane] = v;
    }
    lds_sync();  // every wave's totals and gap shares are there; image `buf` was cleared

    // the unit's totals and this wave's prefix, per last-level bucket
    uint32_t stot[NQ], spre[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const uint32_t v = lane < WAVES ? wt[lane][q] : 0u;
      static_assert(WAVES <= 16, "per-wave totals in one row of lanes");
      const uint32_t i2 = row_sum16(v);  // (values in lanes < WAVES: row 0's prefix)
      stot[q] = __builtin_amdgcn_readlane(i2, WAVES - 1);
      spre[q] = __builtin_amdgcn_readlane(i2 - v, uwid);
    }
    uint32_t c0[NSUB];  // the unit's bucket-0 count per sub-level
#pragma unroll
    for (int j = 1; j <= NSUB; ++j) {
      uint32_t T = 0;
#pragma unroll
      for (int r = 0; r < NQ; r += 1 << j) T += stot[r];
      c0[j - 1] = T;
    }
    if (uwid == 0 && lane == 0) {
      unsigned long long a = 0;
#pragma unroll
      for (int j = 0; j < NSUB; ++j) a |= (unsigned long long)c0[j] << (32 * j);
      if (d.publish) __hip_atomic_store(d.agg + sc, a | kAggFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the run starts of the unit built two steps ago
    uint64_t D[NSUB];
    uint32_t cl[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) cl[j] = c0b[j];
    if (step >= 2) finish(D, cl);
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      in_vgpr(D[j]);
      in_vgpr(cl[j]);
    }
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      c0b[j] = c0a[j];
      c0a[j] = c0[j];
      in_vgpr(c0a[j]);
      in_vgpr(c0b[j]);
      in_vgpr(incl[j]);
    }
#pragma unroll
    for (int j = 0; j < G; ++j) in_vgpr(zsum[j]);

    uint32_t cslot[2 * NQ - 2];
#pragma unroll
    for (int j = 1; j <= JC; ++j) {
      const int nb = 1 << j;
      uint32_t start = 0;
#pragma unroll
      for (int q = 0; q < nb; ++q) {
        uint32_t T = 0, E = 0, Wq = 0;
#pragma unroll
        for (int r = q; r < NQ; r += nb) {
          T += stot[r];
          E += spre[r] + ex[r];
          Wq += tot[r];  // this wave's symbols of the bucket (wave-uniform)
        }
        if (Wq) {
          uint64_t mq = vm;
#pragma unroll
          for (int i = 0; i < j; ++i) mq &= ((q >> i) & 1) ? P0[i] : ~P0[i];
          or_bits(img[buf][j - 1], start + E, bucket_bits(P0[j], mq, Wq, Ec, lut), __popcll(mq));
        }
        cslot[nb - 2 + q] = T;
        start += T;
      }
    }
    // the unit built two steps ago goes out
    if (step >= 2) write_in_place(dbuf, D, cl);
    lds_sync();  // image `buf` is complete; image `dbuf` and part[] were read

    // scratch: word 0 and the words from the bucket-0 run's last two on
    const uint64_t s0 = sc * kSuperSyms;
    const uint32_t nws = uint32_t((min<uint64_t>(kSuperSyms, p.n - s0) + 63) >> 6);
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const uint32_t lo = (c0[j] >> 6) ? (c0[j] >> 6) - 1 : 0;
      if (threadIdx.x < nws && (threadIdx.x == 0 || threadIdx.x >= lo))
        st_stream(p.temp_levels + j * p.temp_stride + sc * kSuperWords + threadIdx.x,
                  brev64(img[buf][j][threadIdx.x]));
      img[dbuf][j][threadIdx.x] = 0;  // written in place above; cleared before the next step's first barrier
    }
    if (uwid == 0) {
      uint32_t v = 0;
#pragma unroll
      for (int i = 0; i < 2 * NQ - 2; ++i) v = lane == i ? cslot[i] : v;
      if (lane < 2 * NQ - 2) gmem(p.counts)[sc * p.nslot + lane] = v;
    }
  }
  // the last two units built: the one two steps back (its gap loaded in the
  // last step), then the last one
  if (step >= 2) {
    __builtin_amdgcn_s_waitcnt(0x0f70);
    share(uint32_t(cur - 2 * G_));
    lds_sync();
    uint64_t D[NSUB];
    finish(D, c0b);
    write_in_place(int((step + 1) % 3), D, c0b);
  }
  if (step >= 1) {
    const uint32_t u = uint32_t(cur - G_);
    issue_gap(u);
    __builtin_amdgcn_s_waitcnt(0x0f70);
    lds_sync();  // part[] was read
    share(u);
    lds_sync();
    uint64_t D[NSUB];
    finish(D, c0a);
    write_in_place(int((step + 2) % 3), D, c0a);
  }

  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (zsum[j]) atomicAdd(&red[j], (unsigned long long)zsum[j]);
  }
#pragma unroll
  for (int dd = 32; dd >= 1; dd >>= 1) my_or |= __shfl_xor(my_or, dd, 64);
  if (lane == 0 && my_or) atomicOr(&red[G], (unsigned long long)my_or);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < G; ++j)
      if (red[j]) atomicAdd(p.zeros_out + j, red[j]);
    if (red[G]) atomicOr(p.sym_or, red[G]);
  }
}

// ---- K1 over super-chunk units of many texts ---------------------------------
// Batches of long texts: text t owns units [base[t], base[t+1]) of up to 16
// chunks (the single-text super-chunk build per unit).  Per-unit metadata
// (first symbol, symbols left in the text, level-0 word, text) comes from a
// table, so that a unit needs one round of independent loads, issued a unit
// ahead.  Workgroup b takes units b, b + G, b + 2G, ..., the next one's
// text prefetched during a unit; zero counts are reduced per unit from the
// waves' bucket totals (one atomic per level and unit).
struct UnitMeta {
  uint64_t text;  // address of the unit's first symbol
  uint64_t rem;   // symbols of the text from the unit's first on
  uint64_t lvl0;  // word of the unit's first level-0 bit in the output block
  uint64_t t;     // text
};

// One wave per text (lanes over its units): the metadata of its units, and
// unit -> text.
__global__ __launch_bounds__(256) void k_unit_meta(const uint64_t* base, const uint64_t* text, const uint64_t* n,
                                                   const uint64_t* out0, uint32_t ntexts, uint32_t width,
                                                   uint64_t unit_syms, UnitMeta* meta, uint32_t* unit_text) {
  const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= ntexts) return;
  for (uint64_t u = base[t] + (threadIdx.x & 63); u < base[t + 1]; u += 64) {
    const uint64_t lu = u - base[t];
    UnitMeta m;
    m.text = text[t] + lu * unit_syms * width;
    m.rem = n[t] - lu * unit_syms;
    m.lvl0 = out0[t] + lu * (unit_syms / 64);
    m.t = t;
    meta[u] = m;
    unit_text[u] = t;
  }
}

template <int G, bool WIDE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_local_super_batch(K1Params p, const UnitMeta* meta) {
  static_assert(G >= 2 && G <= 4, "two to four levels");
  constexpr uint64_t kUnitSyms = uint64_t(WAVES) * kChunkSyms;
  constexpr uint32_t kUnitWords = WAVES * kChunkWords;
  constexpr int NSUB = G - 1;
  constexpr int JC = G - 1;
  constexpr int NQ = 1 << JC;
  constexpr int NF = NQ + 1;       // bucket counts + the last level's zeros
  constexpr int NS = (NF + 1) / 2;
  __shared__ __align__(16) uint8_t lut[65536];
  __shared__ uint64_t img[NSUB][kUnitWords];
  __shared__ uint32_t wt[WAVES][NF];
  __shared__ unsigned long long red_or;
  if (threadIdx.x == 0) red_or = 0;
  for (uint32_t i = threadIdx.x; i < NSUB * kUnitWords; i += 64 * WAVES) (&img[0][0])[i] = 0;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int uwid = __builtin_amdgcn_readfirstlane(wid);
  load_lut<WAVES>(lut, p.lut);  // (ends with a barrier)

  uint64_t my_or = 0;
  const uint64_t U = p.nchunks;  // units
  auto get = [&](uint64_t u) {
    UnitMeta m{0, 0, 0, 0};
    if (u < U) m = gmem(meta)[u];
    return m;
  };
  // this wave's chunk of a unit: symbols and whether it is a whole aligned chunk
  auto chunk_syms = [&](const UnitMeta& m) -> uint32_t {
    const uint64_t c0 = uint64_t(wid) * kChunkSyms;
    return m.rem > c0 ? uint32_t(min<uint64_t>(kChunkSyms, m.rem - c0)) : 0u;
  };
  auto full = [&](const UnitMeta& m) {
    return !WIDE && (m.text & 15) == 0 && m.rem >= uint64_t(wid + 1) * kChunkSyms;
  };
  auto load_chunk = [&](const UnitMeta& m, uint4* d) {
    const auto src = gmem(reinterpret_cast<const uint4*>(m.text + uint64_t(wid) * kChunkSyms + 64ull * lane));
#pragma unroll
    for (int i = 0; i < 4; ++i) d[i] = src[i];
  };
  uint4 pf[4];
  // units blockIdx.x + k gridDim.x: every workgroup runs about the same
  // number of units (their sizes differ little on average), and no queue
  // atomic is waited for at every step
  const uint64_t G_ = gridDim.x;
  uint64_t u = blockIdx.x, un = u + G_;
  UnitMeta cur = get(u), nxt = get(un);
  if (u < U && full(cur)) load_chunk(cur, pf);

  while (u < U) {
    const uint32_t Ec = chunk_syms(cur);
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * lane));

    uint64_t P0[G];
#pragma unroll
    for (int k = 0; k < G; ++k) P0[k] = 0;
    const bool pf_next = un < U && full(nxt);
    if (full(cur)) {
      planes_from_bytes<G>(p, pf, P0, my_or);
      if (pf_next) load_chunk(nxt, pf);  // prefetch
    } else {
      if (cnt > 0) {
        const uint64_t e0 = uint64_t(wid) * kChunkSyms + 64ull * lane;
        const void* tx = reinterpret_cast<const void*>(cur.text);
        if constexpr (!WIDE) {
          load_planes_text<G, uint8_t>(p, tx, e0, cnt, P0, my_or);
        } else {
          switch (p.width) {
            case 2: load_planes_text<G, uint16_t>(p, tx, e0, cnt, P0, my_or); break;
            case 4: load_planes_text<G, uint32_t>(p, tx, e0, cnt, P0, my_or); break;
            default: load_planes_text<G, uint64_t>(p, tx, e0, cnt, P0, my_or); break;
          }
        }
      }
      if (pf_next) load_chunk(nxt, pf);
    }
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (lane < int(nw)) st_stream(p.out_level0 + cur.lvl0 + uint64_t(wid) * kChunkWords + lane, brev64(P0[0]));

    uint64_t M[NQ];
    M[0] = vm;
#pragma unroll
    for (int i = 0; i < JC; ++i) {
#pragma unroll
      for (int q = (1 << i) - 1; q >= 0; --q) {
        M[q + (1 << i)] = M[q] & P0[i];
        M[q] = M[q] & ~P0[i];
      }
    }
    uint32_t f[2 * NS];
#pragma unroll
    for (int q = 0; q < NQ; ++q) f[q] = __popcll(M[q]);
    f[NQ] = __popcll(~P0[G - 1] & vm);
    if (NF < 2 * NS) f[2 * NS - 1] = 0;
    uint32_t ex[2 * NS], tot[2 * NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const uint32_t v = f[2 * s2] | (f[2 * s2 + 1] << 16);
      const uint32_t inc = wave_incl_scan(v);
      const uint32_t tt = __builtin_amdgcn_readlane(inc, 63);
      const uint32_t e = inc - v;
      ex[2 * s2] = e & 0xffffu;
      ex[2 * s2 + 1] = e >> 16;
      tot[2 * s2] = tt & 0xffffu;
      tot[2 * s2 + 1] = tt >> 16;
    }
    if (lane < NF) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < NF; ++q) v = lane == q ? tot[q] : v;
      wt[wid][lane] = v;
    }
    lds_sync();  // every wave's totals are there; the image was cleared

    // the unit after the next: its metadata loads overlap the rest of this unit
    const uint64_t unn = un + G_;
    const UnitMeta nn = get(unn);

    uint32_t stot[NF], spre[NQ];
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      const uint32_t v = lane < WAVES ? wt[lane][q] : 0u;
      static_assert(WAVES <= 16, "per-wave totals in one row of lanes");
      const uint32_t inc = row_sum16(v);  // (values in lanes < WAVES: row 0's prefix)
      stot[q] = __builtin_amdgcn_readlane(inc, WAVES - 1);
      if (q < NQ) spre[q] = __builtin_amdgcn_readlane(inc - v, uwid);
    }
    if (threadIdx.x == 0) {
      // the unit's zeros per level: level j < G-1 has bit j == 0 in the
      // last sub-level's buckets q with bit j clear
      unsigned long long* zd = p.b_zeros + cur.t * kMaxGroup;
#pragma unroll
      for (int j = 0; j < JC; ++j) {
        uint32_t z = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (!((q >> j) & 1)) z += stot[q];
        if (z) atomicAdd(zd + j, (unsigned long long)z);
      }
      if (stot[NQ]) atomicAdd(zd + (G - 1), (unsigned long long)stot[NQ]);
    }
    uint32_t cslot[2 * NQ - 2];
#pragma unroll
    for (int j = 1; j <= JC; ++j) {
      const int nb = 1 << j;
      uint32_t start = 0;
#pragma unroll
      for (int q = 0; q < nb; ++q) {
        uint32_t T = 0, E = 0, Wq = 0;
#pragma unroll
        for (int r = q; r < NQ; r += nb) {
          T += stot[r];
          E += spre[r] + ex[r];
          Wq += tot[r];  // this wave's symbols of the bucket (wave-uniform)
        }
        if (Wq) {  // a bucket the wave does not hold has no bits to place
          uint64_t mq = vm;
#pragma unroll
          for (int i = 0; i < j; ++i) mq &= ((q >> i) & 1) ? P0[i] : ~P0[i];
          or_bits(img[j - 1], start + E, bucket_bits(P0[j], mq, Wq, Ec, lut), __popcll(mq));
        }
        cslot[nb - 2 + q] = T;
        start += T;
      }
    }
    lds_sync();  // the image is complete

    const uint32_t nws = uint32_t((min<uint64_t>(kUnitSyms, cur.rem) + 63) >> 6);
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      if (threadIdx.x < nws)
        st_stream(p.temp_levels + j * p.temp_stride + u * kUnitWords + threadIdx.x, brev64(img[j][threadIdx.x]));
      img[j][threadIdx.x] = 0;  // cleared before the next unit's first barrier
    }
    if (uwid == 0) {
      uint32_t v = 0;
#pragma unroll
      for (int i = 0; i < 2 * NQ - 2; ++i) v = lane == i ? cslot[i] : v;
      if (lane < 2 * NQ - 2) gmem(p.counts)[u * p.nslot + lane] = v;
    }
    u = un;
    cur = nxt;
    un = unn;
    nxt = nn;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) my_or |= __shfl_xor(my_or, d, 64);
  if (lane == 0 && my_or) atomicOr(&red_or, (unsigned long long)my_or);
  __syncthreads();
  if (threadIdx.x == 0 && red_or) atomicOr(p.sym_or, red_or);
}

// ---- Huffman-shaped local build over super-chunks -------------------------
// Level l of a Huffman-shaped structure holds the occurrences whose code is
// longer than l, grouped by node (their l-bit code prefix), text order
// inside a node.  With every symbol mapped to its code left-aligned in G
// bits, a node of level j is a bucket of the plain construction's sub-level
// j (bit i of the bucket index = code bit i), and whether a bucket is a node
// depends on the prefix alone (codes are prefix-free), so the plain
// super-chunk assembly applies with the buckets that are no node left out.
// Sub-level buckets are laid out in slot order (the stitch's order).
template <int G>
__global__ __launch_bounds__(64 * kSuper) void k_huff_super(HuffSuperParams hp, const uint8_t* lut_src) {
  static_assert(G >= 2 && G <= 4, "two to four levels");
  constexpr int WAVES = kSuper;
  constexpr int NSUB = G - 1;
  constexpr int JC = G - 1;
  constexpr int NQ = 1 << JC;
  constexpr int NS = (NQ + 1) / 2;
  __shared__ __align__(16) uint8_t lut[65536];
  __shared__ uint64_t img[NSUB][kSuperWords];
  __shared__ uint32_t wt[WAVES][NQ];
  __shared__ uint8_t cmap[256];
  for (uint32_t i = threadIdx.x; i < NSUB * kSuperWords; i += 64 * WAVES) (&img[0][0])[i] = 0;
  if (threadIdx.x < 256) cmap[threadIdx.x] = hp.code[threadIdx.x];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int uwid = __builtin_amdgcn_readfirstlane(wid);
  load_lut<WAVES>(lut, lut_src);  // (ends with a barrier)
  K1Params kp{};
  kp.L = G;
  kp.l0 = 0;
  kp.np = G;
  // codes of symbols 0..7 as two permute tables
  const uint32_t tlo = uint32_t(cmap[0]) | (uint32_t(cmap[1]) << 8) | (uint32_t(cmap[2]) << 16) |
                       (uint32_t(cmap[3]) << 24);
  const uint32_t thi = uint32_t(cmap[4]) | (uint32_t(cmap[5]) << 8) | (uint32_t(cmap[6]) << 16) |
                       (uint32_t(cmap[7]) << 24);
  auto map4 = [&](uint32_t x) -> uint32_t {
    return uint32_t(cmap[x & 0xff]) | (uint32_t(cmap[(x >> 8) & 0xff]) << 8) |
           (uint32_t(cmap[(x >> 16) & 0xff]) << 16) | (uint32_t(cmap[x >> 24]) << 24);
  };
  const uint64_t nsuper = (hp.nchunks + WAVES - 1) / WAVES;
  const bool aligned = (reinterpret_cast<uintptr_t>(hp.text) & 15) == 0;
  auto full_chunk = [&](uint64_t c) { return aligned && c < hp.nchunks && (c + 1) * kChunkSyms <= hp.n; };
  auto load_chunk = [&](uint64_t c, uint4* d) {
    const u32x4_t* src = reinterpret_cast<const u32x4_t*>(hp.text + c * kChunkSyms + 64ull * lane);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const u32x4_t v = ld_stream(src + i);
      d[i] = make_uint4(v.x, v.y, v.z, v.w);
    }
  };
  uint4 pf[4];
  uint64_t sc = blockIdx.x;
  if (sc < nsuper && full_chunk(sc * WAVES + wid)) load_chunk(sc * WAVES + wid, pf);

  for (; sc < nsuper; sc += gridDim.x) {
    const uint64_t c = sc * WAVES + wid;
    const uint64_t cn = (sc + gridDim.x) * WAVES + wid;
    const bool have = c < hp.nchunks;
    const uint32_t Ec = have ? uint32_t(min<uint64_t>(kChunkSyms, hp.n - c * kChunkSyms)) : 0u;
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * lane));

    uint64_t P0[G];
#pragma unroll
    for (int k = 0; k < G; ++k) P0[k] = 0;
    if (full_chunk(c)) {
      uint4 m[4];
      if (hp.small_alpha) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          m[i] = make_uint4(__builtin_amdgcn_perm(thi, tlo, pf[i].x), __builtin_amdgcn_perm(thi, tlo, pf[i].y),
                            __builtin_amdgcn_perm(thi, tlo, pf[i].z), __builtin_amdgcn_perm(thi, tlo, pf[i].w));
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) m[i] = make_uint4(map4(pf[i].x), map4(pf[i].y), map4(pf[i].z), map4(pf[i].w));
      }
      uint64_t unused = 0;
      planes_from_bytes<G>(kp, m, P0, unused);
      if (full_chunk(cn)) load_chunk(cn, pf);  // prefetch
    } else {
      const auto tx = gmem(hp.text);
      for (int i = 0; i < cnt; ++i) {
        const uint32_t a = cmap[tx[c * kChunkSyms + 64ull * lane + i]];
#pragma unroll
        for (int k = 0; k < G; ++k) P0[k] |= uint64_t((a >> (G - 1 - k)) & 1u) << i;
      }
      if (full_chunk(cn)) load_chunk(cn, pf);
    }
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (lane < int(nw)) st_stream(hp.level0 + c * kChunkWords + lane, brev64(P0[0]));

    uint64_t M[NQ];
    M[0] = vm;
#pragma unroll
    for (int i = 0; i < JC; ++i) {
#pragma unroll
      for (int q = (1 << i) - 1; q >= 0; --q) {
        M[q + (1 << i)] = M[q] & P0[i];
        M[q] = M[q] & ~P0[i];
      }
    }
    uint32_t f[2 * NS];
#pragma unroll
    for (int q = 0; q < NQ; ++q) f[q] = __popcll(M[q]);
    if (NQ < 2 * NS) f[2 * NS - 1] = 0;
    uint32_t ex[2 * NS], tot[2 * NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const uint32_t v = f[2 * s2] | (f[2 * s2 + 1] << 16);
      const uint32_t inc = wave_incl_scan(v);
      const uint32_t t = __builtin_amdgcn_readlane(inc, 63);
      const uint32_t e = inc - v;
      ex[2 * s2] = e & 0xffffu;
      ex[2 * s2 + 1] = e >> 16;
      tot[2 * s2] = t & 0xffffu;
      tot[2 * s2 + 1] = t >> 16;
    }
    if (lane < NQ) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v = lane == q ? tot[q] : v;
      wt[wid][lane] = v;
    }
    lds_sync();

    uint32_t stot[NQ], spre[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const uint32_t v = lane < WAVES ? wt[lane][q] : 0u;
      static_assert(WAVES <= 16, "per-wave totals in one row of lanes");
      const uint32_t inc = row_sum16(v);  // (values in lanes < WAVES: row 0's prefix)
      stot[q] = __builtin_amdgcn_readlane(inc, WAVES - 1);
      spre[q] = __builtin_amdgcn_readlane(inc - v, uwid);
    }
    uint32_t used[NSUB];  // the super-chunk's bits of each sub-level (its nodes)
    uint32_t crow = 0;    // wave 0, lane s: the super-chunk's count of slot s
#pragma unroll
    for (int j = 1; j <= JC; ++j) {
      const int nb = 1 << j;
      uint32_t T[1 << JC];
      used[j - 1] = 0;
#pragma unroll
      for (int q = 0; q < nb; ++q) {
        T[q] = 0;
#pragma unroll
        for (int r = q; r < NQ; r += nb) T[q] += stot[r];
      }
#pragma unroll
      for (int q = 0; q < nb; ++q) {
        const uint32_t rk = hp.rank[j][q];
        if (rk == 0xffu) continue;  // no node
        used[j - 1] += T[q];
        uint32_t start = 0, E = 0, Wq = 0;
#pragma unroll
        for (int q2 = 0; q2 < nb; ++q2) {
          const uint32_t r2 = hp.rank[j][q2];
          if (r2 != 0xffu && r2 < rk) start += T[q2];
        }
#pragma unroll
        for (int r = q; r < NQ; r += nb) {
          E += spre[r] + ex[r];
          Wq += tot[r];  // this wave's symbols of the node (wave-uniform)
        }
        if (Wq) {
          uint64_t mq = vm;
#pragma unroll
          for (int i = 0; i < j; ++i) mq &= ((q >> i) & 1) ? P0[i] : ~P0[i];
          or_bits(img[j - 1], start + E, bucket_bits(P0[j], mq, Wq, Ec, lut), __popcll(mq));
        }
        crow = lane == int(hp.slot0[j] + rk) ? T[q] : crow;
      }
    }
    lds_sync();

    // every slot is a node of one sub-level, each written above
    if (uwid == 0 && uint32_t(lane) < hp.nslot) gmem(hp.counts)[sc * hp.nslot + lane] = crow;
    // only the words that hold node bits are written (the stitch reads no
    // bit outside a node's run)
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      if (threadIdx.x < (used[j] + 63) >> 6)
        st_stream(hp.temp + j * hp.temp_stride + sc * kSuperWords + threadIdx.x, brev64(img[j][threadIdx.x]));
      img[j][threadIdx.x] = 0;
    }
  }
}

// ---- symbol reorder between passes ------------------------------------------
// After a pass over the top G levels of its symbols, the next pass reads the
// symbols in the order of the level below: stable by bucket q = (bit of the
// pass's last level ... bit of its first level), q ascending.  Only the bits
// below the pass are kept, in the narrowest type that holds them.  One wave
// per chunk, 64 symbols per step (one per lane, read coalesced): the step's
// bucket masks come from G ballots; a lane's destination is its bucket's
// running position (held by lane q for bucket q) plus its rank among the
// step's lanes of that bucket.
struct ScatterParams {
  const void* in;
  void* out;
  uint64_t n;
  uint32_t sh;               // the G-bit field of the pass's levels starts at bit sh
  const uint64_t* dstpos;    // [c * 2^G + q]: global position of (chunk, bucket)
  const uint32_t* counts;    // [c * nslot + slot0 + q]: the chunk's bucket sizes
  uint32_t nslot, slot0;
  uint64_t nchunks;
  const uint8_t* lane_sc;    // or nullptr: each lane's field-value counts, as the
                             // pass's local build wrote them (K1Params::lane_sc)
};

template <typename TO>
constexpr int scatter_waves() {
  return 4;
}

// Rank of this lane among the lanes whose bit is set in m (m's bits below
// the lane): v_mbcnt.
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

template <typename TI, typename TO, int G>
__global__ __launch_bounds__(256) void k_scatter(ScatterParams p) {
  constexpr int NQ = 1 << G;
  __shared__ uint64_t runs[4][NQ];  // per wave: running position of every bucket
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint64_t c = uint64_t(blockIdx.x) * 4 + wid;
  if (c >= p.nchunks) return;  // whole waves; no barrier below
  uint64_t* run = runs[wid];
  const auto in = gmem(static_cast<const TI*>(p.in) + c * kChunkSyms);
  const auto out = gmem(static_cast<TO*>(p.out));
  const uint32_t E = uint32_t(min<uint64_t>(kChunkSyms, p.n - c * kChunkSyms));
  const uint64_t low = p.sh >= 64 ? ~0ull : ((1ull << p.sh) - 1);  // bits kept
  if (lane < NQ) run[lane] = p.dstpos[c * NQ + lane];
  wave_sync();
  // 16 steps of 64 symbols per round trip: their loads are issued together
  constexpr int U = 16;
  for (uint32_t t0 = 0; t0 < E; t0 += 64 * U) {
    TI x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = t0 + 64 * u + lane;
      x[u] = i < E ? in[i] : TI(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool valid = t0 + 64 * u + lane < E;
      const uint32_t f = uint32_t(uint64_t(x[u]) >> p.sh) & uint32_t(NQ - 1);
      const uint32_t q = __builtin_bitreverse32(f) >> (32 - G);
      // lanes of this lane's bucket
      uint64_t mine = __ballot(valid);
#pragma unroll
      for (int b = 0; b < G; ++b) {
        const uint64_t B = __ballot(valid && ((q >> b) & 1));
        mine &= ((q >> b) & 1) ? B : ~B;
      }
      const uint32_t rank = mbcnt64(mine);
      // one wave's LDS accesses execute in program order, and the compiler
      // keeps these (possibly aliasing) ones in order: every lane reads its
      // bucket's position before the bucket's last lane advances it, and
      // the next step reads the advanced positions
      const uint64_t base = run[q];
      if (valid) {
        out[base + rank] = TO(uint64_t(x[u]) & low);
        if (rank + 1 == uint32_t(__popcll(mine))) run[q] = base + rank + 1;  // the bucket's last lane
      }
    }
  }
}

// The reorder as a counting sort per chunk: lane i owns the chunk's symbols
// [64 i, 64 i + 64) (held in registers), counts their buckets with LDS adds
// into per-lane counters, one wave scan per bucket turns the counters into
// positions in the chunk's bucket order (stable: lanes in order, symbols in
// order within a lane), and returning LDS adds hand every symbol its
// position in an LDS copy of the chunk, which is then copied out run by run.
// About a dozen instructions per 64 symbols, against some 75 for the
// per-step ranking of k_scatter (measured on the 2^28-symbol Zipf matrix:
// 3.89 ms for the whole build, against 4.74 ms).  Inputs and outputs of 1 to
// 4 bytes; 8-byte symbols use k_scatter.
template <typename TO>
constexpr int scatter_seg_waves() {
  return sizeof(TO) == 1 ? 4 : 2;
}

template <typename TI, typename TO, int G>
__global__ __launch_bounds__(64 * scatter_seg_waves<TO>()) void k_scatter_seg(ScatterParams p) {
  constexpr int NQ = 1 << G;
  constexpr int SW = scatter_seg_waves<TO>();
  constexpr int V = 64 * int(sizeof(TI)) / 16;  // 16-byte loads per lane
  static_assert(sizeof(TI) <= 4 && sizeof(TO) <= 4, "symbols of up to 4 bytes");
  __shared__ TO stage[SW][kChunkSyms];
  __shared__ uint32_t ctr[SW][NQ][64];  // per wave, bucket and lane: count, then next position
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint64_t c = uint64_t(blockIdx.x) * SW + wid;
  if (c >= p.nchunks) return;  // whole waves; no barrier below
  TO* st = stage[wid];
  const TI* src = static_cast<const TI*>(p.in) + c * kChunkSyms + 64ull * lane;
  const auto out = gmem(static_cast<TO*>(p.out));
  const uint32_t E = uint32_t(min<uint64_t>(kChunkSyms, p.n - c * kChunkSyms));
  const int m = max(0, min(64, int(E) - 64 * lane));  // this lane's symbols
  // this lane's field-value counts when the local build left them
  uint4 lsc = make_uint4(0u, 0u, 0u, 0u);
  if (p.lane_sc) lsc = *gmem(reinterpret_cast<const uint4*>(p.lane_sc + (c * 64 + lane) * 16));
  // the lane's symbols: 16-byte loads when the whole segment is there and aligned
  // (kept as dwords, PD symbols each, the first in the low bits)
  constexpr int PD = 4 / int(sizeof(TI));  // symbols per dword
  constexpr uint32_t TB = 8 * sizeof(TI);  // bits per symbol
  constexpr int ND = 64 / PD;              // dwords per lane
  uint32_t d[ND];
  if (m == 64 && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    const auto g = gmem(reinterpret_cast<const uint4*>(src));
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const uint4 w = g[i];
      d[4 * i] = w.x;
      d[4 * i + 1] = w.y;
      d[4 * i + 2] = w.z;
      d[4 * i + 3] = w.w;
    }
  } else {
    const auto g = gmem(src);
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      uint32_t x = 0;
#pragma unroll
      for (int e = 0; e < PD; ++e)
        if (PD * j + e < m) x |= uint32_t(g[PD * j + e]) << (e * TB);
      d[j] = x;
    }
  }
  // Counters are kept per field value f (the G bits at sh); bucket q (the
  // chunk's order) is the bit reversal of f, applied only when the counters
  // are scanned.  A symbol's field and its kept low bits are one bit-field
  // extract each from the dword that holds it (sh + G <= the symbol's bits;
  // an extract of width 0 is 0).
  const uint32_t sh = p.sh;  // < 32
  auto field = [&](int i) -> uint32_t { return __builtin_amdgcn_ubfe(d[i / PD], (i % PD) * TB + sh, G); };
  auto kept = [&](int i) -> TO { return TO(__builtin_amdgcn_ubfe(d[i / PD], (i % PD) * TB, sh)); };
  auto lds_add = [](uint32_t* a) {
    return __hip_atomic_fetch_add(a, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  uint32_t* cl = &ctr[wid][0][lane];  // this lane's counter of field value f at cl[64 f]
  // whole chunks (wave-uniform) without a test per symbol
  const bool whole = E == kChunkSyms;
  if (p.lane_sc) {  // counted by the pass's local build
    const uint32_t lw[4] = {lsc.x, lsc.y, lsc.z, lsc.w};
#pragma unroll
    for (int q = 0; q < NQ; ++q) cl[64 * q] = (lw[q >> 2] >> (8 * (q & 3))) & 0xffu;
  } else {
#pragma unroll
    for (int q = 0; q < NQ; ++q) cl[64 * q] = 0;
    if (whole) {
#pragma unroll
      for (int i = 0; i < 64; ++i) lds_add(cl + 64 * field(i));
    } else {
#pragma unroll
      for (int i = 0; i < 64; ++i)
        if (i < m) lds_add(cl + 64 * field(i));
    }
  }
  wave_sync();
  // counters -> positions: bucket start in the chunk + count in lanes below
  uint32_t run = 0, lst = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int f = int(__builtin_bitreverse32(uint32_t(q)) >> (32 - G));  // (a constant)
    const uint32_t v = cl[64 * f];
    const uint32_t inc = wave_incl_scan(v);
    cl[64 * f] = run + inc - v;
    const uint32_t t = __builtin_amdgcn_readlane(inc, 63);
    lst = lane == q ? run : lst;
    tot = lane == q ? t : tot;
    run += t;
  }
  wave_sync();
  if (whole) {
    // eight positions taken before their symbols are stored, so that the
    // returning adds overlap (a lane's adds to one counter still return
    // in order: one wave's LDS operations execute in program order)
#pragma unroll
    for (int i0 = 0; i0 < 64; i0 += 8) {
      uint32_t pos[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) pos[k] = lds_add(cl + 64 * field(i0 + k));
#pragma unroll
      for (int k = 0; k < 8; ++k) st[pos[k]] = kept(i0 + k);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 64; ++i)
      if (i < m) st[lds_add(cl + 64 * field(i))] = kept(i);
  }
  wave_sync();  // the staged chunk is complete
  const uint64_t gpos = lane < NQ ? p.dstpos[c * NQ + lane] : 0ull;
  // (the bucket sizes the positions were made from bound every run)
  const uint32_t kcnt = lane < NQ ? p.counts[c * p.nslot + p.slot0 + lane] : 0u;
  tot = min(tot, kcnt);
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    const uint32_t len = __builtin_amdgcn_readlane(tot, q);
    const uint32_t s0 = __builtin_amdgcn_readlane(lst, q);
    const uint64_t dst = (uint64_t(__builtin_amdgcn_readlane(uint32_t(gpos >> 32), q)) << 32) |
                         __builtin_amdgcn_readlane(uint32_t(gpos), q);
    // a uniform base and a 32-bit lane offset: scalar-base stores, LDS reads
    // at immediate offsets; whole rounds of 64 without a test per lane
    const auto o = out + dst;
    const TO* s = st + s0;
    const uint32_t nfull = len >> 6;
    uint32_t i = uint32_t(lane);
#pragma unroll 4
    for (uint32_t k = 0; k < nfull; ++k, i += 64) o[i] = s[i];
    if (i < len) o[i] = s[i];
  }
}

// ---- K2 --------------------------------------------------------------------
// Parallel exclusive scan over chunks, in tiles of 1024 chunks:
// k_tile_sums (grid tiles x slots) -> k_tile_scan (grid slots) ->
// k_tile_apply (grid tiles x slots).
constexpr uint32_t kTile = 1024;

template <bool IND>
__global__ __launch_bounds__(256) void k_tile_sums(const uint32_t* counts, uint64_t* tsum,
                                                   uint64_t nchunks, uint32_t nslot, uint64_t ntiles) {
  __shared__ uint64_t ws[4];
  const uint64_t tile = blockIdx.x;
  const uint32_t slot = blockIdx.y;
  uint64_t v = 0;
  for (uint32_t k = threadIdx.x; k < kTile; k += 256) {
    const uint64_t c = tile * kTile + k;
    if (c < nchunks) v += IND ? (counts[c * nslot + slot] != 0) : counts[c * nslot + slot];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) tsum[uint64_t(slot) * ntiles + tile] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(1024) void k_tile_scan(uint64_t* tsum, uint64_t* totals, uint64_t ntiles) {
  __shared__ uint64_t wsum[16];
  const uint32_t slot = blockIdx.x;
  uint64_t* a = tsum + uint64_t(slot) * ntiles;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  uint64_t carry = 0;
  for (uint64_t b = 0; b < ntiles; b += 1024) {
    const uint64_t i = b + t;
    const uint64_t v = i < ntiles ? a[i] : 0;
    uint64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t x = __shfl_up(inc, d, 64);
      if (lane >= d) inc += x;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      if (w < wid) off += wsum[w];
      tot += wsum[w];
    }
    if (i < ntiles) a[i] = carry + off + inc - v;
    carry += tot;
    __syncthreads();
  }
  if (t == 0) totals[slot] = carry;
}

template <bool IND>
__global__ __launch_bounds__(1024) void k_tile_apply(const uint32_t* counts, const uint64_t* tsum,
                                                     uint64_t* pref, uint64_t nchunks, uint32_t nslot,
                                                     uint64_t ntiles) {
  __shared__ uint64_t wsum[16];
  const uint64_t tile = blockIdx.x;
  const uint32_t slot = blockIdx.y;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const uint64_t c = tile * kTile + t;
  const uint64_t v = c < nchunks ? (IND ? uint64_t(counts[c * nslot + slot] != 0) : counts[c * nslot + slot]) : 0;
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t x = __shfl_up(inc, d, 64);
    if (lane >= d) inc += x;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  uint64_t off = tsum[uint64_t(slot) * ntiles + tile];
#pragma unroll
  for (int w = 0; w < 16; ++w)
    if (w < wid) off += wsum[w];
  if (c < nchunks) pref[c * nslot + slot] = off + inc - v;
}

// Sum of per-chunk zero counts of every sub-level (zeros_out pre-cleared);
// grid (tiles of 1024 chunks, sub-levels).
__global__ __launch_bounds__(256) void k_sum_zeros(const uint32_t* ztot, uint64_t nchunks, uint32_t g,
                                                   unsigned long long* zeros_out) {
  __shared__ uint64_t wsum[4];
  const uint32_t j = blockIdx.y;
  uint64_t s = 0;
  for (uint32_t k = threadIdx.x; k < kTile; k += 256) {
    const uint64_t c = uint64_t(blockIdx.x) * kTile + k;
    if (c < nchunks) s += ztot[c * g + j];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&zeros_out[j], (unsigned long long)(wsum[0] + wsum[1] + wsum[2] + wsum[3]));
}

__host__ __device__ inline uint32_t rev_bits(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

// Global start of every bucket of sub-levels 1..jmax; one block per sub-level.
// Chunk-local bucket b of sub-level j holds the symbols whose j-bit digit d
// has reverse_j(d) == b (matrix order).  Matrix: buckets in index order;
// tree: buckets in numeric digit order.
__global__ __launch_bounds__(256) void k_bucket_starts(const uint64_t* totals, uint64_t* gstart,
                                                       int tree, uint64_t* level_total) {
  __shared__ uint64_t tmp[256];
  const uint32_t j = blockIdx.x + 1;
  const uint32_t nb = 1u << j, slot0 = nb - 2;
  const uint32_t t = threadIdx.x;
  if (t < nb) {
    const uint32_t src = tree ? rev_bits(t, j) : t;  // t = position in output order
    tmp[t] = totals[slot0 + src];
  }
  __syncthreads();
  if (t == 0) {
    uint64_t run = 0;
    for (uint32_t i = 0; i < nb; ++i) {
      const uint64_t v = tmp[i];
      tmp[i] = run;
      run += v;
    }
    if (level_total) level_total[j] = run;
  }
  __syncthreads();
  if (t < nb) {
    const uint32_t src = tree ? rev_bits(t, j) : t;
    gstart[slot0 + src] = tmp[t];
  }
}

struct RunGeom {
  uint64_t* const* dst_level;   // per sub-level j (index j): destination level base
  uint64_t* dst_payload;        // payload plane k at dst_payload + k*pay_stride
  uint64_t pay_stride;
  uint32_t npay;
  uint32_t g, jmax, nslot;
  uint64_t n;
};

// Read 64 MSB-first bits starting at (possibly negative, > -64) bit position s.
__device__ __forceinline__ uint64_t window64(const uint64_t* a, int64_t s) {
  const int64_t w = s >> 6;  // arithmetic shift
  const uint32_t sh = uint32_t(s & 63);
  const uint64_t w0 = w >= 0 ? a[w] : 0ull;
  if (sh == 0) return w0;
  const uint64_t w1 = a[w + 1];
  return (w0 << sh) | (w1 >> (64 - sh));
}

// MSB-first mask of bit positions [a, b) within a word, 0 <= a < b <= 64.
__device__ __forceinline__ uint64_t range_mask(uint32_t a, uint32_t b) {
  const uint64_t hi = a == 0 ? ~0ull : (~0ull >> a);
  const uint64_t lo = b >= 64 ? ~0ull : ~(~0ull >> b);
  return hi & lo;
}

__device__ __forceinline__ uint32_t slot_level(uint32_t slot) {
  // slot = 2^j - 2 + b  ->  j
  return 31 - __clz(slot + 2);
}

// Compact run table of one construction pass: for every sub-level j the
// non-empty (chunk, bucket) runs in destination order, entries at
// [(2^j - 2) * nchunks, ...).  A run's length is the distance to the next
// entry's destination (runs of one level tile the level).
struct RunTable {
  const uint32_t* counts;
  const uint64_t* pref;     // exclusive prefix over chunks of counts
  const uint64_t* gstart;   // global bucket starts
  const uint64_t* nepref;   // exclusive prefix over chunks of (count != 0)
  const uint64_t* nebase;   // start of each bucket's entries in its level table
  uint64_t* tab_dst;
  uint64_t* tab_src;
  uint32_t nslot;
  uint64_t nchunks;
};

__global__ void k_make_runs(RunTable p) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= p.nchunks * p.nslot) return;
  const uint64_t c = i / p.nslot;
  const uint32_t slot = uint32_t(i - c * p.nslot);
  const uint32_t len = p.counts[i];
  if (!len) return;
  const uint32_t j = slot_level(slot);
  const uint32_t s0 = (1u << j) - 2;
  uint64_t loc = 0;  // chunk-local start: counts of the chunk's earlier buckets of this level
  for (uint32_t b = s0; b < slot; ++b) loc += p.counts[c * p.nslot + b];
  const uint64_t e = uint64_t(s0) * p.nchunks + p.nebase[slot] + p.nepref[i];
  p.tab_dst[e] = p.gstart[slot] + p.pref[i];
  p.tab_src[e] = c * kChunkSyms + loc;
}

// Destination-major gather: each wave writes 64 consecutive words of one
// level (or of every payload plane), each word exactly once.  The wave finds
// the run holding its first bit with a 64-way ballot search over the
// compact table, then walks the following runs.
struct GatherParams {
  const uint64_t* tab_dst;  // this sub-level's entries
  const uint64_t* tab_src;
  const uint64_t* level_ne; // number of entries (device, index j)
  uint32_t j;
  const uint64_t* src;      // chunk-local plane 0 (planes at src + k*src_stride)
  uint64_t src_stride;
  uint64_t* dst;            // destination plane 0 (planes at dst + k*dst_stride)
  uint64_t dst_stride;
  uint32_t nplanes;
  uint64_t n;
  uint32_t nwin;            // 64-word windows per wave (multiple of kGroup)
};

constexpr uint32_t kGatherWindows = 32;  // 64-word windows per wave (at most)
constexpr uint32_t kGroup = 4;           // windows whose loads are issued together

__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t lane) {
  const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

// Contributions of the cached runs 0..62 to one destination word [ws, we)
// (lanes hold entry i of the cache in bd/bs; run i ends where run i+1
// starts).  Every lane calls it (shuffles read 0 from inactive lanes);
// `pending` lanes own an unfinished word and stay pending when the word
// continues past the cache.
__device__ __forceinline__ uint64_t gather_word(const uint64_t* src, uint64_t bd, uint64_t bs, uint64_t ws,
                                                uint64_t we, bool& pending, uint64_t T0, bool& full0) {
  // last cached entry with start <= ws (entry 0 if none)
  uint32_t idx = 0;
#pragma unroll
  for (uint32_t step = 32; step >= 1; step >>= 1) {
    const uint64_t cand = __shfl(bd, int((idx + step) & 63), 64);
    if (idx + step < 64 && cand <= ws) idx += step;
  }
  uint64_t d0 = __shfl(bd, int(idx), 64);
  uint64_t d1 = __shfl(bd, int(idx + 1 < 64 ? idx + 1 : 63), 64);
  uint64_t s0 = __shfl(bs, int(idx), 64);
  // a word wholly inside one bucket-0 run (d0 < T0) was written in place
  full0 = pending && idx < 63 && d0 <= ws && d1 >= ws + 64 && d0 < T0;
  uint64_t v = 0;
  bool more = pending;
  while (__ballot(more)) {  // wave-uniform loop: shuffles below see every lane
    if (more) {
      if (idx >= 63) {
        more = false;  // run 63 ends outside the cache
      } else {
        const uint64_t a = max(d0, ws), b = min(d1, we);
        if (a < b) {
          const int64_t sb = int64_t(s0) - int64_t(d0) + int64_t(ws);
          v |= window64(src, sb) & range_mask(uint32_t(a - ws), uint32_t(b - ws));
        }
        if (d1 >= we) {
          more = false;
          pending = false;
        }
      }
    }
    const uint32_t nidx = more ? idx + 1 : idx;
    const uint64_t nd1 = __shfl(bd, int(nidx + 1 < 64 ? nidx + 1 : 63), 64);
    const uint64_t ns0 = __shfl(bs, int(nidx), 64);
    if (more) {
      idx = nidx;
      d0 = d1;
      d1 = nd1;
      s0 = ns0;
    }
  }
  return v;
}

// 64-entry run cache of a wave: lane i holds entry base+i (bd = destination
// start, n past the last entry; bs = source start).
struct RunCache {
  const uint64_t* tab_dst;
  const uint64_t* tab_src;
  uint64_t NE, n, base, bd, bs;
  __device__ void load(uint32_t lane) {
    bd = base + lane < NE ? tab_dst[base + lane] : n;
    bs = base + lane < NE ? tab_src[base + lane] : 0;
  }
};

// One window (64 words from w0) whatever its number of runs.
// A window that 63 or more runs overlap (sparse buckets: tiny runs from many
// chunks) is assembled run by run: each lane takes one run, cuts it into
// pieces of single destination words and ORs them into the wave's LDS copy
// of the window; work is proportional to the runs, not to runs x words.
__device__ void gather_window_runs(const uint64_t* src, uint64_t* dst, RunCache& rc, uint64_t w0, uint64_t e0,
                                   unsigned long long* stage, uint64_t T0) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n = rc.n, NE = rc.NE;
  const uint64_t sw0 = w0 << 6, sw1 = min<uint64_t>(sw0 + 4096, n);
  // last run starting before sw1 (64-way ballot search from e0)
  uint64_t lo = e0, hi = NE;
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + 63) >> 6;
    const uint64_t cand = lo + uint64_t(lane + 1) * step;
    const bool pred = cand < hi && rc.tab_dst[cand] < sw1;
    lo += uint64_t(__popcll(__ballot(pred))) * step;
    hi = min(hi, lo + step);
  }
  const uint64_t e1 = lo;
  stage[lane] = 0;
  wave_sync();
  uint64_t fullm = 0;  // bit i: word w0 + i lies wholly inside a bucket-0 run (written in place)
  for (uint64_t eb = e0; eb <= e1; eb += 64) {
    const uint64_t e = eb + lane;
    if (e <= e1) {
      const uint64_t d0 = rc.tab_dst[e], d1 = e + 1 < NE ? rc.tab_dst[e + 1] : n, s0 = rc.tab_src[e];
      if (d0 < T0) {
        const uint64_t fa = max((d0 + 63) >> 6, w0), fb = min(d1 >> 6, w0 + 64);
        if (fa < fb) fullm |= (fb - fa == 64 ? ~0ull : ((1ull << (fb - fa)) - 1)) << (fa - w0);
      }
      const uint64_t a = max(d0, sw0), b = min(d1, sw1);
      for (uint64_t pos = a; pos < b;) {
        const uint32_t off = uint32_t((pos - sw0) & 63);
        const uint32_t len = uint32_t(min<uint64_t>(64 - off, b - pos));
        const uint64_t v = (window64(src, int64_t(s0 + (pos - d0))) & range_mask(0, len)) >> off;
        if (v) atomicOr(&stage[(pos - sw0) >> 6], (unsigned long long)v);
        pos += len;
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int dd = 32; dd >= 1; dd >>= 1) fullm |= __shfl_xor(fullm, dd, 64);
  if (((w0 + lane) << 6) < n && !((fullm >> lane) & 1)) st_stream(dst + w0 + lane, uint64_t(stage[lane]));
  rc.base = e1;  // the window's last run may continue into the next window
  rc.load(lane);
}

// T0: words wholly inside one run that starts below T0 (bucket 0 when the
// local build wrote those words in place) are left alone.
__device__ __forceinline__ void gather_window(const uint64_t* src, uint64_t* dst, RunCache& rc, uint64_t w0,
                                              unsigned long long* stage, uint64_t T0 = 0) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n = rc.n;
  const uint64_t sw0 = w0 << 6, sw1 = min<uint64_t>(sw0 + 4096, n);
  const uint64_t ws = (w0 + lane) << 6, we = min<uint64_t>(ws + 64, n);
  const bool live = ws < n;
  // cached runs overlapping [sw0, sw1): ef..el (bd[0] <= sw0 holds)
  uint32_t ef = __popcll(__ballot(rc.bd <= sw0)) - 1;
  uint32_t el = __popcll(__ballot(rc.bd < sw1)) - 1;
  if (el == 63 && ef > 0) {  // move the cache to start at the window's first run
    rc.base += ef;
    rc.load(lane);
    ef = 0;
    el = __popcll(__ballot(rc.bd < sw1)) - 1;
  }
  if (el == 63) {
    gather_window_runs(src, dst, rc, w0, rc.base + ef, stage, T0);
    return;
  }
  uint64_t v = 0;
  bool full0 = false;
  if (el - ef < 4) {
    for (uint32_t e = ef; e <= el; ++e) {
      const uint64_t d0 = readlane64(rc.bd, e), d1 = readlane64(rc.bd, e + 1), s0 = readlane64(rc.bs, e);
      const uint64_t a = max(d0, ws), b = min(d1, we);
      full0 |= d0 <= ws && d1 >= ws + 64 && d0 < T0;
      if (live && a < b && !full0) {
        const int64_t sb = int64_t(s0) - int64_t(d0) + int64_t(ws);
        v |= window64(src, sb) & range_mask(uint32_t(a - ws), uint32_t(b - ws));
      }
    }
  } else {
    bool pending = live;  // every run of the window is in the cache
    v = gather_word(src, rc.bd, rc.bs, ws, we, pending, T0, full0);
  }
  if (live && !full0) st_stream(dst + w0 + lane, v);
}

// Writes the wave's destination words [W0, W0 + 64*nwin) of `nplanes`
// planes of one level from the level's run table (tab_dst/tab_src, NE
// entries in destination order; the runs tile [0, n)).  Groups of kGroup
// windows whose runs all sit in the cache go together: every lane locates
// the runs of its kGroup words (one per window) with uniform compares, then
// issues all their loads at once.  Other windows go one at a time.
__device__ __forceinline__ void gather_range(const uint64_t* tab_dst, const uint64_t* tab_src, uint64_t NE, const uint64_t* src0,
                             uint64_t src_stride, uint64_t* dst0, uint64_t dst_stride, uint32_t nplanes, uint64_t n,
                             uint64_t W0, uint32_t nwin, unsigned long long* stage) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t X = W0 << 6;
  if (X >= n || NE == 0) return;
  // last entry with dst <= X (64-way ballot search)
  uint64_t lo = 0, hi = NE;
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + 63) >> 6;
    const uint64_t cand = lo + uint64_t(lane + 1) * step;
    const bool pred = cand < hi && tab_dst[cand] <= X;
    const uint32_t cnt = __popcll(__ballot(pred));
    lo += uint64_t(cnt) * step;
    hi = min(hi, lo + step);
  }
  for (uint32_t pl = 0; pl < nplanes; ++pl) {
    const uint64_t* src = src0 + pl * src_stride;
    uint64_t* dst = dst0 + pl * dst_stride;
    RunCache rc{tab_dst, tab_src, NE, n, lo, 0, 0};
    rc.load(lane);
    for (uint32_t g0 = 0; g0 < nwin; g0 += kGroup) {
      const uint64_t sw0 = (W0 + 64ull * g0) << 6;
      if (sw0 >= n) break;
      bool done = false;
      if (g0 + kGroup <= nwin) {
        const uint64_t gend = min<uint64_t>(sw0 + 4096ull * kGroup, n);
        uint32_t ef = __popcll(__ballot(rc.bd <= sw0)) - 1;
        uint32_t el = __popcll(__ballot(rc.bd < gend)) - 1;
        if (el == 63 && ef > 0) {
          rc.base += ef;
          rc.load(lane);
          ef = 0;
          el = __popcll(__ballot(rc.bd < gend)) - 1;
        }
        if (el < 63) {
          uint64_t v[kGroup], d1k[kGroup];
          uint32_t idx[kGroup];
          const uint64_t ws0 = (W0 + 64ull * g0 + lane) << 6;  // word k starts at ws0 + 4096 k
#pragma unroll
          for (int k = 0; k < int(kGroup); ++k) idx[k] = ef;
          for (uint32_t e = ef + 1; e <= el; ++e) {
            const uint64_t bde = readlane64(rc.bd, e);
#pragma unroll
            for (int k = 0; k < int(kGroup); ++k) idx[k] += bde <= ws0 + 4096ull * k ? 1u : 0u;
          }
#pragma unroll
          for (int k = 0; k < int(kGroup); ++k) {
            const uint64_t ws = ws0 + 4096ull * k;
            const int64_t dl =
                int64_t(__shfl(rc.bs, int(idx[k]), 64)) - int64_t(__shfl(rc.bd, int(idx[k]), 64));
            d1k[k] = __shfl(rc.bd, int(idx[k] + 1), 64);
            v[k] = ws < n ? window64(src, dl + int64_t(ws)) : 0;
          }
#pragma unroll
          for (int k = 0; k < int(kGroup); ++k) {
            const uint64_t ws = ws0 + 4096ull * k, we = min<uint64_t>(ws + 64, n);
            bool more = ws < n && d1k[k] < we;
            if (ws < n) v[k] &= range_mask(0, uint32_t(min(d1k[k], we) - ws));
            // the word continues into later runs (run boundaries inside it)
            uint32_t e = idx[k] + 1;
            uint64_t d0 = d1k[k];
            while (__ballot(more)) {
              const uint64_t d1 = __shfl(rc.bd, int(min(e + 1, 63u)), 64);
              const uint64_t s0 = __shfl(rc.bs, int(min(e, 63u)), 64);
              if (more) {
                const int64_t sb = int64_t(s0) - int64_t(d0) + int64_t(ws);
                v[k] |= window64(src, sb) & range_mask(uint32_t(d0 - ws), uint32_t(min(d1, we) - ws));
                more = d1 < we;
                d0 = d1;
                ++e;
              }
            }
            if (ws < n) st_stream(dst + (ws >> 6), v[k]);
          }
          done = true;
        }
      }
      if (!done) {
#pragma unroll 1
        for (uint32_t win = g0; win < min(g0 + kGroup, nwin); ++win) {
          const uint64_t w0 = W0 + 64ull * win;
          if ((w0 << 6) >= n) break;
          gather_window(src, dst, rc, w0, stage);
        }
      }
    }
  }
}

// Two-word lanes: destination steps of 128 words, lane i owns words 2i and
// 2i+1 (16-byte loads and stores).  A step inside one run is a shift copy
// with one source offset for the whole wave; a step that up to 16 runs
// cross is a shift copy per lane when no lane's two words hold more than one
// run boundary; four steps issue their loads together.  Other steps go
// through gather_window (64-word windows, one word per lane).
// T0 > 0: the words wholly inside bucket-0 runs (runs starting below T0) were
// written in place by the local build and are skipped.  STEPS: steps whose
// loads go out together (the batched stitch takes 2: fewer registers, five
// waves a SIMD instead of three).
template <uint32_t STEPS>
__device__ __forceinline__ void gather_range_wide(const uint64_t* tab_dst, const uint64_t* tab_src, uint64_t NE,
                                                  const uint64_t* src, uint64_t* dst, uint64_t n, uint64_t W0,
                                                  uint32_t nstep, uint64_t first, unsigned long long* stage,
                                                  uint64_t T0 = 0) {
  constexpr uint32_t kSteps = STEPS;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t X = W0 << 6;
  if (X >= n || NE == 0) return;
  RunCache rc{tab_dst, tab_src, NE, n, first, 0, 0};  // first: last entry with dst <= X
  rc.load(lane);
  for (uint32_t g0 = 0; g0 < nstep; g0 += kSteps) {
    const uint64_t gw0 = W0 + 128ull * g0;
    if ((gw0 << 6) >= n) break;
    const uint64_t gend = min<uint64_t>((gw0 + 128ull * kSteps) << 6, n);
    uint32_t ef = __popcll(__ballot(rc.bd <= (gw0 << 6))) - 1;
    uint32_t el = __popcll(__ballot(rc.bd < gend)) - 1;
    if (el == 63 && ef > 0) {
      rc.base += ef;
      rc.load(lane);
      el = __popcll(__ballot(rc.bd < gend)) - 1;
    }
    // steps inside one cached run (whose end is cached too); a step wholly
    // inside a bucket-0 run written in place is done; a step that a few run
    // boundaries cross (no run in place) is a shift copy per lane from the
    // run its first bit lies in, when no lane's two words hold more than one
    // boundary: the lane whose words hold one also loads the next run's bits
    constexpr uint32_t kStepRuns = 16;  // runs of a step searched per lane
    uint32_t single = 0, inplace = 0, lanes = 0;
    int64_t dl[kSteps], dlb[kSteps];
    uint64_t bnd[kSteps];  // the next run's first bit when inside the lane's words, else ~0
#pragma unroll
    for (int k = 0; k < int(kSteps); ++k) {
      const uint64_t sw0 = (gw0 + 128ull * k) << 6, sw1 = min<uint64_t>(sw0 + 8192, n);
      dl[k] = 0;
      dlb[k] = 0;
      bnd[k] = ~0ull;
      if (sw0 < n && el < 63) {
        const uint32_t a = __popcll(__ballot(rc.bd <= sw0)) - 1;
        const uint32_t b = __popcll(__ballot(rc.bd < sw1)) - 1;
        if (a == b) {
          const uint64_t d0 = readlane64(rc.bd, a);
          if (d0 < T0) {
            if (readlane64(rc.bd, a + 1) >= sw0 + 8192) inplace |= 1u << k;
            // (a bucket-0 run ending inside the step: the level's last, partial word)
          } else {
            single |= 1u << k;
            dl[k] = int64_t(readlane64(rc.bs, a)) - int64_t(d0);
          }
        } else if (b - a <= kStepRuns && readlane64(rc.bd, a) >= T0) {
          const uint64_t ws = sw0 + 128ull * lane, we = ws + 128;
          uint32_t r = a;  // the lane's first run: the last one starting at or before ws
          for (uint32_t e = a + 1; e <= b; ++e) r += readlane64(rc.bd, e) <= ws ? 1u : 0u;
          // (r + 1 <= b + 1 <= 63: cached)
          const uint64_t d0 = __shfl(rc.bd, int(r), 64), s0 = __shfl(rc.bs, int(r), 64);
          const uint64_t B1 = __shfl(rc.bd, int(r + 1), 64), s1 = __shfl(rc.bs, int(r + 1), 64);
          const uint64_t B2 = __shfl(rc.bd, int(min(r + 2, 63u)), 64);
          const bool one = B1 < we && B1 < n;          // a boundary inside the lane's words
          const bool more = one && B2 < we && B2 < n;  // two or more
          if (!__ballot(ws < n && more)) {
            lanes |= 1u << k;
            dl[k] = int64_t(s0) - int64_t(d0);
            if (one) {
              bnd[k] = B1;
              dlb[k] = int64_t(s1) - int64_t(B1);
            }
          }
        }
      }
    }
    const uint32_t shift_copy = single | lanes;
    u64x2_u lo2[kSteps];
    uint64_t hi1[kSteps], xb[kSteps][3];
#pragma unroll
    for (int k = 0; k < int(kSteps); ++k) {
      const uint64_t ws = ((gw0 + 128ull * k) << 6) + 128ull * lane;
      if (((shift_copy >> k) & 1) && ws < n) {
        const uint64_t q = uint64_t(dl[k] + int64_t(ws)) >> 6;
        lo2[k] = ld_stream2(src + q);
        hi1[k] = gmem(src)[q + 2];
        if (bnd[k] < ws + 128) {
          // the boundary lane: the next run's bits for destination bits
          // ws.. start at source bit pb >= (its first source bit) - 127
          const int64_t pb = dlb[k] + int64_t(ws);
          const int64_t qb = pb >> 6;  // >= -2
          xb[k][0] = qb >= 0 ? gmem(src)[qb] : 0ull;
          xb[k][1] = qb + 1 >= 0 ? gmem(src)[qb + 1] : 0ull;
          xb[k][2] = gmem(src)[qb + 2];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < int(kSteps); ++k) {
      const uint64_t ws = ((gw0 + 128ull * k) << 6) + 128ull * lane;
      if (((shift_copy >> k) & 1) && ws < n) {
        const uint32_t sh = uint32_t(uint64_t(dl[k] + int64_t(ws)) & 63);
        uint64_t v0 = lo2[k].x, v1 = lo2[k].y;
        if (sh) {
          v0 = (v0 << sh) | (lo2[k].y >> (64 - sh));
          v1 = (v1 << sh) | (hi1[k] >> (64 - sh));
        }
        if (bnd[k] < ws + 128) {
          // destination bits [ws, B) from the lane's run, [B, ws + 128) from the next
          const uint32_t shb = uint32_t(uint64_t(dlb[k] + int64_t(ws)) & 63);
          uint64_t b0 = xb[k][0], b1 = xb[k][1];
          if (shb) {
            b0 = (b0 << shb) | (xb[k][1] >> (64 - shb));
            b1 = (b1 << shb) | (xb[k][2] >> (64 - shb));
          }
          const uint32_t cut = uint32_t(bnd[k] - ws);  // 1..127
          if (cut < 64) {
            const uint64_t m = range_mask(0, cut);
            v0 = (v0 & m) | (b0 & ~m);
            v1 = b1;
          } else {
            const uint64_t m = range_mask(0, cut - 64);
            v1 = (v1 & m) | (b1 & ~m);
          }
        }
        uint64_t* d = dst + (ws >> 6);
        if (ws + 128 <= n) {
          u64x2_u o;
          o.x = v0;
          o.y = v1;
          st_stream2(d, o);
        } else {  // the level's end: bits from n on are zero
          if (ws + 64 < n) {
            gmem(d)[0] = v0;
            gmem(d)[1] = v1 & range_mask(0, uint32_t(n - ws - 64));
          } else {
            gmem(d)[0] = v0 & range_mask(0, uint32_t(n - ws));
          }
        }
      }
    }
#pragma unroll 1
    for (uint32_t k = 0; k < kSteps; ++k) {
      if (((shift_copy | inplace) >> k) & 1) continue;
      const uint64_t w0 = gw0 + 128ull * k;
      if ((w0 << 6) >= n) break;
      gather_window(src, dst, rc, w0, stage, T0);
      if (((w0 + 64) << 6) < n) gather_window(src, dst, rc, w0 + 64, stage, T0);
    }
  }
}

// Every sub-level of a pass in one grid (blockIdx.y = sub-level - 1).
struct WideParams {
  const uint64_t* tab_dst;   // level j's entries from ((1 << j) - 2) * units
  const uint64_t* tab_src;
  const uint64_t* level_ne;  // entries per sub-level (device, index j)
  uint32_t* first;           // [(j - 1) * wpl + wave]: the wave's first entry
  const uint64_t* src;       // sub-level j at src + (j - 1) * src_stride
  uint64_t src_stride;
  uint64_t* dst[kMaxGroup];  // destination level of sub-level j (index j)
  uint64_t units, n, wpl;    // run units, symbols, waves per level
  uint32_t nstep;
  uint64_t nbits[kMaxGroup]; // bits of sub-level j when its own (0: n)
  const uint64_t* bucket0;   // or nullptr: [j] = bucket 0's total of sub-level j, whose
                             // runs' whole words the local build wrote in place
  uint32_t edge_blocks;      // with bucket0: the grid's first blocks build the shared words
};

// First table entry of every wave of the wide stitch: the last entry whose
// destination start is at or before the wave's first bit.
__global__ __launch_bounds__(256) void k_wave_starts(WideParams p) {
  const uint64_t k = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  const uint32_t j = blockIdx.y + 1;
  if (k >= p.wpl) return;
  const uint64_t X = k * 128ull * p.nstep * 64;
  const uint64_t* td = p.tab_dst + uint64_t((1u << j) - 2) * p.units;
  const uint64_t NE = p.level_ne[j];
  const uint64_t nj = p.nbits[j] ? p.nbits[j] : p.n;
  uint64_t lo = 0, hi = NE;  // td[lo] <= X < td[hi] (td[0] == 0)
  if (X < nj && NE > 0) {
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (td[mid] <= X) lo = mid;
      else hi = mid;
    }
  }
  p.first[(j - 1) * p.wpl + k] = uint32_t(lo);
}

// With bucket-0 runs written in place (p.bucket0), the waves whose words all
// lie below bucket 0's end have nothing left but the words two runs share:
// they exit, and the grid's first blocks build those words (bucket0_edge).
__device__ __forceinline__ uint64_t bucket0_skip_words(const WideParams& p, uint32_t j) {
  const uint64_t wpw = 128ull * p.nstep;  // words per wave
  return p.bucket0 ? (p.bucket0[j] >> 6) / wpw * wpw : 0;
}

// The words below the skipped waves' end that bucket-0 runs share (those
// not wholly inside one run, which the local build wrote in place): one
// thread per run, which builds the word holding the run's end when the run
// also holds that word's first bit, from the following runs as far as the
// word goes.  Every other word there lies inside one run.  Run e of
// sub-level j (bucket-0 runs are a level's first entries, at most one per
// unit).
__device__ __forceinline__ void bucket0_edge(const WideParams& p, uint32_t j, uint64_t e) {
  const uint64_t NE = p.level_ne[j];
  if (e >= NE) return;
  const uint64_t toff = uint64_t((1u << j) - 2) * p.units;
  const uint64_t* td = p.tab_dst + toff;
  const uint64_t* ts = p.tab_src + toff;
  const uint64_t nb = p.nbits[j] ? p.nbits[j] : p.n;
  const uint64_t T0 = p.bucket0[j];
  uint64_t d0 = td[e];
  if (d0 >= T0) return;
  uint64_t d1 = e + 1 < NE ? td[e + 1] : nb;
  if ((d1 & 63) == 0) return;
  const uint64_t w = (d1 - 1) >> 6, ws = w << 6;
  if (ws < d0 || w >= bucket0_skip_words(p, j)) return;
  const uint64_t* src = p.src + (j - 1) * p.src_stride;
  uint64_t v = 0, ee = e, s0 = ts[e];
  for (;;) {  // (the word lies below T0: bucket-0 runs only)
    const uint64_t a = max(d0, ws), b = min(d1, ws + 64);
    if (a < b) v |= window64(src, int64_t(s0) - int64_t(d0) + int64_t(ws)) & range_mask(uint32_t(a - ws), uint32_t(b - ws));
    if (d1 >= ws + 64 || ++ee >= NE) break;
    d0 = d1;
    d1 = ee + 1 < NE ? td[ee + 1] : nb;
    s0 = ts[ee];
  }
  p.dst[j][w] = v;
}

// With bucket-0 runs in place, blocks [0, edge_blocks) of each sub-level
// build the words two bucket-0 runs share (one thread per run); they run
// while the blocks of the skipped waves are dispatched and exit.
__global__ __launch_bounds__(256) void k_gather_wide(WideParams p) {
  __shared__ unsigned long long stage[4][64];
  const uint32_t j = blockIdx.y + 1;
  if (blockIdx.x < p.edge_blocks) {
    bucket0_edge(p, j, uint64_t(blockIdx.x) * 256 + threadIdx.x);
    return;
  }
  const uint64_t wave = uint64_t(blockIdx.x - p.edge_blocks) * 4 + (threadIdx.x >> 6);
  if (wave >= p.wpl) return;  // whole waves; no barrier below
  if ((wave + 1) * 128ull * p.nstep <= bucket0_skip_words(p, j)) return;
  const uint64_t toff = uint64_t((1u << j) - 2) * p.units;
  gather_range_wide<4>(p.tab_dst + toff, p.tab_src + toff, p.level_ne[j], p.src + (j - 1) * p.src_stride, p.dst[j],
                    p.nbits[j] ? p.nbits[j] : p.n, wave * 128 * p.nstep, p.nstep,
                    p.first[(j - 1) * p.wpl + wave], stage[threadIdx.x >> 6], p.bucket0 ? p.bucket0[j] : 0);
}

__global__ __launch_bounds__(256) void k_gather(GatherParams p) {
  __shared__ unsigned long long stage[4][64];
  const uint64_t W0 = (uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 64 * p.nwin;
  gather_range(p.tab_dst, p.tab_src, p.level_ne[p.j], p.src, p.src_stride, p.dst, p.dst_stride, p.nplanes, p.n,
               W0, p.nwin, stage[threadIdx.x >> 6]);
}

// Sub-levels 1..g-1 of one pass in one launch (blockIdx.y = j - 1): level
// j's entries at ((1 << j) - 2) * nchunks of the tables, its unit-local bits
// at src + (j - 1) * src_stride, one plane each.
struct GatherLevels {
  GatherParams gp;   // j, dst and nplanes unused
  uint64_t nchunks;
  uint64_t* dst[kMaxGroup];
};

__global__ __launch_bounds__(256) void k_gather_levels(GatherLevels q) {
  __shared__ unsigned long long stage[4][64];
  const GatherParams& p = q.gp;
  const uint32_t j = blockIdx.y + 1;
  const uint64_t toff = uint64_t((1u << j) - 2) * q.nchunks;
  const uint64_t W0 = (uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 64 * p.nwin;
  gather_range(p.tab_dst + toff, p.tab_src + toff, p.level_ne[j], p.src + (j - 1) * p.src_stride, 0, q.dst[j], 0,
               1, p.n, W0, p.nwin, stage[threadIdx.x >> 6]);
}

// Batch: waves are assigned to texts by wave_base (per text), level j.
struct BatchGatherParams {
  const uint64_t* wave_base;   // ntexts + 1
  const uint32_t* wave_text;   // text of every gather wave
  const uint64_t* chunk_base;  // ntexts + 1
  const uint64_t* text_n;
  const uint64_t* out0;        // word offset of each text's level 0
  const uint64_t* stride;      // words between a text's levels
  const uint64_t* level_ne;    // [t * 8 + j]
  const uint64_t* tab_dst;
  const uint64_t* tab_src;
  const uint64_t* src;         // chunk-local level j at src + (j-1)*src_stride
  uint64_t src_stride;
  uint64_t* out;
  uint32_t ntexts, j;
  uint64_t nchunks;
  uint32_t nwin;               // windows per wave (multiple of kGroup)
};

__global__ __launch_bounds__(256) void k_gather_batch(BatchGatherParams p) {
  __shared__ unsigned long long stage[4][64];
  const uint64_t wv = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (wv >= p.wave_base[p.ntexts]) return;
  const uint32_t t = p.wave_text[wv];
  const uint64_t W0 = (wv - p.wave_base[t]) * 64 * p.nwin;
  const uint64_t toff = uint64_t((1u << p.j) - 2) * p.nchunks + p.chunk_base[t] * (1ull << p.j);
  gather_range(p.tab_dst + toff, p.tab_src + toff, p.level_ne[uint64_t(t) * 8 + p.j],
               p.src + (p.j - 1) * p.src_stride, 0, p.out + p.out0[t] + p.j * p.stride[t], 0, 1, p.text_n[t], W0,
               p.nwin, stage[threadIdx.x >> 6]);
}

// Batch stitch over super-chunk runs: waves belong to texts by wave_base
// (steps of 128 words, nwin steps per wave; the text of every wave from a
// table); a wave finds its first run in its text's table slice with a 64-way
// ballot search.
__global__ __launch_bounds__(256) void k_gather_batch_wide(BatchGatherParams p) {
  __shared__ unsigned long long stage[4][64];
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wv = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (wv >= p.wave_base[p.ntexts]) return;  // whole waves; no barrier below
  const uint32_t t = p.wave_text[wv];
  const uint64_t W0 = (wv - p.wave_base[t]) * 128 * p.nwin;
  const uint64_t toff = uint64_t((1u << p.j) - 2) * p.nchunks + p.chunk_base[t] * (1ull << p.j);
  const uint64_t* td = p.tab_dst + toff;
  const uint64_t NE = p.level_ne[uint64_t(t) * 8 + p.j];
  const uint64_t n = p.text_n[t];
  const uint64_t X = W0 << 6;
  if (X >= n || NE == 0) return;
  uint64_t lo = 0, hi = NE;  // last entry with dst <= X
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + 63) >> 6;
    const uint64_t cand = lo + uint64_t(lane + 1) * step;
    const bool pred = cand < hi && td[cand] <= X;
    lo += uint64_t(__popcll(__ballot(pred))) * step;
    hi = min(hi, lo + step);
  }
  gather_range_wide<2>(td, p.tab_src + toff, NE, p.src + (p.j - 1) * p.src_stride, p.out + p.out0[t] + p.j * p.stride[t],
                    n, W0, p.nwin, lo, stage[threadIdx.x >> 6]);
}

// Fused offsets for the run table: per (tile of 1024 chunks, slot) the sums of
// counts and of non-empty indicators ...
__global__ __launch_bounds__(256) void k_tile_sums2(const uint32_t* counts, uint64_t* tsum, uint64_t* tnz,
                                                    uint64_t nchunks, uint32_t nslot, uint64_t ntiles) {
  __shared__ uint64_t ws[2][4];
  const uint64_t tile = blockIdx.x;
  const uint32_t slot = blockIdx.y;
  uint64_t v = 0, z = 0;
  for (uint32_t k = threadIdx.x; k < kTile; k += 256) {
    const uint64_t c = tile * kTile + k;
    if (c < nchunks) {
      const uint32_t x = counts[c * nslot + slot];
      v += x;
      z += x != 0;
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    v += __shfl_xor(v, d, 64);
    z += __shfl_xor(z, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    ws[0][threadIdx.x >> 6] = v;
    ws[1][threadIdx.x >> 6] = z;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tsum[uint64_t(slot) * ntiles + tile] = ws[0][0] + ws[0][1] + ws[0][2] + ws[0][3];
    tnz[uint64_t(slot) * ntiles + tile] = ws[1][0] + ws[1][1] + ws[1][2] + ws[1][3];
  }
}

// ... then (after the tile sums are scanned and the bucket starts known) every
// non-empty (chunk, bucket) run is written into its level's table.
// Slots from sc_slot0 on are the buckets of the order after the pass: their
// global start per (chunk, bucket) goes to dstpos (for the symbol reorder)
// instead of a run table.
__global__ __launch_bounds__(1024) void k_apply_runs(const uint32_t* counts, const uint64_t* tsum,
                                                     const uint64_t* tnz, const uint64_t* gstart,
                                                     const uint64_t* nebase, uint64_t* tab_dst, uint64_t* tab_src,
                                                     uint64_t nchunks, uint32_t nslot, uint64_t ntiles,
                                                     uint64_t* dstpos, uint32_t sc_slot0, uint64_t csyms) {
  __shared__ uint64_t wsum[2][16];
  const uint64_t tile = blockIdx.x;
  const uint32_t slot = blockIdx.y;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const uint64_t c = tile * kTile + t;
  const uint32_t x = c < nchunks ? counts[c * nslot + slot] : 0;
  uint64_t iv = x, iz = x != 0;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t a = __shfl_up(iv, d, 64), b = __shfl_up(iz, d, 64);
    if (lane >= d) {
      iv += a;
      iz += b;
    }
  }
  if (lane == 63) {
    wsum[0][wid] = iv;
    wsum[1][wid] = iz;
  }
  __syncthreads();
  if (!x) return;
  uint64_t ov = tsum[uint64_t(slot) * ntiles + tile], oz = tnz[uint64_t(slot) * ntiles + tile];
#pragma unroll
  for (int w = 0; w < 16; ++w)
    if (w < wid) {
      ov += wsum[0][w];
      oz += wsum[1][w];
    }
  const uint64_t pref = ov + iv - x, nzpref = oz + iz - 1;
  if (slot >= sc_slot0) {
    dstpos[c * (nslot - sc_slot0) + (slot - sc_slot0)] = gstart[slot] + pref;
    return;
  }
  const uint32_t j = slot_level(slot);
  const uint32_t s0 = (1u << j) - 2;
  uint64_t loc = 0;  // chunk-local start: counts of the chunk's earlier buckets of this level
  for (uint32_t b = s0; b < slot; ++b) loc += counts[c * nslot + b];
  const uint64_t e = uint64_t(s0) * nchunks + nebase[slot] + nzpref;
  tab_dst[e] = gstart[slot] + pref;
  tab_src[e] = c * csyms + loc;
}

// The same run tables in one kernel after k_tile_sums2, for passes of few
// buckets and tiles (every block scans the unscanned tile sums it needs):
// the bucket starts of the block's level, its tile's prefix, the entries,
// the entry counts per level and the first entry of every wave of the wide
// stitch (a run [d, d + x) holds the wave starts k * wave_bits inside it).
struct FusedRuns {
  const uint32_t* counts;
  const uint64_t* tsum;  // [slot * ntiles + tile], unscanned
  const uint64_t* tnz;
  uint64_t* tab_dst;
  uint64_t* tab_src;
  uint64_t* level_ne;
  uint32_t* wave_first;  // [(j - 1) * wpl + k]
  uint64_t units, ntiles, csyms, wave_bits, wpl;
  uint32_t nslot;
  int tree;
  // slots from sc_slot0 on are the next pass's order (a reorder follows):
  // their (unit, bucket) positions go to dstpos[u * (nslot - sc_slot0) + b]
  // instead of run tables; sc_slot0 == nslot: none
  uint64_t* dstpos;
  uint32_t sc_slot0;
  uint64_t* bucket0;  // or nullptr: bucket 0's total of sub-level j at [j]
};

__global__ __launch_bounds__(1024) void k_runs_fused(FusedRuns p) {
  __shared__ uint64_t wsum[2][16];
  __shared__ uint64_t base[2];  // bucket start, entry start (this slot, this tile)
  const uint64_t tile = blockIdx.x;
  const uint32_t slot = blockIdx.y;
  const uint32_t j = slot_level(slot), s0 = (1u << j) - 2, nb = 1u << j, b = slot - s0;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (wid == 0) {
    // totals of the level's buckets before this one (output order) and the
    // sums of this slot's earlier tiles
    const uint32_t pb = p.tree ? rev_bits(b, j) : b;
    uint64_t gv = 0, gz = 0, lz = 0, pv = 0, pz = 0, t0 = 0;
    for (uint32_t bb = 0; bb < nb; ++bb) {
      const uint32_t pos = p.tree ? rev_bits(bb, j) : bb;
      const uint64_t* sv = p.tsum + uint64_t(s0 + bb) * p.ntiles;
      const uint64_t* sz = p.tnz + uint64_t(s0 + bb) * p.ntiles;
      for (uint64_t i = lane; i < p.ntiles; i += 64) {
        const uint64_t v = sv[i], z = sz[i];
        lz += z;
        if (bb == 0) t0 += v;
        if (pos < pb) {
          gv += v;
          gz += z;
        }
        if (bb == b && i < tile) {
          pv += v;
          pz += z;
        }
      }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      gv += __shfl_xor(gv, d, 64);
      gz += __shfl_xor(gz, d, 64);
      lz += __shfl_xor(lz, d, 64);
      pv += __shfl_xor(pv, d, 64);
      pz += __shfl_xor(pz, d, 64);
      t0 += __shfl_xor(t0, d, 64);
    }
    if (lane == 0) {
      base[0] = gv + pv;
      base[1] = gz + pz;
      if (tile == 0 && b == 0) {
        p.level_ne[j] = lz;
        if (p.bucket0) p.bucket0[j] = t0;
      }
    }
  }
  const uint64_t c = tile * kTile + t;
  const uint32_t x = c < p.units ? p.counts[c * p.nslot + slot] : 0;
  uint64_t iv = x, iz = x != 0;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t a = __shfl_up(iv, d, 64), bz = __shfl_up(iz, d, 64);
    if (lane >= d) {
      iv += a;
      iz += bz;
    }
  }
  if (lane == 63) {
    wsum[0][wid] = iv;
    wsum[1][wid] = iz;
  }
  __syncthreads();
  if (!x) return;
  uint64_t ov = base[0], oz = base[1];
#pragma unroll
  for (int w = 0; w < 16; ++w)
    if (w < wid) {
      ov += wsum[0][w];
      oz += wsum[1][w];
    }
  const uint64_t d = ov + iv - x, e = oz + iz - 1;  // destination start, entry in the level
  if (slot >= p.sc_slot0) {
    p.dstpos[c * (p.nslot - p.sc_slot0) + (slot - p.sc_slot0)] = d;
    return;
  }
  uint64_t loc = 0;
  for (uint32_t bb = s0; bb < slot; ++bb) loc += p.counts[c * p.nslot + bb];
  p.tab_dst[uint64_t(s0) * p.units + e] = d;
  p.tab_src[uint64_t(s0) * p.units + e] = c * p.csyms + loc;
  if (p.wave_first) {
    for (uint64_t k = (d + p.wave_bits - 1) / p.wave_bits; k * p.wave_bits < d + x && k < p.wpl; ++k)
      p.wave_first[uint64_t(j - 1) * p.wpl + k] = uint32_t(e);
  }
}

// ---- batch offsets --------------------------------------------------------
// In-tile inclusive scans of one slot's counts and non-empty indicators over
// the tile's 1024 chunks (one thread per chunk), plus the tile's offsets:
// the chunk's global exclusive prefixes (pref, nzpref) and its count x.
__device__ __forceinline__ void tile_prefix(const uint32_t* counts, const uint64_t* tsum, const uint64_t* tnz,
                                            uint64_t nchunks, uint32_t nslot, uint64_t ntiles, uint64_t* wsum,
                                            uint64_t& pref, uint64_t& nzpref, uint32_t& x) {
  const uint64_t tile = blockIdx.x;
  const uint32_t slot = blockIdx.y;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const uint64_t c = tile * kTile + t;
  x = c < nchunks ? counts[c * nslot + slot] : 0;
  uint64_t iv = x, iz = x != 0;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t a = __shfl_up(iv, d, 64), b = __shfl_up(iz, d, 64);
    if (lane >= d) {
      iv += a;
      iz += b;
    }
  }
  if (lane == 63) {
    wsum[wid] = iv;
    wsum[16 + wid] = iz;
  }
  __syncthreads();
  uint64_t ov = tsum[uint64_t(slot) * ntiles + tile], oz = tnz[uint64_t(slot) * ntiles + tile];
#pragma unroll
  for (int w = 0; w < 16; ++w)
    if (w < wid) {
      ov += wsum[w];
      oz += wsum[16 + w];
    }
  pref = ov + iv - x;
  nzpref = oz + iz - (x != 0);
}

// Batch: global prefixes at every text's first chunk and inclusive ones at
// its last chunk.
__global__ __launch_bounds__(1024) void k_batch_text_bounds(const uint32_t* counts, const uint64_t* tsum,
                                                            const uint64_t* tnz, const uint32_t* chunk_text,
                                                            const uint64_t* chunk_base, uint64_t* tb_v,
                                                            uint64_t* tb_z, uint64_t* te_v, uint64_t* te_z,
                                                            uint64_t nchunks, uint32_t nslot, uint64_t ntiles) {
  __shared__ uint64_t wsum[32];
  uint64_t pref, nzpref;
  uint32_t x;
  tile_prefix(counts, tsum, tnz, nchunks, nslot, ntiles, wsum, pref, nzpref, x);
  const uint64_t c = uint64_t(blockIdx.x) * kTile + threadIdx.x;
  if (c >= nchunks) return;
  const uint32_t slot = blockIdx.y, t = chunk_text[c];
  if (chunk_base[t] == c) {
    tb_v[uint64_t(t) * nslot + slot] = pref;
    tb_z[uint64_t(t) * nslot + slot] = nzpref;
  }
  if (chunk_base[t + 1] == c + 1) {
    te_v[uint64_t(t) * nslot + slot] = pref + x;
    te_z[uint64_t(t) * nslot + slot] = nzpref + (x != 0);
  }
}

__global__ void k_batch_text_totals(const uint64_t* tb_v, const uint64_t* tb_z, const uint64_t* te_v,
                                    const uint64_t* te_z, const uint64_t* chunk_base, uint64_t* ttot,
                                    uint64_t* tnz, uint32_t ntexts, uint32_t nslot) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= uint64_t(ntexts) * nslot) return;
  const uint32_t t = uint32_t(i / nslot);
  const bool empty = chunk_base[t] == chunk_base[t + 1];
  ttot[i] = empty ? 0 : te_v[i] - tb_v[i];
  tnz[i] = empty ? 0 : te_z[i] - tb_z[i];
}

// Batch run tables: every non-empty (chunk, bucket) run, text-local prefix =
// global prefix - the text's base.
__global__ __launch_bounds__(1024) void k_batch_apply_runs(const uint32_t* counts, const uint64_t* tsum,
                                                           const uint64_t* tnz, const uint32_t* chunk_text,
                                                           const uint64_t* chunk_base, const uint64_t* tb_v,
                                                           const uint64_t* tb_z, const uint64_t* gstart,
                                                           const uint64_t* nebase, uint64_t* tab_dst,
                                                           uint64_t* tab_src, uint64_t nchunks, uint32_t nslot,
                                                           uint64_t ntiles, uint64_t csyms) {
  __shared__ uint64_t wsum[32];
  uint64_t pref, nzpref;
  uint32_t x;
  tile_prefix(counts, tsum, tnz, nchunks, nslot, ntiles, wsum, pref, nzpref, x);
  const uint64_t c = uint64_t(blockIdx.x) * kTile + threadIdx.x;
  if (c >= nchunks || !x) return;
  const uint32_t slot = blockIdx.y, t = chunk_text[c];
  const uint64_t ts = uint64_t(t) * nslot + slot;
  const uint32_t j = slot_level(slot);
  const uint32_t s0 = (1u << j) - 2;
  uint64_t loc = 0;
  for (uint32_t b = s0; b < slot; ++b) loc += counts[c * nslot + b];
  const uint64_t e = uint64_t(s0) * nchunks + chunk_base[t] * (1ull << j) + nebase[ts] + (nzpref - tb_z[ts]);
  tab_dst[e] = gstart[ts] + (pref - tb_v[ts]);
  tab_src[e] = c * csyms + loc;
}

// Bucket starts of every (text, sub-level); grid (ntexts, jmax), 64 threads.
__global__ __launch_bounds__(64) void k_bucket_starts_batch(const uint64_t* ttot, const uint64_t* tnz,
                                                            uint64_t* gstart, uint64_t* nebase, uint64_t* level_ne,
                                                            uint32_t nslot, int tree) {
  const uint32_t t = blockIdx.x, j = blockIdx.y + 1;
  if (threadIdx.x != 0) return;
  const uint32_t nb = 1u << j, slot0 = nb - 2;
  const uint64_t* tt = ttot + uint64_t(t) * nslot + slot0;
  const uint64_t* tz = tnz + uint64_t(t) * nslot + slot0;
  uint64_t run = 0, runz = 0;
  for (uint32_t q = 0; q < nb; ++q) {  // q = position in output order
    const uint32_t b = tree ? rev_bits(q, j) : q;
    gstart[uint64_t(t) * nslot + slot0 + b] = run;
    nebase[uint64_t(t) * nslot + slot0 + b] = runz;
    run += tt[b];
    runz += tz[b];
  }
  level_ne[uint64_t(t) * 8 + j] = runz;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u32(unsigned int* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The text of every batch gather wave: waves [wave_base[t], wave_base[t+1]).
__global__ __launch_bounds__(256) void k_wave_text(const uint64_t* wave_base, uint32_t ntexts, uint32_t* wave_text) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntexts) return;
  for (uint64_t w = wave_base[t]; w < wave_base[t + 1]; ++w) wave_text[w] = t;
}

// Positions of every chunk's buckets in the order after a super-chunk pass:
// the unit's bucket position plus the counts of the unit's earlier chunks.
// One block per unit, thread (k, q) = chunk k of the unit, bucket q.
__global__ __launch_bounds__(256) void k_chunk_dstpos(const uint64_t* unit_pos, const uint32_t* chunk_sc,
                                                      uint64_t* chunk_pos, uint64_t nchunks, uint32_t nq,
                                                      uint32_t cpu) {
  const uint64_t u = blockIdx.x;
  const uint32_t k = threadIdx.x / nq, q = threadIdx.x % nq;
  const uint64_t c = u * cpu + k;
  if (k >= cpu || c >= nchunks) return;
  uint64_t pos = unit_pos[u * nq + q];
  for (uint32_t k2 = 0; k2 < k; ++k2) pos += chunk_sc[(u * cpu + k2) * nq + q];
  chunk_pos[c * nq + q] = pos;
}

}  // namespace

void launch_scan_slots(hipStream_t s, const uint32_t* counts, uint64_t* pref, uint64_t* totals,
                       uint64_t nchunks, uint32_t nslot, uint64_t* tile_scratch, bool indicator) {
  const uint64_t ntiles = (nchunks + kTile - 1) / kTile;
  if (indicator)
    hipLaunchKernelGGL(k_tile_sums<true>, dim3(uint32_t(ntiles), nslot), dim3(256), 0, s, counts, tile_scratch,
                       nchunks, nslot, ntiles);
  else
    hipLaunchKernelGGL(k_tile_sums<false>, dim3(uint32_t(ntiles), nslot), dim3(256), 0, s, counts, tile_scratch,
                       nchunks, nslot, ntiles);
  hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, tile_scratch, totals, ntiles);
  if (indicator)
    hipLaunchKernelGGL(k_tile_apply<true>, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, counts, tile_scratch,
                       pref, nchunks, nslot, ntiles);
  else
    hipLaunchKernelGGL(k_tile_apply<false>, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, counts, tile_scratch,
                       pref, nchunks, nslot, ntiles);
  PWM_HIP_TRY(hipGetLastError());
}

uint64_t scan_scratch_words(uint64_t nchunks, uint32_t nslot) {
  return ((nchunks + kTile - 1) / kTile) * std::max<uint32_t>(nslot, 1);
}

namespace {

// PWMHIP_GENERIC_K1=1 routes every pass through the generic local build
// (for comparisons).
bool small_k1_disabled() {
  const char* e = std::getenv("PWMHIP_GENERIC_K1");  // read at every call
  return e && e[0] == '1';
}

// PWMHIP_CHUNK_RUNS=1: one run per chunk (no super-chunk assembly).
bool super_disabled() {
  const char* e = std::getenv("PWMHIP_CHUNK_RUNS");  // read at every call
  return e && e[0] == '1';
}

// PWMHIP_STITCH_ALL=1: single texts of 2 or 3 byte levels leave every word
// of the sub-levels to the stitch (no bucket-0 runs written in place).
bool direct_disabled() {
  const char* e = std::getenv("PWMHIP_STITCH_ALL");  // read at every call
  return e && e[0] == '1';
}

// Cooperative: the workgroups wait for each other's counts, so all of them
// must be resident (one per CU: the kernel's LDS allows no second).
template <int G>
void launch_direct(pwmhip_ctx* ctx, const K1Params& prm, const DirectParams& dp) {
  const uint64_t nsuper = (prm.nchunks + kSuper - 1) / kSuper;
  const uint64_t grid = std::min<uint64_t>({nsuper, uint64_t(ctx->num_cus), uint64_t(kGapEntries) + 1});
  K1Params a0 = prm;
  DirectParams a1 = dp;
  void* args[] = {&a0, &a1};
  PWM_HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_local_direct<G>), dim3(uint32_t(grid)),
                                         dim3(64 * kSuper), args, 0, ctx->stream));
}

// PWMHIP_NARROW_GATHER=1: one destination word per lane in the stitch.
bool wide_gather_disabled() {
  const char* e = std::getenv("PWMHIP_NARROW_GATHER");  // read at every call
  return e && e[0] == '1';
}

// 128-word steps per wave of the wide stitch (multiple of 4; PWMHIP_WIDE_STEPS).
// dflt: 8 for one-pass builds, 4 for the passes of larger alphabets (shorter
// runs: more waves in flight paid off, 2.90 against 2.91 ms on the 20-level
// Zipf matrix; the 3-level DNA stitch ran 0.128 against 0.137 ms at 8).
uint32_t wide_gather_steps(uint32_t dflt) {
  const char* e = std::getenv("PWMHIP_WIDE_STEPS");  // read at every call
  const long x = e ? std::atol(e) : long(dflt);
  return uint32_t(x >= 4 && x <= 256 ? (x + 3) / 4 * 4 : dflt);
}

// PWMHIP_SCATTER_RANKED=1: the symbol reorder ranks 64 symbols per step
// (k_scatter) instead of the per-lane counting sort.
bool scatter_seg_disabled() {
  const char* e = std::getenv("PWMHIP_SCATTER_RANKED");  // read at every call
  return e && e[0] == '1';
}

// Chunks per super-chunk unit of the plain builds: 16 (one workgroup of 16
// waves per CU) or 8 (two workgroups of 8 per CU); PWMHIP_SUPER_WAVES.
uint32_t super_waves() {
  const char* e = std::getenv("PWMHIP_SUPER_WAVES");  // read at every call
  return e && std::atoi(e) == 8 ? 8u : 16u;
}

template <int G, int WAVES>
void launch_super_w(pwmhip_ctx* ctx, const K1Params& prm) {
  const uint64_t nsuper = (prm.nchunks + WAVES - 1) / WAVES;
  const uint64_t grid = std::min<uint64_t>(nsuper, uint64_t(ctx->num_cus) * (16 / WAVES));
  if (prm.width == 1)
    hipLaunchKernelGGL((k_local_super<G, false, WAVES>), dim3(uint32_t(grid)), dim3(64 * WAVES), 0, ctx->stream, prm);
  else
    hipLaunchKernelGGL((k_local_super<G, true, WAVES>), dim3(uint32_t(grid)), dim3(64 * WAVES), 0, ctx->stream, prm);
  PWM_HIP_TRY(hipGetLastError());
}

// A pass over W-chunk units followed by a reorder (next-order counts).
template <int G, int W>
void launch_super_sc(pwmhip_ctx* ctx, const K1Params& prm) {
  const uint64_t nsuper = (prm.nchunks + W - 1) / W;
  const uint64_t grid = std::min<uint64_t>(nsuper, uint64_t(ctx->num_cus) * (16 / W));
  if (prm.width == 1)
    hipLaunchKernelGGL((k_local_super<G, false, W, true>), dim3(uint32_t(grid)), dim3(64 * W), 0, ctx->stream, prm);
  else
    hipLaunchKernelGGL((k_local_super<G, true, W, true>), dim3(uint32_t(grid)), dim3(64 * W), 0, ctx->stream, prm);
  PWM_HIP_TRY(hipGetLastError());
}

// One pass of g (2..4) levels over W-chunk units, with (sc) or without the
// next-order counts.
template <int W>
void launch_pass_units(pwmhip_ctx* ctx, const K1Params& prm, uint32_t g, bool sc) {
  switch (g * 2 + (sc ? 1 : 0)) {
    case 4: launch_super_w<2, W>(ctx, prm); break;
    case 5: launch_super_sc<2, W>(ctx, prm); break;
    case 6: launch_super_w<3, W>(ctx, prm); break;
    case 7: launch_super_sc<3, W>(ctx, prm); break;
    case 8: launch_super_w<4, W>(ctx, prm); break;
    default: launch_super_sc<4, W>(ctx, prm); break;
  }
}

template <int G>
void launch_super(pwmhip_ctx* ctx, const K1Params& prm) {
  if (super_waves() == 8) launch_super_w<G, 8>(ctx, prm);
  else launch_super_w<G, 16>(ctx, prm);
}

}  // namespace

uint64_t wide_stitch_first_words(uint64_t max_bits, uint32_t nsub) {
  const uint32_t nstep = wide_gather_steps(8);
  const uint64_t wpl = (words_for(max_bits) + 128ull * nstep - 1) / (128ull * nstep);
  return std::max<uint64_t>(1, wpl * nsub);
}

void launch_wide_stitch(pwmhip_ctx* ctx, const WideStitch& w) {
  if (w.nsub == 0) return;
  WideParams wp{};
  wp.tab_dst = w.tab_dst;
  wp.tab_src = w.tab_src;
  wp.level_ne = w.level_ne;
  wp.first = w.first;
  wp.src = w.src;
  wp.src_stride = w.src_stride;
  uint64_t maxb = 0;
  for (uint32_t j = 1; j <= w.nsub; ++j) {
    wp.dst[j] = w.dst[j];
    wp.nbits[j] = std::max<uint64_t>(w.nbits[j], 1);
    maxb = std::max(maxb, w.nbits[j]);
  }
  wp.units = w.units;
  wp.n = maxb;
  wp.nstep = wide_gather_steps(8);
  wp.wpl = (words_for(maxb) + 128ull * wp.nstep - 1) / (128ull * wp.nstep);
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_wave_starts, dim3(uint32_t((wp.wpl + 255) / 256), w.nsub), dim3(256), 0, s, wp);
  hipLaunchKernelGGL(k_gather_wide, dim3(uint32_t((wp.wpl + 3) / 4), w.nsub), dim3(256), 0, s, wp);
  PWM_HIP_TRY(hipGetLastError());
}

void launch_huff_super(pwmhip_ctx* ctx, const HuffSuperParams& hp) {
  static_assert(kSuperChunkSyms == kSuperSyms, "one super-chunk size");
  const uint64_t nsuper = (hp.nchunks + kSuper - 1) / kSuper;
  const uint32_t grid = uint32_t(std::min<uint64_t>(nsuper, uint64_t(ctx->num_cus)));
  const uint8_t* lut = compress_table(ctx);
  switch (hp.levels) {
    case 2: hipLaunchKernelGGL(k_huff_super<2>, dim3(grid), dim3(64 * kSuper), 0, ctx->stream, hp, lut); break;
    case 3: hipLaunchKernelGGL(k_huff_super<3>, dim3(grid), dim3(64 * kSuper), 0, ctx->stream, hp, lut); break;
    case 4: hipLaunchKernelGGL(k_huff_super<4>, dim3(grid), dim3(64 * kSuper), 0, ctx->stream, hp, lut); break;
    default: throw status_error{PWMHIP_ERR_INVALID, "Huffman super-chunk build: 2..4 levels"};
  }
  PWM_HIP_TRY(hipGetLastError());
}

namespace {

template <int W>
void launch_super_batch_w(pwmhip_ctx* ctx, const K1Params& prm, uint32_t L, const UnitMeta* meta) {
  const uint32_t grid =
      uint32_t(std::min<uint64_t>(std::max<uint64_t>(prm.nchunks, 1), uint64_t(ctx->num_cus) * (16 / W)));
  const bool wide = prm.width != 1;
  const dim3 gr(grid), bl(64 * W);
  hipStream_t s = ctx->stream;
  switch (L * 2 + (wide ? 1 : 0)) {
    case 4: hipLaunchKernelGGL((k_local_super_batch<2, false, W>), gr, bl, 0, s, prm, meta); break;
    case 5: hipLaunchKernelGGL((k_local_super_batch<2, true, W>), gr, bl, 0, s, prm, meta); break;
    case 6: hipLaunchKernelGGL((k_local_super_batch<3, false, W>), gr, bl, 0, s, prm, meta); break;
    case 7: hipLaunchKernelGGL((k_local_super_batch<3, true, W>), gr, bl, 0, s, prm, meta); break;
    case 8: hipLaunchKernelGGL((k_local_super_batch<4, false, W>), gr, bl, 0, s, prm, meta); break;
    default: hipLaunchKernelGGL((k_local_super_batch<4, true, W>), gr, bl, 0, s, prm, meta); break;
  }
  PWM_HIP_TRY(hipGetLastError());
}

void launch_super_batch(pwmhip_ctx* ctx, const K1Params& prm, uint32_t L, const UnitMeta* meta, uint32_t waves) {
  if (waves == 8) launch_super_batch_w<8>(ctx, prm, L, meta);
  else launch_super_batch_w<16>(ctx, prm, L, meta);
}

// Chunks per unit of a batch: every text's last unit is on average half
// empty, so with many texts for their chunks (the half units saved by 8
// chunks, T x 4 slots, above 1/16 of all chunks) units of 8 chunks, else
// 16 (longer runs for the stitch).  20,000 DNA texts of median 64 Ki
// symbols: local build 1.48 -> 1.29 ms, stitch 0.47 -> 0.56 ms.
// PWMHIP_SUPER_WAVES=8 or 16 fixes it.
uint32_t batch_unit_waves(uint64_t T, uint64_t total_chunks) {
  const char* e = std::getenv("PWMHIP_SUPER_WAVES");  // read at every call
  if (e) return std::atoi(e) == 8 ? 8u : 16u;
  return T * 4 * 16 > total_chunks ? 8u : 16u;
}

template <int NP>
void launch_k1(pwmhip_ctx* ctx, const K1Params& prm) {
  constexpr int WAVES = waves_for<NP>();
  const uint64_t grid = std::min<uint64_t>((prm.nchunks + WAVES - 1) / WAVES, uint64_t(ctx->num_cus));
  hipLaunchKernelGGL(k_local_build<NP>, dim3(uint32_t(grid)), dim3(64 * WAVES), 0, ctx->stream, prm);
  PWM_HIP_TRY(hipGetLastError());
}

template <int G, int SC>
void launch_small(pwmhip_ctx* ctx, const K1Params& prm) {
  const uint64_t grid = std::min<uint64_t>((prm.nchunks + 15) / 16, uint64_t(ctx->num_cus));
  if (prm.width == 1)
    hipLaunchKernelGGL((k_local_small<G, SC, false>), dim3(uint32_t(grid)), dim3(64 * 16), 0, ctx->stream, prm);
  else
    hipLaunchKernelGGL((k_local_small<G, SC, true>), dim3(uint32_t(grid)), dim3(64 * 16), 0, ctx->stream, prm);
  PWM_HIP_TRY(hipGetLastError());
}

// A pass of g <= 4 levels from the text, with (sc) or without the bucket
// counts of the order after it.
void dispatch_small(pwmhip_ctx* ctx, const K1Params& prm, bool sc) {
  switch (prm.g * 2 + (sc ? 1 : 0)) {
    case 2: launch_small<1, 0>(ctx, prm); return;
    case 3: launch_small<1, 1>(ctx, prm); return;
    case 4: launch_small<2, 0>(ctx, prm); return;
    case 5: launch_small<2, 1>(ctx, prm); return;
    case 6: launch_small<3, 0>(ctx, prm); return;
    case 7: launch_small<3, 1>(ctx, prm); return;
    case 8: launch_small<4, 0>(ctx, prm); return;
    default: launch_small<4, 1>(ctx, prm); return;
  }
}

void dispatch_k1(pwmhip_ctx* ctx, const K1Params& prm) {
  const uint32_t np = prm.np;
  if (!prm.planes_in && prm.g == np && np <= 4 && !small_k1_disabled()) {
    dispatch_small(ctx, prm, false);
    return;
  }
  if (np <= 1) launch_k1<1>(ctx, prm);
  else if (np <= 2) launch_k1<2>(ctx, prm);
  else if (np <= 3) launch_k1<3>(ctx, prm);
  else if (np <= 4) launch_k1<4>(ctx, prm);
  else if (np <= 8) launch_k1<8>(ctx, prm);
  else if (np <= 12) launch_k1<12>(ctx, prm);
  else if (np <= 16) launch_k1<16>(ctx, prm);
  else if (np <= 20) launch_k1<20>(ctx, prm);
  else throw status_error{PWMHIP_ERR_UNSUPPORTED, "more than 20 levels"};
}

}  // namespace

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
uint64_t layout_structure(pwmhip_ws* ws, uint64_t levels, const uint64_t* bits) {
  uint64_t off = 0;
  for (uint64_t l = 0; l < levels; ++l) {
    ws->level_offset[l] = off;
    ws->level_bits[l] = bits[l];
    off += level_stride(bits[l]);
  }
  ws->levels = levels;
  return std::max<uint64_t>(off, 32);
}

void alloc_structure(pwmhip_ws* ws, uint64_t levels, const uint64_t* bits, uint64_t* reuse,
                     uint64_t reuse_bytes) {
  ws->alloc_bytes = layout_structure(ws, levels, bits) * 8;
  if (reuse) {
    if (reuse_bytes >= ws->alloc_bytes) {
      ws->d_words = reuse;
      ws->alloc_bytes = reuse_bytes;
      return;
    }
    (void)hipFree(reuse);
  }
  void* p = nullptr;
  PWM_HIP_TRY(hipMalloc(&p, ws->alloc_bytes));
  ws->d_words = static_cast<uint64_t*>(p);
}

const uint8_t* compress_table(pwmhip_ctx* ctx) {
  if (!ctx->d_lut) {
    void* p = nullptr;
    PWM_HIP_TRY(hipMalloc(&p, 65536));
    hipLaunchKernelGGL(k_make_lut, dim3(256), dim3(256), 0, ctx->stream, static_cast<uint8_t*>(p));
    PWM_HIP_TRY(hipGetLastError());
    ctx->d_lut = static_cast<uint8_t*>(p);
  }
  return ctx->d_lut;
}

namespace {

// Windows per gather wave: enough waves to fill the chip (about 64 per CU).
uint32_t gather_windows(uint64_t W, uint32_t num_cus) {
  const uint64_t target = uint64_t(num_cus) * 64;
  uint64_t nwin = (W + 64 * target - 1) / (64 * target);
  nwin = std::min<uint64_t>(std::max<uint64_t>(nwin, kGroup), kGatherWindows);
  return uint32_t((nwin + kGroup - 1) / kGroup * kGroup);
}

template <typename TI, typename TO, int G>
void launch_scatter_k(hipStream_t s, const ScatterParams& sp) {
  if constexpr (sizeof(TI) <= 4 && sizeof(TO) <= 4) {
    if (!scatter_seg_disabled()) {
      constexpr int SW = scatter_seg_waves<TO>();
      hipLaunchKernelGGL((k_scatter_seg<TI, TO, G>), dim3(uint32_t((sp.nchunks + SW - 1) / SW)), dim3(64 * SW), 0,
                         s, sp);
      return;
    }
  }
  constexpr int WB = scatter_waves<TO>();
  const dim3 grid(uint32_t((sp.nchunks + WB - 1) / WB)), block(64 * WB);
  hipLaunchKernelGGL((k_scatter<TI, TO, G>), grid, block, 0, s, sp);
}

template <typename TI, int G>
void launch_scatter_o(hipStream_t s, uint32_t wout, const ScatterParams& sp) {
  switch (wout) {
    case 1: launch_scatter_k<TI, uint8_t, G>(s, sp); break;
    case 2: launch_scatter_k<TI, uint16_t, G>(s, sp); break;
    case 4: launch_scatter_k<TI, uint32_t, G>(s, sp); break;
    default: launch_scatter_k<TI, uint64_t, G>(s, sp); break;
  }
}

template <int G>
void launch_scatter_g(hipStream_t s, uint32_t win, uint32_t wout, const ScatterParams& sp) {
  switch (win) {
    case 1: launch_scatter_o<uint8_t, G>(s, wout, sp); break;
    case 2: launch_scatter_o<uint16_t, G>(s, wout, sp); break;
    case 4: launch_scatter_o<uint32_t, G>(s, wout, sp); break;
    default: launch_scatter_o<uint64_t, G>(s, wout, sp); break;
  }
}

void launch_scatter(hipStream_t s, uint32_t win, uint32_t wout, uint32_t g, const ScatterParams& sp) {
  switch (g) {
    case 1: launch_scatter_g<1>(s, win, wout, sp); break;
    case 2: launch_scatter_g<2>(s, win, wout, sp); break;
    case 3: launch_scatter_g<3>(s, win, wout, sp); break;
    default: launch_scatter_g<4>(s, win, wout, sp); break;
  }
  PWM_HIP_TRY(hipGetLastError());
}

// Bytes per symbol that hold `bits` bits.
uint32_t width_for_bits(uint32_t bits) { return bits <= 8 ? 1 : bits <= 16 ? 2 : bits <= 32 ? 4 : 8; }

}  // namespace

// Trees above one pass: the matrix levels' buckets moved into numeric order.
// Its time (host-side, it ends with a synchronisation) counts as stitch
// time of the build.
static void tree_order_timed(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width, uint32_t L,
                             pwmhip_ws* ws) {
  const auto t0 = std::chrono::steady_clock::now();
  wm_to_wt_device(ctx, d_text, n, width, L, ws);
  const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  ctx->last_ms[2] += dt;
  ctx->last_ms[3] += dt;
}

// Passes of at most 4 levels over the text; between passes the symbols are
// reordered into the next pass's level order, so every pass sees few
// buckets per chunk (long runs) whatever the number of levels.  Trees are
// built as matrices and their buckets moved into numeric order at the end.
void build_passes_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width, uint32_t L, bool tree,
                         bool check_range, pwmhip_ws* ws) {
  hipStream_t s = ctx->stream;
  const uint64_t nchunks = (n + kChunkSyms - 1) / kChunkSyms;
  const uint64_t stride = level_stride(n);
  const uint32_t P = (L + 3) / 4;
  std::vector<uint32_t> gs;
  for (uint32_t i = 0; i < P; ++i) gs.push_back(L / P + (i < L % P ? 1 : 0));
  const uint32_t gmax = gs[0];
  const uint32_t nslot_max = (2u << gmax) - 2u;
  const uint64_t ntiles_max = (nchunks + kTile - 1) / kTile;
  const size_t sz_cnt = std::max<size_t>(1, nchunks * nslot_max);
  const size_t text_bytes = (n * width + 255) & ~size_t(255);
  // runs per 16-chunk unit (the single-text super-chunk K1, with the
  // next-order counts per chunk) and the wide stitch
  // (with 8 or more 16-chunk units per CU; below that the per-unit barriers
  // cost more than the longer runs save: 64 Mi bytes, 8 levels, 0.383 ms in
  // chunks against 0.398 ms in units).  PWMHIP_PASS_UNITS=1 / 0, read at
  // every call, takes units / chunks whatever the size.
  const char* pu = std::getenv("PWMHIP_PASS_UNITS");
  const bool sup = !small_k1_disabled() && !super_disabled() && !wide_gather_disabled() &&
                   (pu ? pu[0] == '1' : (nchunks + 15) / 16 >= 8ull * ctx->num_cus);
  // 16 chunks per unit for byte symbols; 8 (two workgroups per CU, whose
  // loads and compute overlap) for wider ones
  auto pass_waves = [&](uint32_t w) { return w == 1 ? 16u : 8u; };
  const uint64_t units_max = (nchunks + 7) / 8;
  const uint64_t tstride =
      sup ? std::max((nchunks + 15) / 16 * (16 * kChunkSyms / 64), units_max * (8 * kChunkSyms / 64)) + 64 : stride;
  const uint32_t wide_nstep = wide_gather_steps(4);
  const uint64_t wide_wpl = (words_for(n) + 128ull * wide_nstep - 1) / (128ull * wide_nstep);
  auto carve = [&](arena& A, std::vector<void*>& v) {
    v.clear();
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, size_t(gmax - 1) * tstride)));  // 0 temp
    v.push_back(A.take<uint32_t>(sz_cnt));                       // 1 counts
    v.push_back(A.take<uint64_t>(sz_cnt));                       // 2 tnz scratch
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, ntiles_max * nslot_max)));  // 3 tsum
    v.push_back(A.take<uint64_t>(nslot_max));                    // 4 totals
    v.push_back(A.take<uint64_t>(nslot_max));                    // 5 netot
    v.push_back(A.take<uint64_t>(nslot_max));                    // 6 gstart
    v.push_back(A.take<uint64_t>(nslot_max));                    // 7 nebase
    v.push_back(A.take<uint64_t>(kMaxGroup + 1));                // 8 level_ne
    v.push_back(A.take<uint64_t>(sz_cnt));                       // 9 tab_dst
    v.push_back(A.take<uint64_t>(sz_cnt));                       // 10 tab_src
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, nchunks << gmax)));  // 11 dstpos
    v.push_back(A.take<uint64_t>(64));                           // 12 zeros
    v.push_back(A.take<unsigned long long>(1));                  // 13 sym_or
    v.push_back(A.take<uint8_t>(P > 1 ? text_bytes : 1));        // 14 reorder buffer A
    v.push_back(A.take<uint8_t>(P > 2 ? text_bytes : 1));        // 15 reorder buffer B
    v.push_back(A.take<uint64_t>(sup ? std::max<size_t>(1, units_max << gmax) : 1));  // 16 unit positions
    v.push_back(A.take<uint32_t>(sup ? std::max<size_t>(1, nchunks << gmax) : 1));  // 17 chunk counts
    v.push_back(A.take<uint32_t>(sup ? std::max<size_t>(1, size_t(gmax) * wide_wpl) : 1));  // 18 wave firsts
    v.push_back(A.take<uint8_t>(sup && P > 1 ? std::max<size_t>(1, nchunks * 1024) : 1));  // 19 lane counts
  };
  std::vector<void*> b;
  {
    arena sizing;
    sizing.cap = ~size_t(0) >> 1;
    carve(sizing, b);
    ctx->scratch.reset();
    ctx->scratch.ensure(sizing.used);
  }
  carve(ctx->scratch, b);
  auto U64 = [&](int i) { return static_cast<uint64_t*>(b[i]); };
  uint32_t* counts = static_cast<uint32_t*>(b[1]);
  PWM_HIP_TRY(hipMemsetAsync(b[13], 0, 8, s));
  PWM_HIP_TRY(hipMemsetAsync(b[12], 0, 64 * 8, s));
  const uint8_t* lut = compress_table(ctx);

  PWM_HIP_TRY(hipEventRecord(ctx->ev[0], s));
  const void* cur = d_text;
  uint32_t cw = width;  // bytes per symbol of the current pass's input
  uint32_t l0 = 0;
  for (uint32_t i = 0; i < P; ++i) {
    const uint32_t g = gs[i];
    const bool last = i + 1 == P;
    const uint32_t nslot = last ? (1u << g) - 2u : (2u << g) - 2u;
    const uint32_t R = L - l0;  // significant bits of the current symbols
    K1Params k1{};
    k1.text = cur;
    k1.width = cw;
    k1.n = n;
    k1.L = R;
    k1.l0 = 0;
    k1.g = g;
    k1.np = g;
    k1.out_level0 = ws->d_words + ws->level_offset[l0];
    k1.temp_levels = U64(0);
    k1.temp_payload = U64(0);
    k1.temp_stride = tstride;
    k1.counts = counts;
    k1.nslot = nslot;
    k1.zeros_out = reinterpret_cast<unsigned long long*>(U64(12) + l0);
    k1.sym_or = static_cast<unsigned long long*>(b[13]);
    k1.nchunks = nchunks;
    k1.lut = lut;
    const uint32_t pw = pass_waves(cw);  // chunks per unit of this pass
    const uint64_t units = (nchunks + pw - 1) / pw;
    const uint64_t nu = sup ? units : nchunks;  // run units
    const uint64_t usyms = sup ? pw * kChunkSyms : kChunkSyms;
    if (sup) {
      k1.chunk_sc = static_cast<uint32_t*>(b[17]);
      k1.lane_sc = last ? nullptr : static_cast<uint8_t*>(b[19]);
      if (pw == 16) launch_pass_units<16>(ctx, k1, g, !last);
      else launch_pass_units<8>(ctx, k1, g, !last);
    } else {
      dispatch_small(ctx, k1, !last);
    }
    if (i == 0) PWM_HIP_TRY(hipEventRecord(ctx->ev[1], s));
    if (nslot) {
      const uint64_t ntiles = (nu + kTile - 1) / kTile;
      hipLaunchKernelGGL(k_tile_sums2, dim3(uint32_t(ntiles), nslot), dim3(256), 0, s, counts, U64(3), U64(2),
                         nu, nslot, ntiles);
      // every block scans the tile sums it needs: bucket starts, run tables,
      // entry counts, the next order's positions and the wide stitch's wave
      // starts from one launch (each block reads all tiles of its level)
      const bool fused = ntiles <= 1024;
      if (fused) {
        FusedRuns fr{};
        fr.counts = counts;
        fr.tsum = U64(3);
        fr.tnz = U64(2);
        fr.tab_dst = U64(9);
        fr.tab_src = U64(10);
        fr.level_ne = U64(8);
        fr.wave_first = sup ? static_cast<uint32_t*>(b[18]) : nullptr;
        fr.units = nu;
        fr.ntiles = ntiles;
        fr.csyms = usyms;
        fr.wave_bits = 128ull * 64 * wide_nstep;
        fr.wpl = wide_wpl;
        fr.nslot = nslot;
        fr.tree = 0;
        fr.dstpos = sup ? U64(16) : U64(11);
        fr.sc_slot0 = last ? nslot : (1u << g) - 2u;
        hipLaunchKernelGGL(k_runs_fused, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, fr);
      } else {
        hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, U64(3), U64(4), ntiles);
        hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, U64(2), U64(5), ntiles);
        const uint32_t jb = last ? g - 1 : g;  // sub-levels with bucket starts
        hipLaunchKernelGGL(k_bucket_starts, dim3(jb), dim3(256), 0, s, U64(4), U64(6), 0, (uint64_t*)nullptr);
        hipLaunchKernelGGL(k_bucket_starts, dim3(jb), dim3(256), 0, s, U64(5), U64(7), 0, U64(8));
        hipLaunchKernelGGL(k_apply_runs, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, counts, U64(3), U64(2),
                           U64(6), U64(7), U64(9), U64(10), nu, nslot, ntiles, sup ? U64(16) : U64(11),
                           last ? nslot : (1u << g) - 2u, usyms);
      }
      if (sup && !last)
        hipLaunchKernelGGL(k_chunk_dstpos, dim3(uint32_t(units)), dim3(256), 0, s, U64(16),
                           static_cast<const uint32_t*>(b[17]), U64(11), nchunks, 1u << g, pw);
      if (sup) {
        WideParams wp{};
        wp.tab_dst = U64(9);
        wp.tab_src = U64(10);
        wp.level_ne = U64(8);
        wp.first = static_cast<uint32_t*>(b[18]);
        wp.src = U64(0);
        wp.src_stride = tstride;
        for (uint32_t j = 1; j < g; ++j) wp.dst[j] = ws->d_words + ws->level_offset[l0 + j];
        wp.units = nu;
        wp.n = n;
        wp.wpl = wide_wpl;
        wp.nstep = wide_nstep;
        if (!fused)
          hipLaunchKernelGGL(k_wave_starts, dim3(uint32_t((wide_wpl + 255) / 256), g - 1), dim3(256), 0, s, wp);
        if (i == 0) PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
        hipLaunchKernelGGL(k_gather_wide, dim3(uint32_t((wide_wpl + 3) / 4), g - 1), dim3(256), 0, s, wp);
      } else if (i == 0) {
        PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
      }
      const uint64_t W = words_for(n);
      const uint32_t nwin = gather_windows(W, ctx->num_cus);
      const uint32_t gblocks = uint32_t((W + 256ull * nwin - 1) / (256ull * nwin));
      if (!sup && g > 1) {
        GatherLevels gl{};
        gl.gp.nwin = nwin;
        gl.gp.tab_dst = U64(9);
        gl.gp.tab_src = U64(10);
        gl.gp.level_ne = U64(8);
        gl.gp.n = n;
        gl.gp.src = U64(0);
        gl.gp.src_stride = stride;
        gl.nchunks = nchunks;
        for (uint32_t j = 1; j < g; ++j) gl.dst[j] = ws->d_words + ws->level_offset[l0 + j];
        hipLaunchKernelGGL(k_gather_levels, dim3(gblocks, g - 1), dim3(256), 0, s, gl);
      }
      if (!last) {
        ScatterParams sp{};
        sp.in = cur;
        sp.out = b[14 + (i & 1)];
        sp.n = n;
        sp.sh = R - g;
        sp.dstpos = U64(11);
        sp.counts = sup ? static_cast<const uint32_t*>(b[17]) : counts;
        sp.nslot = sup ? (1u << g) : nslot;
        sp.slot0 = sup ? 0u : (1u << g) - 2u;
        sp.nchunks = nchunks;
        sp.lane_sc = sup ? static_cast<const uint8_t*>(b[19]) : nullptr;
        const uint32_t wout = width_for_bits(R - g);
        launch_scatter(s, cw, wout, g, sp);
        cur = sp.out;
        cw = wout;
      }
    } else if (i == 0) {
      PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
    }
    PWM_HIP_TRY(hipGetLastError());
    l0 += g;
  }
  PWM_HIP_TRY(hipEventRecord(ctx->ev[3], s));
  if (ctx->pinned_words < 65) {
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    void* hp = nullptr;
    PWM_HIP_TRY(hipHostMalloc(&hp, 128 * 8, hipHostMallocDefault));
    ctx->pinned = static_cast<uint64_t*>(hp);
    ctx->pinned_words = 128;
  }
  PWM_HIP_TRY(hipMemcpyAsync(ctx->pinned, b[12], L * 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipMemcpyAsync(ctx->pinned + 64, b[13], 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  for (int k = 0; k < 3; ++k) {
    PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]));
    ctx->last_ms[k] = ms;
  }
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
  ctx->last_ms[3] = ms;
  if (check_range && L < 64 && (ctx->pinned[64] >> L) != 0)
    throw status_error{PWMHIP_ERR_SYMBOL_RANGE, "symbol >= 2^levels"};
  uint64_t zeros[64];
  std::memcpy(zeros, ctx->pinned, L * 8);
  if (tree) tree_order_timed(ctx, d_text, n, width, L, ws);
  ws->n_zeros = tree ? 0 : L;
  for (uint32_t l = 0; l < L; ++l) ws->zeros[l] = tree ? 0 : zeros[l];
}

namespace {

// Symbols per streamed slice: a multiple of the super-chunk
// (PWMHIP_STREAM_SLICE_SYMS overrides the 256 Mi default).
uint64_t stream_slice_syms() {
  const char* e = std::getenv("PWMHIP_STREAM_SLICE_SYMS");
  uint64_t v = e ? std::strtoull(e, nullptr, 10) : (uint64_t(256) << 20);
  const uint64_t us = uint64_t(super_waves()) * kChunkSyms;
  v = std::max<uint64_t>(v, us);
  return v / us * us;
}

// Local build of a host text in slices: slice k is copied into staging
// buffer k % 2 on the copy stream while the build of slice k - 1 runs on
// the context's stream; a buffer is refilled once the build that read it
// has finished.  Every slice but the last is whole super-chunks, so its
// outputs (level 0, chunk-local rows, counts) sit at fixed offsets.
void stream_local_build(pwmhip_ctx* ctx, const void* h_text, uint64_t n, uint32_t width, uint32_t L,
                        const K1Params& k1) {
  if (!ctx->copy_stream) PWM_HIP_TRY(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
  for (auto& e : ctx->sev)
    if (!e) PWM_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const uint64_t slice = stream_slice_syms();
  const uint64_t bytes = std::min<uint64_t>(slice, n) * width;
  ctx->text_scratch.ensure(2 * ((bytes + 255) & ~uint64_t(255)) + 64);
  char* buf[2] = {static_cast<char*>(ctx->text_scratch.base),
                  static_cast<char*>(ctx->text_scratch.base) + ((bytes + 255) & ~uint64_t(255))};
  const char* src = static_cast<const char*>(h_text);
  uint64_t k = 0;
  for (uint64_t off = 0; off < n; off += slice, ++k) {
    const uint32_t b = uint32_t(k & 1);
    const uint64_t m = std::min<uint64_t>(slice, n - off);
    if (k >= 2) PWM_HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->sev[2 + b], 0));
    PWM_HIP_TRY(hipMemcpyAsync(buf[b], src + off * width, m * width, hipMemcpyHostToDevice, ctx->copy_stream));
    PWM_HIP_TRY(hipEventRecord(ctx->sev[b], ctx->copy_stream));
    PWM_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->sev[b], 0));
    K1Params ks = k1;
    ks.text = buf[b];
    ks.n = m;
    ks.nchunks = (m + kChunkSyms - 1) / kChunkSyms;
    ks.out_level0 = k1.out_level0 + off / 64;
    const uint64_t us = uint64_t(super_waves()) * kChunkSyms;
    ks.temp_levels = k1.temp_levels + (off / us) * (us / 64);
    ks.counts = k1.counts + (off / us) * k1.nslot;
    switch (L) {
      case 2: launch_super<2>(ctx, ks); break;
      case 3: launch_super<3>(ctx, ks); break;
      default: launch_super<4>(ctx, ks); break;
    }
    PWM_HIP_TRY(hipEventRecord(ctx->sev[2 + b], ctx->stream));
  }
}

}  // namespace

void build_plain_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width,
                        uint32_t L, bool tree, bool check_range, pwmhip_ws* ws, const void* h_stream,
                        const NoAlloc* na) {
  if (na && (h_stream || L < 2 || L > 4))
    throw status_error{PWMHIP_ERR_UNSUPPORTED, "allocation-free builds: 2 to 4 levels, device text"};
  if (h_stream && (L < 2 || L > 4 || small_k1_disabled() || super_disabled()))
    throw status_error{PWMHIP_ERR_UNSUPPORTED, "streamed texts: 2 to 4 levels"};
  if (L > 4 && !small_k1_disabled()) {
    build_passes_device(ctx, d_text, n, width, L, tree, check_range, ws);
    return;
  }
  hipStream_t s = ctx->stream;
  const uint64_t nchunks = (n + kChunkSyms - 1) / kChunkSyms;
  const uint64_t stride = level_stride(n);
  const bool multi = L > uint32_t(kMaxGroup);
  const bool tree_direct = tree && !multi;
  // 2..4 levels from a text: runs per super-chunk; otherwise per chunk
  const bool sup = L >= 2 && L <= 4 && !small_k1_disabled() && !super_disabled();
  const uint32_t sw = super_waves();
  const uint64_t units = sup ? (nchunks + sw - 1) / sw : nchunks;
  const uint64_t unit_syms = sup ? uint64_t(sw) * kChunkSyms : kChunkSyms;
  const uint64_t tstride = sup ? units * (unit_syms / 64) + 64 : stride;
  // wide stitch: 128-word steps, wide_nstep per wave, one grid for all sub-levels
  const bool wide = sup && !wide_gather_disabled();
  const uint32_t wide_nstep = wide_gather_steps(8);
  const uint64_t wide_wpl = (words_for(n) + 128ull * wide_nstep - 1) / (128ull * wide_nstep);
  // 2 or 3 byte levels: bucket-0 runs written in place by the local build
  // (16-chunk units, offsets in one fused launch, the wide stitch)
  const bool direct = wide && width == 1 && (L == 2 || L == 3) && !h_stream && !na && super_waves() == 16 &&
                      (units + kTile - 1) / kTile <= 1024 && !direct_disabled();

  // group plan (at most 64 / kMaxGroup groups; a fixed array, no heap)
  struct {
    uint32_t v[PWMHIP_MAX_LEVELS / kMaxGroup];
    uint32_t n = 0;
    size_t size() const { return n; }
    uint32_t operator[](size_t i) const { return v[i]; }
  } groups;
  for (uint32_t l = 0; l < L; l += kMaxGroup) groups.v[groups.n++] = std::min<uint32_t>(kMaxGroup, L - l);

  // scratch sizing
  const uint32_t gmax = groups[0];
  const uint32_t npay0 = L - gmax;
  const uint32_t nslot_max = (2u << gmax) - 2u;
  // Every scratch buffer is taken twice: once to size the arena, once for real.
  const size_t sz_temp = size_t(std::max<uint32_t>(1, gmax - 1)) * tstride;
  const size_t sz_pay = std::max<size_t>(1, size_t(npay0) * stride);
  const size_t sz_pong = std::max<size_t>(1, size_t(npay0 > kMaxGroup ? npay0 - kMaxGroup : 0) * stride);
  const size_t sz_cnt = std::max<size_t>(1, units * nslot_max);
  const size_t sz_tile = scan_scratch_words(units, nslot_max);
  const size_t sz_tab = std::max<size_t>(1, units * nslot_max);
  uint32_t* wave_first = nullptr;
  unsigned long long* agg = nullptr;  // direct: published bucket-0 counts per unit
  uint64_t* bucket0 = nullptr;        // direct: bucket 0's total per sub-level
  auto carve = [&](arena& A, uint64_t*& temp_levels, uint64_t*& temp_payload, uint64_t*& S_ping,
                   uint64_t*& S_pong, uint32_t*& counts, uint64_t*& pref, uint32_t*& ztot, uint64_t*& totals,
                   uint64_t*& gstart, uint64_t*& zeros_dev, unsigned long long*& sym_or, uint64_t*& tiles,
                   uint64_t*& nepref, uint64_t*& netot, uint64_t*& nebase, uint64_t*& level_ne, uint64_t*& tab_dst,
                   uint64_t*& tab_src) {
    temp_levels = A.take<uint64_t>(sz_temp);
    temp_payload = A.take<uint64_t>(sz_pay);
    S_ping = A.take<uint64_t>(sz_pay);
    S_pong = A.take<uint64_t>(sz_pong);
    counts = A.take<uint32_t>(sz_cnt);
    pref = A.take<uint64_t>(sz_cnt);
    ztot = A.take<uint32_t>(nchunks * kMaxGroup);
    totals = A.take<uint64_t>(nslot_max);
    gstart = A.take<uint64_t>(nslot_max);
    zeros_dev = A.take<uint64_t>(64);
    sym_or = A.take<unsigned long long>(1);
    tiles = A.take<uint64_t>(sz_tile);
    nepref = A.take<uint64_t>(sz_cnt);
    netot = A.take<uint64_t>(nslot_max);
    nebase = A.take<uint64_t>(nslot_max);
    level_ne = A.take<uint64_t>(kMaxGroup + 1);
    tab_dst = A.take<uint64_t>(sz_tab);
    tab_src = A.take<uint64_t>(sz_tab);
    wave_first = A.take<uint32_t>(std::max<size_t>(1, wide ? size_t(gmax) * wide_wpl : 1));
    agg = A.take<unsigned long long>(direct ? units : 1);
    bucket0 = A.take<uint64_t>(kMaxGroup + 1);
  };
  uint64_t *temp_levels, *temp_payload, *S_ping, *S_pong, *pref, *totals, *gstart, *zeros_dev, *tiles;
  uint64_t *nepref, *netot, *nebase, *level_ne, *tab_dst, *tab_src;
  uint32_t *counts, *ztot;
  unsigned long long* sym_or;
  {
    arena sizing;  // dry run on an unbounded arena to measure
    sizing.cap = ~size_t(0) >> 1;
    carve(sizing, temp_levels, temp_payload, S_ping, S_pong, counts, pref, ztot, totals, gstart, zeros_dev,
          sym_or, tiles, nepref, netot, nebase, level_ne, tab_dst, tab_src);
    if (na) {
      if (sizing.used > na->cap) throw status_error{PWMHIP_ERR_BUDGET, "text too long for the static scratch"};
    } else {
      ctx->scratch.reset();
      ctx->scratch.ensure(sizing.used);
    }
  }
  arena na_view;  // the static scratch, carved but never allocated or released
  if (na) {
    na_view.base = na->scratch;
    na_view.cap = na->cap;
  }
  carve(na ? na_view : ctx->scratch, temp_levels, temp_payload, S_ping, S_pong, counts, pref, ztot, totals, gstart,
        zeros_dev, sym_or, tiles, nepref, netot, nebase, level_ne, tab_dst, tab_src);
  PWM_HIP_TRY(hipMemsetAsync(sym_or, 0, 8, s));
  PWM_HIP_TRY(hipMemsetAsync(zeros_dev, 0, 64 * 8, s));  // (direct: [63] fault)
  if (direct) PWM_HIP_TRY(hipMemsetAsync(agg, 0, units * 8, s));

  PWM_HIP_TRY(hipEventRecord(ctx->ev[0], s));
  const uint64_t* S_in = nullptr;
  uint64_t* S_out = S_ping;
  uint64_t* S_next = S_pong;
  uint32_t l0 = 0;
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const uint32_t g = groups[gi];
    const uint32_t np = L - l0;
    const uint32_t npay = np - g;
    const uint32_t jmax = npay ? g : g - 1;
    const uint32_t nslot = jmax ? (2u << jmax) - 2u : 0;

    K1Params k1{};
    k1.text = S_in ? nullptr : d_text;
    k1.width = width;
    k1.planes_in = S_in;
    k1.in_stride = stride;
    k1.n = n;
    k1.L = L;
    k1.l0 = l0;
    k1.g = g;
    k1.np = np;
    k1.out_level0 = ws->d_words + ws->level_offset[l0];
    k1.temp_levels = temp_levels;
    k1.temp_payload = temp_payload;
    k1.temp_stride = tstride;
    k1.counts = counts;
    k1.nslot = nslot;
    k1.zeros_out = reinterpret_cast<unsigned long long*>(zeros_dev + l0);
    k1.sym_or = sym_or;
    k1.nchunks = nchunks;
    k1.lut = na ? na->lut : compress_table(ctx);
    if (h_stream) {
      stream_local_build(ctx, h_stream, n, width, L, k1);
    } else if (direct) {
      DirectParams dp{};
      dp.agg = agg;
      dp.fault = reinterpret_cast<unsigned long long*>(zeros_dev + 63);
      for (uint32_t j = 1; j < L; ++j) dp.dst[j] = ws->d_words + ws->level_offset[j];
      // PWMHIP_TEST_UNPUBLISHED=1 (tests of the give-up path): no unit
      // publishes its counts and a wait gives up after 16 reloads, so the
      // build ends with PWMHIP_ERR_TIMEOUT and every kernel exits normally
      const char* tu = std::getenv("PWMHIP_TEST_UNPUBLISHED");  // read at every call
      const bool unpublished = tu && tu[0] == '1';
      dp.publish = unpublished ? 0u : 1u;
      dp.spin_limit = unpublished ? 16u : kSpinLimit;
      if (L == 2) launch_direct<2>(ctx, k1, dp);
      else launch_direct<3>(ctx, k1, dp);
    } else if (sup) {
      switch (L) {
        case 2: launch_super<2>(ctx, k1); break;
        case 3: launch_super<3>(ctx, k1); break;
        default: launch_super<4>(ctx, k1); break;
      }
    } else {
      dispatch_k1(ctx, k1);
    }
    if (gi == 0) PWM_HIP_TRY(hipEventRecord(ctx->ev[1], s));

    if (nslot) {
      const uint64_t nchunks = units;  // run units from here on
      const uint64_t ntiles = (nchunks + kTile - 1) / kTile;
      uint64_t* tsum = tiles;
      uint64_t* tnz = pref;  // tile sums of the indicators (pref is free scratch here)
      hipLaunchKernelGGL(k_tile_sums2, dim3(uint32_t(ntiles), nslot), dim3(256), 0, s, counts, tsum, tnz, nchunks,
                         nslot, ntiles);
      if (wide && ntiles <= 1024) {
        FusedRuns fr{};
        fr.counts = counts;
        fr.tsum = tsum;
        fr.tnz = tnz;
        fr.tab_dst = tab_dst;
        fr.tab_src = tab_src;
        fr.level_ne = level_ne;
        fr.wave_first = wave_first;
        fr.units = nchunks;
        fr.ntiles = ntiles;
        fr.csyms = unit_syms;
        fr.wave_bits = 128ull * 64 * wide_nstep;
        fr.wpl = wide_wpl;
        fr.nslot = nslot;
        fr.tree = tree_direct ? 1 : 0;
        fr.sc_slot0 = nslot;
        fr.bucket0 = direct ? bucket0 : nullptr;
        hipLaunchKernelGGL(k_runs_fused, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, fr);
      } else {
        hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, tsum, totals, ntiles);
        hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, tnz, netot, ntiles);
        hipLaunchKernelGGL(k_bucket_starts, dim3(jmax), dim3(256), 0, s, totals, gstart, tree_direct ? 1 : 0,
                           (uint64_t*)nullptr);
        hipLaunchKernelGGL(k_bucket_starts, dim3(jmax), dim3(256), 0, s, netot, nebase, tree_direct ? 1 : 0,
                           level_ne);
        hipLaunchKernelGGL(k_apply_runs, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, counts, tsum, tnz, gstart,
                           nebase, tab_dst, tab_src, nchunks, nslot, ntiles, (uint64_t*)nullptr, nslot, unit_syms);
        if (wide) {
          WideParams wp{};
          wp.tab_dst = tab_dst;
          wp.level_ne = level_ne;
          wp.first = wave_first;
          wp.units = nchunks;
          wp.n = n;
          wp.wpl = wide_wpl;
          wp.nstep = wide_nstep;
          hipLaunchKernelGGL(k_wave_starts, dim3(uint32_t((wide_wpl + 255) / 256), jmax), dim3(256), 0, s, wp);
        }
      }
      if (gi == 0) PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
      const uint64_t W = words_for(n);
      const uint32_t nwin = gather_windows(W, ctx->num_cus);
      const uint32_t gblocks = uint32_t((W + 256ull * nwin - 1) / (256ull * nwin));
      for (uint32_t j = 1; j <= jmax; ++j) {
        GatherParams gp{};
        gp.nwin = nwin;
        const uint64_t toff = uint64_t((1u << j) - 2) * nchunks;
        gp.tab_dst = tab_dst + toff;
        gp.tab_src = tab_src + toff;
        gp.level_ne = level_ne;
        gp.j = j;
        gp.n = n;
        if (j < g) {
          gp.src = temp_levels + (j - 1) * tstride;
          gp.dst = ws->d_words + ws->level_offset[l0 + j];
          gp.nplanes = 1;
        } else {
          gp.src = temp_payload;
          gp.src_stride = stride;
          gp.dst = S_out;
          gp.dst_stride = stride;
          gp.nplanes = npay;
        }
        if (!wide) hipLaunchKernelGGL(k_gather, dim3(gblocks), dim3(256), 0, s, gp);
      }
      if (wide) {
        WideParams wp{};
        wp.tab_dst = tab_dst;
        wp.tab_src = tab_src;
        wp.level_ne = level_ne;
        wp.first = wave_first;
        wp.src = temp_levels;
        wp.src_stride = tstride;
        for (uint32_t j = 1; j <= jmax; ++j) wp.dst[j] = ws->d_words + ws->level_offset[l0 + j];
        wp.units = nchunks;
        wp.n = n;
        wp.wpl = wide_wpl;
        wp.nstep = wide_nstep;
        wp.bucket0 = direct ? bucket0 : nullptr;
        // (bucket-0 runs: the first entries of a level, at most one per unit)
        wp.edge_blocks = direct ? uint32_t((nchunks + 255) / 256) : 0u;
        hipLaunchKernelGGL(k_gather_wide, dim3(wp.edge_blocks + uint32_t((wide_wpl + 3) / 4), jmax), dim3(256), 0, s,
                           wp);
      }
    } else if (gi == 0) {
      PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
    }
    PWM_HIP_TRY(hipGetLastError());
    l0 += g;
    S_in = S_out;
    std::swap(S_out, S_next);
  }
  PWM_HIP_TRY(hipEventRecord(ctx->ev[3], s));

  // results: zeros per level and the symbol OR (the allocation-free build
  // reads them back into host stack memory instead of the pinned buffer)
  uint64_t na_res[65] = {};
  if (!na && ctx->pinned_words < 65) {
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    void* hp = nullptr;
    PWM_HIP_TRY(hipHostMalloc(&hp, 128 * 8, hipHostMallocDefault));
    ctx->pinned = static_cast<uint64_t*>(hp);
    ctx->pinned_words = 128;
  }
  uint64_t* res = na ? na_res : ctx->pinned;
  PWM_HIP_TRY(hipMemcpyAsync(res, zeros_dev, (direct ? 64 : L) * 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipMemcpyAsync(res + 64, sym_or, 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  ctx->last_ms[0] = ms;
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
  ctx->last_ms[1] = ms;
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
  ctx->last_ms[2] = ms;
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
  ctx->last_ms[3] = ms;
  if (direct && res[63])
    throw status_error{PWMHIP_ERR_TIMEOUT, "local build: a wait for another unit's counts gave up"};
  if (check_range && L < 64 && (res[64] >> L) != 0)
    throw status_error{PWMHIP_ERR_SYMBOL_RANGE, "symbol >= 2^levels"};
  // trees above one pass: matrix levels with buckets moved into numeric order
  if (tree && multi) tree_order_timed(ctx, d_text, n, width, L, ws);
  ws->n_zeros = tree ? 0 : L;
  for (uint32_t l = 0; l < L; ++l) ws->zeros[l] = tree ? 0 : res[l];
}

// ---------------------------------------------------------------------------
// batched construction: many texts, one launch per phase
// ---------------------------------------------------------------------------
namespace {

struct BatchTune {
  uint32_t grab;  // chunks per work-queue grab
  uint32_t nwin;  // gather windows per wave
};

const BatchTune kTuneCandidates[] = {{4, 8}, {16, 8}, {64, 8}, {4, 32}, {16, 32}, {64, 32}};
const BatchTune kTuneCandidatesUnits[] = {{1, 2}, {1, 4}, {1, 8}, {1, 16}};

}  // namespace

void build_batch_device(pwmhip_ctx* ctx, const void* const* d_texts, const uint64_t* sizes, uint64_t T,
                        uint32_t width, uint32_t L, bool tree, bool check_range, pwmhip_ws* outs, void* reuse,
                        uint64_t reuse_bytes, uint32_t flags) {
  struct ReuseGuard {  // an earlier block not taken over is freed on every exit
    void*& p;
    ~ReuseGuard() {
      if (p) (void)hipFree(p);
    }
  } reuse_guard{reuse};
  hipStream_t s = ctx->stream;
  const uint32_t g = L;  // one pass (L <= kMaxGroup)
  const uint32_t jmax = g - 1;
  const uint32_t nslot = jmax ? (2u << jmax) - 2u : 0;
  std::vector<uint64_t> cb(T + 1), out0(T), stride(T), wave_base(T + 1);
  std::vector<uint64_t> tptr(T), tn(T);
  // texts of 32 Ki symbols and more on average: units of 16 chunks (the
  // super-chunk build and stitch); otherwise one unit per chunk
  uint64_t total_syms = 0;
  for (uint64_t t = 0; t < T; ++t) total_syms += sizes[t];
  const bool sunits = L >= 2 && L <= 4 && T > 0 && total_syms / T >= 32768 && !small_k1_disabled() &&
                      !super_disabled();
  const uint32_t unit_waves = batch_unit_waves(T, (total_syms + kChunkSyms - 1) / kChunkSyms);
  const uint64_t unit_syms = sunits ? uint64_t(unit_waves) * kChunkSyms : kChunkSyms;
  uint64_t words = 0;
  for (uint64_t t = 0; t < T; ++t) {
    cb[t + 1] = cb[t] + (sizes[t] + unit_syms - 1) / unit_syms;
    stride[t] = level_stride(sizes[t]);
    out0[t] = words;
    words += sizes[t] ? stride[t] * L : 0;
    tptr[t] = reinterpret_cast<uint64_t>(d_texts[t]);
    tn[t] = sizes[t];
  }
  const uint64_t C = cb[T];
  // one allocation for all structures, owned by the first
  const uint64_t block_bytes = std::max<uint64_t>(words, 32) * 8;
  void* block = nullptr;
  uint64_t owned_bytes = block_bytes;
  if (reuse && reuse_bytes >= block_bytes) {
    block = reuse;
    owned_bytes = reuse_bytes;
    reuse = nullptr;
  } else {
    if (reuse) (void)hipFree(reuse);
    reuse = nullptr;
    PWM_HIP_TRY(hipMalloc(&block, block_bytes));
  }
  struct BlockGuard {  // frees the block unless the build completes
    void* p;
    pwmhip_ws* outs;
    uint64_t T;
    ~BlockGuard() {
      if (!p) return;
      (void)hipFree(p);
      for (uint64_t t = 0; t < T; ++t) std::memset(&outs[t], 0, sizeof(pwmhip_ws));
    }
  } guard{block, outs, T};
  for (uint64_t t = 0; t < T; ++t) {
    pwmhip_ws& ws = outs[t];
    ws.levels = sizes[t] ? L : 0;
    ws.text_size = sizes[t];
    ws.is_tree = tree;
    ws.d_words = static_cast<uint64_t*>(block) + out0[t];
    for (uint32_t l = 0; l < L; ++l) {
      ws.level_offset[l] = l * stride[t];
      ws.level_bits[l] = sizes[t];
    }
    ws.alloc_bytes = 0;
  }
  outs[0].d_words = static_cast<uint64_t*>(block);
  outs[0].alloc_bytes = owned_bytes;
  if (C == 0) {
    guard.p = nullptr;
    return;
  }

  // gather waves per text for a number of windows per wave
  auto waves_for_nwin = [&](uint32_t nwin) {
    const uint64_t per_wave = (sunits ? 128ull : 64ull) * nwin;
    uint64_t wv = 0;
    for (uint64_t t = 0; t < T; ++t) {
      wave_base[t] = wv;
      wv += (words_for(sizes[t]) + per_wave - 1) / per_wave;
    }
    wave_base[T] = wv;
    return wv;
  };

  const uint64_t tstride = C * (unit_syms / 64) + 64;
  const uint64_t ntiles = (C + kTile - 1) / kTile;
  const size_t sz_cnt = std::max<size_t>(1, C * nslot);
  const uint64_t wv_max = waves_for_nwin(1);  // (the most gather waves of any geometry)
  arena sizing;
  sizing.cap = ~size_t(0) >> 1;
  auto carve = [&](arena& A, std::vector<void*>& v) {
    v.clear();
    v.push_back(A.take<uint64_t>(T + 1));                 // 0 chunk_base
    v.push_back(A.take<uint64_t>(T));                     // 1 text ptr
    v.push_back(A.take<uint64_t>(T));                     // 2 n
    v.push_back(A.take<uint64_t>(T));                     // 3 out0
    v.push_back(A.take<uint64_t>(T));                     // 4 stride
    v.push_back(A.take<uint64_t>(T + 1));                 // 5 wave_base
    v.push_back(A.take<uint64_t>(T * kMaxGroup));         // 6 zeros
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, size_t(L - 1)) * tstride));  // 7 temp
    v.push_back(A.take<uint32_t>(sz_cnt));                // 8 counts
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, ntiles * nslot)));  // 9 tile sums
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, ntiles * nslot)));  // 10 tile indicator sums
    v.push_back(A.take<uint64_t>(1));                     // 11 (unused)
    v.push_back(A.take<uint64_t>(std::max<uint32_t>(nslot, 1)));          // 12 totals
    v.push_back(A.take<uint64_t>(std::max<uint32_t>(nslot, 1)));          // 13 netot
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, T * nslot)));        // 14 ttot
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, T * nslot)));        // 15 tnz
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, T * nslot)));        // 16 gstart
    v.push_back(A.take<uint64_t>(std::max<size_t>(1, T * nslot)));        // 17 nebase
    v.push_back(A.take<uint64_t>(T * 8));                 // 18 level_ne
    v.push_back(A.take<uint64_t>(sz_cnt));                // 19 tab_dst
    v.push_back(A.take<uint64_t>(sz_cnt));                // 20 tab_src
    v.push_back(A.take<unsigned long long>(1));           // 21 sym_or
    v.push_back(A.take<unsigned int>(1));                 // 22 work counter
    v.push_back(A.take<uint32_t>(std::max<size_t>(1, C))); // 23 chunk -> text
    for (int k = 0; k < 4; ++k)                            // 24..27 text bounds
      v.push_back(A.take<uint64_t>(std::max<size_t>(1, T * nslot)));
    v.push_back(A.take<UnitMeta>(sunits ? std::max<size_t>(1, C) : 1));  // 28 unit metadata
    v.push_back(A.take<uint32_t>(std::max<uint64_t>(1, wv_max)));         // 29 text of every gather wave
  };
  std::vector<void*> b;
  carve(sizing, b);
  ctx->scratch.reset();
  ctx->scratch.ensure(sizing.used);
  carve(ctx->scratch, b);
  auto U64 = [&](int i) { return static_cast<uint64_t*>(b[i]); };
  const uint8_t* lut = compress_table(ctx);  // before any capture

  // the six metadata arrays are consecutive in the arena: one pinned copy
  // with the same layout (zero counts come back through the same buffer)
  const size_t span = size_t(static_cast<char*>(b[5]) - static_cast<char*>(b[0])) + (T + 1) * 8;
  const size_t zbytes = T * kMaxGroup * 8;
  if (ctx->hstage_bytes < std::max(span, zbytes + 64)) {
    if (ctx->hstage) (void)hipHostFree(ctx->hstage);
    ctx->hstage = nullptr;
    ctx->hstage_bytes = 0;
    void* hp = nullptr;
    const size_t want = std::max(span, zbytes + 64) * 2;
    PWM_HIP_TRY(hipHostMalloc(&hp, want, hipHostMallocDefault));
    ctx->hstage = hp;
    ctx->hstage_bytes = want;
  }
  char* hs = static_cast<char*>(ctx->hstage);
  auto put = [&](int k, const std::vector<uint64_t>& v, size_t cnt) {
    std::memcpy(hs + (static_cast<char*>(b[k]) - static_cast<char*>(b[0])), v.data(), cnt * 8);
  };
  // metadata for a gather geometry; the previous use of the staging buffer
  // has completed (every upload is followed by a synchronisation)
  auto upload = [&](uint32_t nwin) {
    const uint64_t wv = waves_for_nwin(nwin);
    put(0, cb, T + 1);
    put(1, tptr, T);
    put(2, tn, T);
    put(3, out0, T);
    put(4, stride, T);
    put(5, wave_base, T + 1);
    PWM_HIP_TRY(hipMemcpyAsync(b[0], hs, span, hipMemcpyHostToDevice, s));
    return wv;
  };

  // the device sequence: clears (kernels, so that it can be captured), local
  // build, offsets, gathers
  auto enqueue = [&](const BatchTune& tu, uint64_t wv, bool events) {
    const uint64_t nz = T * kMaxGroup;
    hipLaunchKernelGGL(k_fill_u64, dim3(uint32_t((nz + 255) / 256)), dim3(256), 0, s, U64(6), nz, uint64_t(0));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, U64(21), uint64_t(1), uint64_t(0));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, static_cast<unsigned int*>(b[22]), 1u, 0u);
    if (events) PWM_HIP_TRY(hipEventRecord(ctx->ev[0], s));
    K1Params k1{};
    k1.width = width;
    k1.n = 0;
    k1.L = L;
    k1.l0 = 0;
    k1.g = g;
    k1.np = L;
    k1.out_level0 = static_cast<uint64_t*>(block);
    k1.temp_levels = U64(7);
    k1.temp_payload = U64(7);
    k1.temp_stride = tstride;
    k1.counts = static_cast<uint32_t*>(b[8]);
    k1.nslot = nslot;
    k1.sym_or = static_cast<unsigned long long*>(b[21]);
    k1.nchunks = C;
    k1.lut = lut;
    k1.b_chunk_base = U64(0);
    k1.b_text = U64(1);
    k1.b_n = U64(2);
    k1.b_out0 = U64(3);
    k1.b_zeros = static_cast<unsigned long long*>(b[6]);
    k1.ntexts = uint32_t(T);
    k1.work = static_cast<unsigned int*>(b[22]);
    k1.grab = tu.grab;
    k1.b_chunk_text = static_cast<uint32_t*>(b[23]);
    if (sunits) {
      hipLaunchKernelGGL(k_unit_meta, dim3(uint32_t((T + 3) / 4)), dim3(256), 0, s, U64(0), U64(1), U64(2),
                         U64(3), uint32_t(T), width, unit_syms, static_cast<UnitMeta*>(b[28]),
                         static_cast<uint32_t*>(b[23]));
      launch_super_batch(ctx, k1, L, static_cast<const UnitMeta*>(b[28]), unit_waves);
    } else {
      dispatch_k1(ctx, k1);
    }
    if (events) PWM_HIP_TRY(hipEventRecord(ctx->ev[1], s));
    if (nslot) {
      const uint32_t* counts = static_cast<const uint32_t*>(b[8]);
      // tile sums, their scans, the text bounds, per-text bucket starts, then
      // the run tables (counts read twice, no per-chunk prefix arrays)
      const uint32_t* ctext = static_cast<const uint32_t*>(b[23]);
      hipLaunchKernelGGL(k_tile_sums2, dim3(uint32_t(ntiles), nslot), dim3(256), 0, s, counts, U64(9), U64(10), C,
                         nslot, ntiles);
      hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, U64(9), U64(12), ntiles);
      hipLaunchKernelGGL(k_tile_scan, dim3(nslot), dim3(1024), 0, s, U64(10), U64(13), ntiles);
      hipLaunchKernelGGL(k_batch_text_bounds, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, counts, U64(9),
                         U64(10), ctext, U64(0), U64(24), U64(25), U64(26), U64(27), C, nslot, ntiles);
      const uint64_t tsl = T * nslot;
      hipLaunchKernelGGL(k_batch_text_totals, dim3(uint32_t((tsl + 255) / 256)), dim3(256), 0, s, U64(24),
                         U64(25), U64(26), U64(27), U64(0), U64(14), U64(15), uint32_t(T), nslot);
      hipLaunchKernelGGL(k_bucket_starts_batch, dim3(uint32_t(T), jmax), dim3(64), 0, s, U64(14), U64(15),
                         U64(16), U64(17), U64(18), nslot, tree ? 1 : 0);
      hipLaunchKernelGGL(k_batch_apply_runs, dim3(uint32_t(ntiles), nslot), dim3(1024), 0, s, counts, U64(9),
                         U64(10), ctext, U64(0), U64(24), U64(25), U64(16), U64(17), U64(19), U64(20), C, nslot,
                         ntiles, unit_syms);
      if (events) PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
      hipLaunchKernelGGL(k_wave_text, dim3(uint32_t((T + 255) / 256)), dim3(256), 0, s, U64(5), uint32_t(T),
                         static_cast<uint32_t*>(b[29]));
      for (uint32_t j = 1; j <= jmax; ++j) {
        BatchGatherParams bp{};
        bp.wave_base = U64(5);
        bp.wave_text = static_cast<const uint32_t*>(b[29]);
        bp.chunk_base = U64(0);
        bp.text_n = U64(2);
        bp.out0 = U64(3);
        bp.stride = U64(4);
        bp.level_ne = U64(18);
        bp.tab_dst = U64(19);
        bp.tab_src = U64(20);
        bp.src = U64(7);
        bp.src_stride = tstride;
        bp.out = static_cast<uint64_t*>(block);
        bp.ntexts = uint32_t(T);
        bp.j = j;
        bp.nchunks = C;
        bp.nwin = tu.nwin;
        if (sunits) hipLaunchKernelGGL(k_gather_batch_wide, dim3(uint32_t((wv + 3) / 4)), dim3(256), 0, s, bp);
        else hipLaunchKernelGGL(k_gather_batch, dim3(uint32_t((wv + 3) / 4)), dim3(256), 0, s, bp);
      }
    } else if (events) {
      PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
    }
    PWM_HIP_TRY(hipGetLastError());
    if (events) PWM_HIP_TRY(hipEventRecord(ctx->ev[3], s));
  };

  // tunables of this length / level bucket (mean text length, rounded to a
  // power of two, and the number of levels)
  uint32_t lg = 0;
  for (uint64_t m = std::max<uint64_t>(1, words_for(total_syms) / T); m > 1; m >>= 1) ++lg;
  const uint64_t bucket =
      (uint64_t(L) << 8) | lg | (sunits ? 1ull << 16 : 0ull) | (sunits && unit_waves == 8 ? 1ull << 17 : 0ull);
  // super-chunk units: one unit per grab, `nwin` = 128-word steps per wave
  BatchTune tu = sunits ? BatchTune{1, 8} : BatchTune{16, kGatherWindows};
  auto it = ctx->tune.find(bucket);
  if (it != ctx->tune.end()) {
    tu = BatchTune{it->second.first, it->second.second};
  } else if (flags & PWMHIP_FLAG_AUTOTUNE) {
    float best = 0;
    const BatchTune* cands = sunits ? kTuneCandidatesUnits : kTuneCandidates;
    const size_t ncand = sunits ? sizeof(kTuneCandidatesUnits) / sizeof(BatchTune)
                                : sizeof(kTuneCandidates) / sizeof(BatchTune);
    for (size_t ci = 0; ci < ncand; ++ci) {
      const BatchTune& cand = cands[ci];
      const uint64_t wv = upload(cand.nwin);
      enqueue(cand, wv, true);
      PWM_HIP_TRY(hipStreamSynchronize(s));
      float ms = 0;
      PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
      if (best == 0 || ms < best) {
        best = ms;
        tu = cand;
      }
    }
    ctx->tune[bucket] = {tu.grab, tu.nwin};
  }

  const uint64_t wv = upload(tu.nwin);
  bool timed = true;
  if (flags & PWMHIP_FLAG_GRAPH) {
    // replay when the geometry and every address the kernels see repeat
    const uint64_t key[12] = {T,
                              C,
                              wv,
                              width,
                              L,
                              uint64_t(tree),
                              reinterpret_cast<uint64_t>(block),
                              reinterpret_cast<uint64_t>(ctx->scratch.base),
                              sizing.used,
                              tu.grab,
                              tu.nwin,
                              reinterpret_cast<uint64_t>(lut)};
    if (!ctx->gexec || std::memcmp(key, ctx->gkey, sizeof(key)) != 0) {
      if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
      if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
      ctx->gexec = nullptr;
      ctx->graph = nullptr;
      PWM_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(tu, wv, false);
      } catch (...) {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(s, &broken);
        if (broken) (void)hipGraphDestroy(broken);
        throw;
      }
      PWM_HIP_TRY(hipStreamEndCapture(s, &ctx->graph));
      PWM_HIP_TRY(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
      std::memcpy(ctx->gkey, key, sizeof(key));
    }
    PWM_HIP_TRY(hipEventRecord(ctx->ev[0], s));
    PWM_HIP_TRY(hipGraphLaunch(ctx->gexec, s));
    PWM_HIP_TRY(hipEventRecord(ctx->ev[3], s));
    timed = false;  // phases are not separated inside the graph
  } else {
    enqueue(tu, wv, true);
  }
  // (the metadata upload above is ordered before these downloads on the stream)
  uint64_t* zeros = reinterpret_cast<uint64_t*>(hs);
  PWM_HIP_TRY(hipMemcpyAsync(zeros, b[6], zbytes, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipMemcpyAsync(zeros + T * kMaxGroup, b[21], 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipStreamSynchronize(s));
  const uint64_t sym_or = zeros[T * kMaxGroup];
  float ms = 0;
  for (int i = 0; i < 3; ++i) {
    ms = 0;
    if (timed) PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
    ctx->last_ms[i] = ms;
  }
  PWM_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
  ctx->last_ms[3] = ms;
  if (check_range && L < 64 && (sym_or >> L) != 0)
    throw status_error{PWMHIP_ERR_SYMBOL_RANGE, "symbol >= 2^levels"};
  for (uint64_t t = 0; t < T; ++t) {
    outs[t].n_zeros = (tree || !sizes[t]) ? 0 : L;
    for (uint32_t l = 0; l < L; ++l) outs[t].zeros[l] = (tree || !sizes[t]) ? 0 : zeros[t * kMaxGroup + l];
  }
  guard.p = nullptr;
}

}  // namespace pwm

// ---------------------------------------------------------------------------
// The allocation-free build (pwmhip_build_noalloc): plain builds of 2 to 4
// levels from a device text into caller-provided output memory, with every
// temporary in one static device array.  Layout of the array: the 64 KiB
// compress table, then the build's temporaries (up to kNoAllocTempBytes).
// Each device that loads the library holds the array until the process ends.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocLutBytes = size_t(1) << 16;
constexpr size_t kNoAllocTempBytes = size_t(128) << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint8_t scratch[kNoAllocLutBytes + kNoAllocTempBytes];

namespace pwm {
namespace {
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

void noalloc_setup(pwmhip_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices) return;  // (the entry point reports it)
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (g_noalloc_ptr[ctx->device]) return;
  // with the context's device current: each device has its own copy of the
  // array at its own address
  void* p = nullptr;
  PWM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  hipLaunchKernelGGL(k_make_lut, dim3(256), dim3(256), 0, ctx->stream, static_cast<uint8_t*>(p));
  PWM_HIP_TRY(hipGetLastError());
  PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  g_noalloc_ptr[ctx->device] = p;
}

}  // namespace pwm

extern "C" int pwmhip_build_noalloc(pwmhip_ctx* ctx, const void* d_text, uint64_t size, uint32_t word_width,
                                    uint32_t levels, uint32_t algo, uint32_t flags, uint64_t* d_out,
                                    uint64_t out_words, pwmhip_ws* out) {
  if (!ctx || !out) return PWMHIP_ERR_INVALID;
  if (word_width != 1 && word_width != 2 && word_width != 4 && word_width != 8) return PWMHIP_ERR_INVALID;
  if (algo > PWMHIP_ALGO_NAIVE) return PWMHIP_ERR_UNSUPPORTED;  // plain shapes only
  if (levels < 2 || levels > 4 || levels > 8 * word_width) return PWMHIP_ERR_UNSUPPORTED;
  if (flags & ~uint32_t(PWMHIP_FLAG_TREE | PWMHIP_FLAG_CHECK_RANGE | PWMHIP_FLAG_TEXT_ON_DEVICE))
    return PWMHIP_ERR_INVALID;
  if (size && (!d_text || !d_out)) return PWMHIP_ERR_INVALID;
  const bool tree = (flags & PWMHIP_FLAG_TREE) != 0;
  pwmhip_ws ws{};
  ws.is_tree = tree;
  ws.text_size = size;
  if (size == 0) {  // wx_ppc.hpp:39-45: empty structure
    *out = ws;
    return PWMHIP_OK;
  }
  const uint64_t bits[4] = {size, size, size, size};
  // the layout of pwmhip_build's output, in the caller's memory (not owned:
  // alloc_bytes stays 0, pwmhip_ws_free frees nothing of it)
  if (pwm::layout_structure(&ws, levels, bits) > out_words) return PWMHIP_ERR_INVALID;
  ws.d_words = d_out;
  const int st = pwm::guarded([&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    void* base = ctx->device < kNoAllocMaxDevices ? pwm::g_noalloc_ptr[ctx->device] : nullptr;
    if (!base) throw pwm::status_error{PWMHIP_ERR_UNSUPPORTED, "no static scratch on this device"};
    const pwm::NoAlloc na{static_cast<uint8_t*>(base) + kNoAllocLutBytes, kNoAllocTempBytes,
                          static_cast<const uint8_t*>(base)};
    pwm::build_plain_device(ctx, d_text, size, word_width, levels, tree, (flags & PWMHIP_FLAG_CHECK_RANGE) != 0,
                            &ws, nullptr, &na);
    if (!tree && pwm::zeros_last_is_zero(algo)) ws.zeros[levels - 1] = 0;
  });
  if (st != PWMHIP_OK) return st;
  *out = ws;
  return PWMHIP_OK;
}
