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
//  K2  k_scan_slots / k_bucket_starts / k_zero_edges: per (level, bucket)
//      exclusive scan over the chunks, global bucket starts (matrix: bucket
//      index order = bit-reversed prefix, permutation.hpp:36-51; tree:
//      numeric prefix order), and zeroing of destination words that several
//      runs share.
//  K3  k_stitch: every (chunk, level, bucket) run of bits is copied to its
//      global position (the copy_bits analogue, merge.hpp:19-190); covered
//      words are stored, shared edge words are OR-ed atomically.
//
// More than kMaxGroup levels are built in passes of up to kMaxGroup levels;
// the partitioned remaining planes of one pass are the next pass's input.
// Trees with more than kMaxGroup levels are built as matrices and then have
// each level's buckets moved into numeric order (bitops.hip).

#include "pwm_internal.hpp"

#include <algorithm>

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
  __syncthreads();
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
  __syncthreads();
}

// rank1 in an LDS plane with per-word exclusive one-counts `ooff`.
__device__ __forceinline__ uint32_t rank1_lds(const uint64_t* plane, const uint32_t* ooff, uint32_t pos) {
  const uint32_t w = pos >> 6, r = pos & 63;
  if (r == 0) return ooff[w];
  return ooff[w] + __popcll(plane[w] & ((1ull << r) - 1));
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
  uint32_t* ztot;            // ztot[c*g + j]
  unsigned long long* sym_or;
  uint64_t nchunks;
};

// Gather bit s (0 = least significant) of 4 packed bytes into 4 bits.
__device__ __forceinline__ uint32_t gather4(uint32_t x4, uint32_t s) {
  return (((x4 >> s) & 0x01010101u) * 0x10204080u) >> 28;
}

template <int NP, typename T>
__device__ __forceinline__ void load_planes_text(const K1Params& p, uint64_t e0, int cnt,
                                                 uint64_t* acc, uint64_t& my_or) {
  // symbol bit of plane k: s_k = L-1-l0-k ; byte index s_k>>3, bit s_k&7
  const T* tx = static_cast<const T*>(p.text);
#pragma unroll 1
  for (int q = 0; q < 16; ++q) {
    uint64_t v[4];
    if (cnt >= 4 * q + 4) {
      if constexpr (sizeof(T) == 1) {
        const uint32_t x = *reinterpret_cast<const uint32_t*>(tx + e0 + 4 * q);
        v[0] = x & 0xff; v[1] = (x >> 8) & 0xff; v[2] = (x >> 16) & 0xff; v[3] = x >> 24;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = tx[e0 + 4 * q + e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = (4 * q + e < cnt) ? uint64_t(tx[e0 + 4 * q + e]) : 0;
    }
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
  }
}

template <int NP>
__global__ __launch_bounds__(P) void k_local_build(K1Params p) {
  __shared__ uint8_t lut[65536];
  __shared__ uint64_t buf[2][NP][P];
  __shared__ uint32_t ooff[P + 1];
  __shared__ uint32_t bst[2][(1 << kMaxGroup) + 1];
  __shared__ uint32_t scr[2 * P / 64];

  const int t = threadIdx.x;
  for (int i = t; i < 65536; i += P) lut[i] = uint8_t(compress8(uint32_t(i) & 255u, uint32_t(i) >> 8));
  uint64_t my_or = 0;

  for (uint64_t c = blockIdx.x; c < p.nchunks; c += gridDim.x) {
    __syncthreads();
    const uint64_t base = c * kChunkSyms;
    const uint32_t Ec = uint32_t(min<uint64_t>(kChunkSyms, p.n - base));
    const uint32_t nw = (Ec + 63) >> 6;
    const int cnt = max(0, min(64, int(Ec) - 64 * t));

    uint64_t acc[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) acc[k] = 0;
    if (p.planes_in) {
      if (t < int(nw)) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
          if (k < int(p.np)) acc[k] = brev64(p.planes_in[k * p.in_stride + c * P + t]);
      }
    } else if (cnt > 0) {
      const uint64_t e0 = base + 64ull * t;
      switch (p.width) {
        case 1: load_planes_text<NP, uint8_t>(p, e0, cnt, acc, my_or); break;
        case 2: load_planes_text<NP, uint16_t>(p, e0, cnt, acc, my_or); break;
        case 4: load_planes_text<NP, uint32_t>(p, e0, cnt, acc, my_or); break;
        default: load_planes_text<NP, uint64_t>(p, e0, cnt, acc, my_or); break;
      }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (k < int(p.np)) buf[0][k][t] = acc[k];
    const uint64_t vm = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    if (t == 0) {
      bst[0][0] = 0;
      bst[0][1] = Ec;
    }
    __syncthreads();

    int cur = 0;
    uint32_t nb = 1;
    for (uint32_t j = 0; j < p.g; ++j) {
      const uint64_t m = buf[cur][j][t];
      if (t < int(nw)) {
        if (j == 0) p.out_level0[c * P + t] = brev64(m);
        else p.temp_levels[(j - 1) * p.temp_stride + c * P + t] = brev64(m);
      }
      const uint64_t m1 = m & vm, m0 = ~m & vm;
      const uint32_t o = __popcll(m1), z = __popcll(m0);
      uint32_t zoff, oof, Zt, Ot;
      block_exscan2<P>(z, o, zoff, oof, Zt, Ot, scr);
      if (t == 0) p.ztot[c * p.g + j] = Zt;
      const bool more = (j + 1 < p.g) || (p.np > p.g);
      if (!more) break;
      ooff[t] = oof;
      if (t == P - 1) ooff[P] = Ot;
      const int nxt = cur ^ 1;
#pragma unroll
      for (int k = 1; k < NP; ++k)
        if (k > int(j) && k < int(p.np)) buf[nxt][k][t] = 0;
      __syncthreads();
#pragma unroll
      for (int k = 1; k < NP; ++k) {
        if (k > int(j) && k < int(p.np)) {
          const uint64_t x = buf[cur][k][t];
          or_bits(buf[nxt][k], zoff, compress64(x, m0, lut), z);
          or_bits(buf[nxt][k], Zt + oof, compress64(x, m1, lut), o);
        }
      }
      // bucket counts of sub-level j+1: zeros and ones of plane j inside
      // every current bucket; new order = zeros of all buckets, then ones.
      uint32_t zc = 0, oc = 0;
      if (t < int(nb)) {
        const uint32_t s = bst[cur][t], e = bst[cur][t + 1];
        oc = rank1_lds(buf[cur][j], ooff, e) - rank1_lds(buf[cur][j], ooff, s);
        zc = (e - s) - oc;
        const uint32_t slot = (2u << j) - 2u;
        p.counts[c * p.nslot + slot + t] = zc;
        p.counts[c * p.nslot + slot + nb + t] = oc;
      }
      uint32_t ez, eo, tz, to;
      block_exscan2<P>(zc, oc, ez, eo, tz, to, scr);
      if (t < int(nb)) {
        bst[nxt][t] = ez;
        bst[nxt][nb + t] = tz + eo;
      }
      if (t == 0) bst[nxt][2 * nb] = Ec;
      __syncthreads();
      cur = nxt;
      nb <<= 1;
    }
    if (t < int(nw)) {
#pragma unroll
      for (int k = 0; k < NP; ++k)
        if (k >= int(p.g) && k < int(p.np))
          p.temp_payload[(k - p.g) * p.temp_stride + c * P + t] = brev64(buf[cur][k][t]);
    }
  }
  if (!p.planes_in) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) my_or |= __shfl_xor(my_or, d, 64);
    if ((t & 63) == 0 && my_or) atomicOr(p.sym_or, (unsigned long long)my_or);
  }
}

// ---- K2 --------------------------------------------------------------------
// Exclusive scan over chunks of every slot's count; one block per slot.
__global__ __launch_bounds__(1024) void k_scan_slots(const uint32_t* counts, uint64_t* pref,
                                                     uint64_t* totals, uint64_t nchunks,
                                                     uint32_t nslot) {
  __shared__ uint64_t wsum[16];
  const uint32_t slot = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  uint64_t carry = 0;
  for (uint64_t b = 0; b < nchunks; b += 1024) {
    const uint64_t c = b + t;
    const uint64_t v = c < nchunks ? counts[c * nslot + slot] : 0;
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
      const uint64_t x = wsum[w];
      if (w < wid) off += x;
      tot += x;
    }
    if (c < nchunks) pref[c * nslot + slot] = carry + off + inc - v;
    carry += tot;
    __syncthreads();
  }
  if (t == 0) totals[slot] = carry;
}

// Sum of per-chunk zero counts for every sub-level; one block per sub-level.
__global__ __launch_bounds__(1024) void k_sum_zeros(const uint32_t* ztot, uint64_t nchunks, uint32_t g,
                                                    uint64_t* zeros_out) {
  __shared__ uint64_t wsum[16];
  const uint32_t j = blockIdx.x;
  uint64_t s = 0;
  for (uint64_t c = threadIdx.x; c < nchunks; c += 1024) s += ztot[c * g + j];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t tot = 0;
    for (int w = 0; w < 16; ++w) tot += wsum[w];
    zeros_out[j] = tot;
  }
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
                                                       int tree) {
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

struct StitchParams {
  const uint32_t* counts;
  const uint64_t* pref;
  const uint64_t* gstart;
  uint64_t* dst_level[kMaxGroup];    // sub-level j -> destination level (j >= 1)
  const uint64_t* src_temp;          // chunk-local sub-level j at src_temp + (j-1)*stride
  uint64_t temp_stride;
  uint64_t* dst_payload;             // payload plane k at dst_payload + k*dst_pay_stride
  uint64_t dst_pay_stride;
  const uint64_t* src_payload;       // chunk-local payload k at src_payload + k*temp_stride
  uint32_t npay, g, jmax, nslot;
  uint64_t n, nchunks;
};

__device__ __forceinline__ uint32_t slot_level(uint32_t slot) {
  // slot = 2^j - 2 + b  ->  j
  return 31 - __clz(slot + 2);
}

// Zero the destination words that a run shares with other runs.
__global__ void k_zero_edges(StitchParams p) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= p.nchunks * p.nslot) return;
  const uint64_t c = i / p.nslot;
  const uint32_t slot = uint32_t(i % p.nslot);
  const uint64_t len = p.counts[i];
  if (!len) return;
  const uint32_t j = slot_level(slot);
  const uint64_t dst = p.gstart[slot] + p.pref[i];
  const uint64_t end = dst + len;
  const bool head = (dst & 63) != 0;
  const bool tail = (end & 63) != 0 && end != p.n;
  if (!head && !tail) return;
  (void)c;
  if (j < p.g) {
    uint64_t* d = p.dst_level[j];
    if (head) d[dst >> 6] = 0;
    if (tail) d[(end - 1) >> 6] = 0;
  } else {
    for (uint32_t k = 0; k < p.npay; ++k) {
      uint64_t* d = p.dst_payload + k * p.dst_pay_stride;
      if (head) d[dst >> 6] = 0;
      if (tail) d[(end - 1) >> 6] = 0;
    }
  }
}

// Read 64 MSB-first bits starting at (possibly negative) bit position s.
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

// Copy one destination word of a run.
__device__ __forceinline__ void stitch_word(uint64_t* dst, const uint64_t* src, uint64_t dbit,
                                            uint64_t sbit, uint64_t len, uint64_t w, uint64_t n) {
  const uint64_t ws = w << 6;
  const uint64_t a = max(ws, dbit), b = min(ws + 64, dbit + len);
  const uint64_t v = window64(src, int64_t(sbit) - int64_t(dbit - ws)) &
                     range_mask(uint32_t(a - ws), uint32_t(b - ws));
  const uint64_t wend = min(ws + 64, n);
  if (a == ws && b == wend) dst[w] = v;
  else if (v) atomicOr(reinterpret_cast<unsigned long long*>(&dst[w]), (unsigned long long)v);
}

// One block per chunk: runs of the chunk flattened over destination words.
__global__ __launch_bounds__(256) void k_stitch(StitchParams p) {
  constexpr int MAXS = (2 << kMaxGroup) - 2;
  __shared__ uint64_t s_src[MAXS], s_dst[MAXS], s_len[MAXS];
  __shared__ uint64_t s_wpre[MAXS + 1];
  __shared__ uint32_t s_mult[MAXS];
  const uint64_t c = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t ns = p.nslot;
  // local starts: exclusive scan of counts within each sub-level (tiny)
  for (uint32_t slot = t; slot < ns; slot += blockDim.x) {
    const uint32_t j = slot_level(slot);
    const uint32_t s0 = (1u << j) - 2;
    uint64_t loc = 0;
    for (uint32_t b = s0; b < slot; ++b) loc += p.counts[c * ns + b];
    const uint64_t len = p.counts[c * ns + slot];
    const uint64_t dst = p.gstart[slot] + p.pref[c * ns + slot];
    s_src[slot] = c * kChunkSyms + loc;
    s_dst[slot] = dst;
    s_len[slot] = len;
    s_mult[slot] = j < p.g ? 1u : p.npay;
  }
  __syncthreads();
  if (t == 0) {
    uint64_t run = 0;
    for (uint32_t s = 0; s < ns; ++s) {
      s_wpre[s] = run;
      const uint64_t len = s_len[s];
      const uint64_t nwd = len ? ((s_dst[s] + len - 1) >> 6) - (s_dst[s] >> 6) + 1 : 0;
      run += nwd * s_mult[s];
    }
    s_wpre[ns] = run;
  }
  __syncthreads();
  const uint64_t total = s_wpre[ns];
  for (uint64_t i = t; i < total; i += blockDim.x) {
    // find slot: largest s with s_wpre[s] <= i
    uint32_t lo = 0, hi = ns;  // invariant s_wpre[lo] <= i < s_wpre[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_wpre[mid] <= i) lo = mid;
      else hi = mid;
    }
    const uint32_t s = lo;
    const uint64_t len = s_len[s], dbit = s_dst[s], sbit = s_src[s];
    const uint64_t nwd = ((dbit + len - 1) >> 6) - (dbit >> 6) + 1;
    const uint64_t r = i - s_wpre[s];
    const uint64_t k = r / nwd, wi = r % nwd;
    const uint64_t w = (dbit >> 6) + wi;
    const uint32_t j = slot_level(s);
    if (j < p.g) {
      stitch_word(p.dst_level[j], p.src_temp + (j - 1) * p.temp_stride, dbit, sbit, len, w, p.n);
    } else {
      stitch_word(p.dst_payload + k * p.dst_pay_stride, p.src_payload + k * p.temp_stride, dbit, sbit,
                  len, w, p.n);
    }
  }
}

}  // namespace

void launch_scan_slots(hipStream_t s, const uint32_t* counts, uint64_t* pref, uint64_t* totals,
                       uint64_t nchunks, uint32_t nslot) {
  hipLaunchKernelGGL(k_scan_slots, dim3(nslot), dim3(1024), 0, s, counts, pref, totals, nchunks, nslot);
  PWM_HIP_TRY(hipGetLastError());
}

namespace {

template <int NP>
void launch_k1(pwmhip_ctx* ctx, const K1Params& prm) {
  const int per_cu = (NP <= 4) ? 2 : 1;
  const uint64_t grid = std::min<uint64_t>(prm.nchunks, uint64_t(ctx->num_cus) * per_cu);
  hipLaunchKernelGGL(k_local_build<NP>, dim3(uint32_t(grid)), dim3(P), 0, ctx->stream, prm);
  PWM_HIP_TRY(hipGetLastError());
}

void dispatch_k1(pwmhip_ctx* ctx, const K1Params& prm) {
  const uint32_t np = prm.np;
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
void alloc_structure(pwmhip_ws* ws, uint64_t levels, const uint64_t* bits, uint64_t* reuse,
                     uint64_t reuse_bytes) {
  uint64_t off = 0;
  for (uint64_t l = 0; l < levels; ++l) {
    ws->level_offset[l] = off;
    ws->level_bits[l] = bits[l];
    off += level_stride(bits[l]);
  }
  ws->levels = levels;
  ws->alloc_bytes = std::max<uint64_t>(off, 32) * 8;
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

void build_plain_device(pwmhip_ctx* ctx, const void* d_text, uint64_t n, uint32_t width,
                        uint32_t L, bool tree, bool check_range, pwmhip_ws* ws) {
  hipStream_t s = ctx->stream;
  const uint64_t nchunks = (n + kChunkSyms - 1) / kChunkSyms;
  const uint64_t stride = level_stride(n);
  const bool multi = L > uint32_t(kMaxGroup);
  const bool tree_direct = tree && !multi;

  // group plan
  std::vector<uint32_t> groups;
  for (uint32_t l = 0; l < L; l += kMaxGroup) groups.push_back(std::min<uint32_t>(kMaxGroup, L - l));

  // scratch sizing
  const uint32_t gmax = groups[0];
  const uint32_t npay0 = L - gmax;
  const uint32_t nslot_max = (2u << gmax) - 2u;
  // Every scratch buffer is taken twice: once to size the arena, once for real.
  const size_t sz_temp = size_t(std::max<uint32_t>(1, gmax - 1)) * stride;
  const size_t sz_pay = std::max<size_t>(1, size_t(npay0) * stride);
  const size_t sz_pong = std::max<size_t>(1, size_t(npay0 > kMaxGroup ? npay0 - kMaxGroup : 0) * stride);
  const size_t sz_cnt = std::max<size_t>(1, nchunks * nslot_max);
  auto carve = [&](arena& A, uint64_t*& temp_levels, uint64_t*& temp_payload, uint64_t*& S_ping,
                   uint64_t*& S_pong, uint32_t*& counts, uint64_t*& pref, uint32_t*& ztot, uint64_t*& totals,
                   uint64_t*& gstart, uint64_t*& zeros_dev, unsigned long long*& sym_or) {
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
  };
  uint64_t *temp_levels, *temp_payload, *S_ping, *S_pong, *pref, *totals, *gstart, *zeros_dev;
  uint32_t *counts, *ztot;
  unsigned long long* sym_or;
  {
    arena sizing;  // dry run on an unbounded arena to measure
    sizing.cap = ~size_t(0) >> 1;
    carve(sizing, temp_levels, temp_payload, S_ping, S_pong, counts, pref, ztot, totals, gstart, zeros_dev,
          sym_or);
    ctx->scratch.reset();
    ctx->scratch.ensure(sizing.used);
  }
  carve(ctx->scratch, temp_levels, temp_payload, S_ping, S_pong, counts, pref, ztot, totals, gstart, zeros_dev,
        sym_or);
  PWM_HIP_TRY(hipMemsetAsync(sym_or, 0, 8, s));

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
    k1.temp_stride = stride;
    k1.counts = counts;
    k1.nslot = nslot;
    k1.ztot = ztot;
    k1.sym_or = sym_or;
    k1.nchunks = nchunks;
    dispatch_k1(ctx, k1);
    if (gi == 0) PWM_HIP_TRY(hipEventRecord(ctx->ev[1], s));

    hipLaunchKernelGGL(k_sum_zeros, dim3(g), dim3(1024), 0, s, ztot, nchunks, g, zeros_dev + l0);
    if (nslot) {
      hipLaunchKernelGGL(k_scan_slots, dim3(nslot), dim3(1024), 0, s, counts, pref, totals, nchunks, nslot);
      hipLaunchKernelGGL(k_bucket_starts, dim3(jmax), dim3(256), 0, s, totals, gstart,
                         tree_direct ? 1 : 0);
      StitchParams sp{};
      sp.counts = counts;
      sp.pref = pref;
      sp.gstart = gstart;
      for (uint32_t j = 1; j < g; ++j) sp.dst_level[j] = ws->d_words + ws->level_offset[l0 + j];
      sp.src_temp = temp_levels;
      sp.temp_stride = stride;
      sp.dst_payload = S_out;
      sp.dst_pay_stride = stride;
      sp.src_payload = temp_payload;
      sp.npay = npay;
      sp.g = g;
      sp.jmax = jmax;
      sp.nslot = nslot;
      sp.n = n;
      sp.nchunks = nchunks;
      const uint64_t nthreads = nchunks * nslot;
      hipLaunchKernelGGL(k_zero_edges, dim3(uint32_t((nthreads + 255) / 256)), dim3(256), 0, s, sp);
      if (gi == 0) PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
      hipLaunchKernelGGL(k_stitch, dim3(uint32_t(nchunks)), dim3(256), 0, s, sp);
    } else if (gi == 0) {
      PWM_HIP_TRY(hipEventRecord(ctx->ev[2], s));
    }
    PWM_HIP_TRY(hipGetLastError());
    l0 += g;
    S_in = S_out;
    std::swap(S_out, S_next);
  }
  PWM_HIP_TRY(hipEventRecord(ctx->ev[3], s));

  // results: zeros per level and the symbol OR
  if (ctx->pinned_words < 65) {
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    void* hp = nullptr;
    PWM_HIP_TRY(hipHostMalloc(&hp, 128 * 8, hipHostMallocDefault));
    ctx->pinned = static_cast<uint64_t*>(hp);
    ctx->pinned_words = 128;
  }
  PWM_HIP_TRY(hipMemcpyAsync(ctx->pinned, zeros_dev, L * 8, hipMemcpyDeviceToHost, s));
  PWM_HIP_TRY(hipMemcpyAsync(ctx->pinned + 64, sym_or, 8, hipMemcpyDeviceToHost, s));
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
  if (check_range && L < 64 && (ctx->pinned[64] >> L) != 0)
    throw status_error{PWMHIP_ERR_SYMBOL_RANGE, "symbol >= 2^levels"};
  // trees above one pass: matrix levels with buckets moved into numeric order
  if (tree && multi) wm_to_wt_device(ctx, d_text, n, width, L, ws);
  ws->n_zeros = tree ? 0 : L;
  for (uint32_t l = 0; l < L; ++l) ws->zeros[l] = tree ? 0 : ctx->pinned[l];
}

}  // namespace pwm
