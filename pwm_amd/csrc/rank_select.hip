// pwm_amd/csrc/rank_select.hip -- rank / select directories over the levels
// of a built structure and batched access / rank / select queries.
//
// Replaces query_support<> (queries/query_support.hpp:18-186) with
// bin_rank_popcnt (bin_rank_popcnt.hpp:32-134) and bin_select_popcnt
// (bin_select_popcnt.hpp:32-252).  Same semantics (rank1(i) = ones in [0, i),
// select1(k) = 0-based position of the k-th one, select(0) = 0), different
// layout, valid for levels below 2^48 bits (the reference's uint32 L1 counts
// and 2^32-1-bit L0 blocks break at 2^32 bits, which every level of a 4 Gi-
// symbol text reaches):
//   superblock = 64 Ki bits : absolute uint64 count of ones before it
//   block      = 512 bits   : uint16 count of ones in the superblock before it
//                             (one 64-byte line of level words)
//   samples    : superblock of every 8192nd one / zero (narrows select's
//                binary search)
// Queries run one lane per query; a symbol query walks the levels with the
// matrix / tree recursions of query_support.hpp:72-178.

#include "pwm_internal.hpp"

#include <algorithm>
#include <cstring>
#include <new>

struct pwmhip_rs {
  uint64_t levels = 0;
  uint64_t n = 0;
  int is_tree = 0;
  void* mem = nullptr;          // one device allocation
  void* d_desc = nullptr;       // device array of level descriptors
  uint64_t true_zeros[PWMHIP_MAX_LEVELS] = {};
  pwm::HuffQuery* d_huff = nullptr;  // Huffman-shaped: code tables (device)
};

namespace pwm {

namespace {

constexpr uint32_t kSbBits = 1u << 16;
constexpr uint32_t kBlkBits = 512;
constexpr uint32_t kBlkPerSb = kSbBits / kBlkBits;  // 128
constexpr uint32_t kSample = 8192;
// select samples hold uint32 superblock indices and the block-count kernel's
// grid has one workgroup per superblock: levels stay below 2^32 superblocks
constexpr uint64_t kMaxLevelBits = uint64_t(1) << 48;

struct LevelRS {
  const uint64_t* words;
  uint64_t nbits;
  uint64_t* sb;      // nsb + 1 entries (last = total)
  uint16_t* blk;     // nblk entries
  uint32_t* s1;      // ns1 entries
  uint32_t* s0;      // ns0 entries
  uint64_t ones;
  uint64_t nsb, nblk, ns1, ns0;
};

// Block counts inside each superblock; one workgroup (128 threads) per
// superblock.
__global__ __launch_bounds__(128) void k_rs_blocks(LevelRS v) {
  __shared__ uint32_t sm[2];
  const uint64_t s = blockIdx.x;
  const uint32_t t = threadIdx.x;
  const uint64_t b = s * kBlkPerSb + t;
  const uint64_t nw = (v.nbits + 63) >> 6;
  uint32_t c = 0;
  if (b < v.nblk) {
    const uint64_t w0 = b * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (w0 + k < nw) c += __popcll(v.words[w0 + k]);
  }
  // exclusive scan over the 128 blocks
  uint32_t inc = c;
  const uint32_t lane = t & 63, wid = t >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t x = __shfl_up(inc, d, 64);
    if (lane >= d) inc += x;
  }
  if (lane == 63) sm[wid] = inc;
  __syncthreads();
  const uint32_t off = wid ? sm[0] : 0;
  if (b < v.nblk) v.blk[b] = uint16_t(off + inc - c);
  if (t == 127) v.sb[s] = sm[0] + sm[1];  // superblock total, scanned later
}

// Exclusive scan of superblock totals (single workgroup).
__global__ __launch_bounds__(1024) void k_rs_sb_scan(LevelRS v) {
  __shared__ uint64_t ws[16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  uint64_t carry = 0;
  for (uint64_t base = 0; base < v.nsb; base += 1024) {
    const uint64_t i = base + t;
    const uint64_t x = i < v.nsb ? v.sb[i] : 0;
    uint64_t inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    if (lane == 63) ws[wid] = inc;
    __syncthreads();
    uint64_t off = 0, tot = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wid) off += ws[w];
      tot += ws[w];
    }
    if (i < v.nsb) v.sb[i] = carry + off + inc - x;
    carry += tot;
    __syncthreads();
  }
  if (t == 0) v.sb[v.nsb] = carry;
}

// Samples: superblock holding the (j*8192+1)-th one / zero.
__global__ void k_rs_samples(LevelRS v) {
  const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= v.nsb) return;
  const uint64_t o0 = v.sb[s], o1 = v.sb[s + 1];
  for (uint64_t j = (o0 + kSample - 1) / kSample; j * kSample + 1 <= o1 && j < v.ns1; ++j)
    if (j * kSample + 1 > o0) v.s1[j] = uint32_t(s);
  const uint64_t b0 = s * kSbBits, b1 = min<uint64_t>((s + 1) * kSbBits, v.nbits);
  const uint64_t z0 = b0 - o0, z1 = (b1 > b0 ? b1 : b0) - o1;
  for (uint64_t j = (z0 + kSample - 1) / kSample; j * kSample + 1 <= z1 && j < v.ns0; ++j)
    if (j * kSample + 1 > z0) v.s0[j] = uint32_t(s);
}

__device__ __forceinline__ uint64_t rank1(const LevelRS& v, uint64_t i) {
  if (i >= v.nbits) return v.ones;
  const uint64_t s = i >> 16, b = i >> 9;
  uint64_t r = v.sb[s] + v.blk[b];
  const uint64_t we = i >> 6;
  for (uint64_t w = b << 3; w < we; ++w) r += __popcll(v.words[w]);
  if (i & 63) r += __popcll(v.words[we] >> (64 - (i & 63)));
  return r;
}

__device__ __forceinline__ bool bit_at(const LevelRS& v, uint64_t i) {
  return (v.words[i >> 6] >> (63 - (i & 63))) & 1ull;
}

// Position (MSB-first) of the r-th set bit (r >= 1) of x.
__device__ __forceinline__ uint32_t word_select(uint64_t x, uint32_t r) {
  uint32_t pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    const uint64_t top = x >> (64 - width);  // the next `width` bits
    const uint32_t c = __popcll(top);
    if (c < r) {
      r -= c;
      x <<= width;
      pos += width;
    }
  }
  return pos;
}

// 0-based position of the k-th one (value 1) or zero (value 0); k >= 1 and
// k <= count is the caller's job.
template <int VALUE>
__device__ uint64_t select_k(const LevelRS& v, uint64_t k) {
  auto sb_cnt = [&](uint64_t s) -> uint64_t {  // ones/zeros before superblock s
    const uint64_t o = v.sb[s];
    return VALUE ? o : (s * uint64_t(kSbBits) - o);
  };
  const uint64_t j = (k - 1) / kSample;
  const uint32_t* samp = VALUE ? v.s1 : v.s0;
  const uint64_t ns = VALUE ? v.ns1 : v.ns0;
  const uint64_t count = VALUE ? v.ones : v.nbits - v.ones;
  uint64_t lo = j < ns ? samp[j] : 0;
  uint64_t hi = ((j + 1) * kSample + 1 <= count) ? samp[j + 1] : v.nsb - 1;
  // last superblock s in [lo, hi] with sb_cnt(s) < k
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (sb_cnt(mid) < k) lo = mid;
    else hi = mid - 1;
  }
  const uint64_t s = lo;
  uint64_t r = k - sb_cnt(s);
  uint64_t blo = s * kBlkPerSb, bhi = min<uint64_t>(blo + kBlkPerSb, v.nblk) - 1;
  const uint64_t bbase = blo;
  auto blk_cnt = [&](uint64_t b) -> uint64_t {
    const uint64_t o = v.blk[b];
    return VALUE ? o : ((b - bbase) * kBlkBits - o);
  };
  while (blo < bhi) {
    const uint64_t mid = (blo + bhi + 1) >> 1;
    if (blk_cnt(mid) < r) blo = mid;
    else bhi = mid - 1;
  }
  r -= blk_cnt(blo);
  const uint64_t nw = (v.nbits + 63) >> 6;
  for (uint64_t w = blo << 3; w < nw; ++w) {
    const uint64_t x = VALUE ? v.words[w] : ~v.words[w];
    const uint32_t c = __popcll(x);
    if (r <= c) return (w << 6) + word_select(x, uint32_t(r));
    r -= c;
  }
  return ~0ull;  // fewer than k (callers check the count first)
}

struct QueryParams {
  const LevelRS* lv;
  uint32_t L;
  int tree;
  uint64_t n;
  const uint64_t* zeros;  // true zero counts per level
  const HuffQuery* huff;  // Huffman-shaped code tables or nullptr
  const uint32_t* op;
  const uint64_t* a;
  const uint64_t* b;
  const uint32_t* level;  // bit-vector queries
  uint64_t* res;
  uint64_t q;
};

constexpr uint64_t kNone = ~0ull;

__global__ void k_bv_query(QueryParams p) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= p.q) return;
  const uint32_t l = p.level[i];
  if (l >= p.L) {
    p.res[i] = kNone;
    return;
  }
  const LevelRS v = p.lv[l];
  const uint64_t x = p.a[i];
  uint64_t r = kNone;
  switch (p.op[i]) {
    case 0: r = rank1(v, min(x, v.nbits)); break;
    case 1: { const uint64_t y = min(x, v.nbits); r = y - rank1(v, y); } break;
    case 2: r = x == 0 ? 0 : (x <= v.ones ? select_k<1>(v, x) : kNone); break;
    case 3: r = x == 0 ? 0 : (x <= v.nbits - v.ones ? select_k<0>(v, x) : kNone); break;
  }
  p.res[i] = r;
}

// Symbol of the complete code (len, word), or kNone if it is no code.
__device__ __forceinline__ uint64_t huff_leaf(const HuffQuery* hq, uint64_t len, uint64_t word) {
  const uint64_t key = (len << 48) | word;
  uint32_t lo = 0, hi = hq->nleaf;  // first entry >= key
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (hq->leaf_key[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < hq->nleaf && hq->leaf_key[lo] == key ? uint64_t(hq->leaf_sym[lo]) : kNone;
}

// Symbol queries (query_support.hpp:72-178 recursions, iterative) on one
// structure: levels lv[0..L), true zero counts, text length n.  With code
// tables (Huffman-shaped, huff_naive's layout) a symbol's path is its code:
// the same recursions over its code length; the code assignment of the
// reference (huff_codes.hpp:141-285) keeps, at every level, the codes that
// end there behind the ones that go on, so positions below a level are
// unchanged by the missing leaves.
__device__ uint64_t sym_query(const LevelRS* lv, uint32_t L, int tree, uint64_t n, const uint64_t* zeros,
                              const HuffQuery* hq, uint32_t op, uint64_t a, uint64_t b) {
  uint64_t res = kNone;
  if (op == 0) {  // access(a)
    if (a < n) {
      uint64_t sym = 0, i = a, s = 0, e = n;
      for (uint32_t l = 0; l < L; ++l) {
        const LevelRS& v = lv[l];
        if (!tree) {
          const bool bit = bit_at(v, i);
          sym = (sym << 1) | bit;
          if (hq) {
            const uint64_t leaf = huff_leaf(hq, l + 1, sym);
            if (leaf != kNone) {
              sym = leaf;
              break;
            }
          }
          const uint64_t r1 = rank1(v, i);
          i = bit ? zeros[l] + r1 : i - r1;
        } else {
          const bool bit = bit_at(v, s + i);
          sym = (sym << 1) | bit;
          if (hq) {
            const uint64_t leaf = huff_leaf(hq, l + 1, sym);
            if (leaf != kNone) {
              sym = leaf;
              break;
            }
          }
          const uint64_t r1s = rank1(v, s), r1e = rank1(v, e), r1i = rank1(v, s + i);
          const uint64_t zn = (e - s) - (r1e - r1s);
          if (bit) { i = r1i - r1s; s = s + zn; }
          else { i = i - (r1i - r1s); e = s + zn; }
        }
      }
      res = sym;
    }
  } else {
    // the symbol's path: its code (Huffman-shaped) or its L bits
    uint64_t c = a;
    bool absent = L < 64 && (c >> L) != 0;
    if (hq) {
      absent = c >= hq->nsym || hq->len[c] == 0;
      if (!absent) {
        L = uint32_t(hq->len[c]);
        c = hq->word[c];
      }
    }
    if (absent) {
      res = op == 1 ? 0 : (b == 0 ? 0 : kNone);
    } else if (!tree) {
      // rank: (index, offset) pair (rank_matrix); select walks back up
      uint64_t idx = op == 1 ? min(b, n) : n, off = 0;
      for (uint32_t l = 0; l < L; ++l) {
        const LevelRS& v = lv[l];
        const bool bit = (c >> (L - 1 - l)) & 1;
        const uint64_t ri = rank1(v, idx), ro = rank1(v, off);
        if (bit) { idx = zeros[l] + ri; off = zeros[l] + ro; }
        else { idx = idx - ri; off = off - ro; }
      }
      if (op == 1) res = idx - off;
      else if (b == 0) res = 0;
      else if (b <= idx - off) {
        uint64_t q = off + b - 1;  // position among the last level's successors
        for (int l = int(L) - 1; l >= 0; --l) {
          const LevelRS& v = lv[l];
          const bool bit = (c >> (L - 1 - l)) & 1;
          q = bit ? select_k<1>(v, q - zeros[l] + 1) : select_k<0>(v, q + 1);
        }
        res = q;
      }
    } else {
      // tree: node ranges down the path of c
      uint64_t st[PWMHIP_MAX_LEVELS];
      uint64_t s = 0, e = n, i = op == 1 ? min(b, n) : 0;
      for (uint32_t l = 0; l < L; ++l) {
        const LevelRS& v = lv[l];
        st[l] = s;
        const bool bit = (c >> (L - 1 - l)) & 1;
        const uint64_t r1s = rank1(v, s), r1e = rank1(v, e), r1i = rank1(v, s + i);
        const uint64_t zn = (e - s) - (r1e - r1s);
        if (bit) { i = r1i - r1s; s = s + zn; }
        else { i = i - (r1i - r1s); e = s + zn; }
      }
      if (op == 1) res = i;
      else if (b == 0) res = 0;
      else if (b <= e - s) {
        uint64_t q = b - 1;  // position inside the leaf
        for (int l = int(L) - 1; l >= 0; --l) {
          const LevelRS& v = lv[l];
          const bool bit = (c >> (L - 1 - l)) & 1;
          const uint64_t s0 = st[l];
          q = bit ? select_k<1>(v, rank1(v, s0) + q + 1) - s0 : select_k<0>(v, (s0 - rank1(v, s0)) + q + 1) - s0;
        }
        res = q;
      }
    }
  }
  return res;
}

__global__ void k_sym_query(QueryParams p) {
  const uint64_t qi = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (qi >= p.q) return;
  p.res[qi] = sym_query(p.lv, p.L, p.tree, p.n, p.zeros, p.huff, p.op[qi], p.a[qi], p.b[qi]);
}

// Queries over many resident structures: query i reads structure qidx[i].
__global__ void k_multi_query(const QueryIndexDesc* desc, const uint32_t* qidx, const uint32_t* op,
                              const uint64_t* a, const uint64_t* b, uint64_t* res, uint64_t q) {
  const uint64_t qi = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (qi >= q) return;
  const QueryIndexDesc d = desc[qidx[qi]];
  res[qi] = sym_query(static_cast<const LevelRS*>(d.lv), d.L, d.tree, d.n, d.zeros, d.huff, op[qi], a[qi], b[qi]);
}

}  // namespace

}  // namespace pwm

using namespace pwm;

namespace pwm {

QueryIndexDesc query_desc(const pwmhip_rs* rs) {
  QueryIndexDesc d{};
  d.lv = rs->d_desc;
  d.zeros = reinterpret_cast<const uint64_t*>(static_cast<const char*>(rs->d_desc) +
                                              ((rs->levels * sizeof(LevelRS) + 255) & ~size_t(255)));
  d.n = rs->n;
  d.L = uint32_t(rs->levels);
  d.tree = rs->is_tree;
  d.huff = rs->d_huff;
  return d;
}

void launch_multi_query(hipStream_t s, const QueryIndexDesc* d_desc, const uint32_t* d_qidx, const uint32_t* op,
                        const uint64_t* a, const uint64_t* b, uint64_t* res, uint64_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_multi_query, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, d_desc, d_qidx, op, a, b,
                     res, n);
  PWM_HIP_TRY(hipGetLastError());
}

}  // namespace pwm


extern "C" {

int pwmhip_rs_build(pwmhip_ctx* ctx, const pwmhip_ws* ws, pwmhip_rs** out) {
  if (!ctx || !ws || !out) return PWMHIP_ERR_INVALID;
  *out = nullptr;
  // Huffman-shaped with the reference's huff_pc node borders: not a layout
  // the symbol recursions can walk
  if (ws->is_huffman == 1) return PWMHIP_ERR_UNSUPPORTED;
  const bool huff = ws->is_huffman != 0;
  if (ws->levels > PWMHIP_MAX_LEVELS) return PWMHIP_ERR_INVALID;
  for (uint64_t l = 0; l < ws->levels; ++l)
    if (ws->level_bits[l] >= kMaxLevelBits) return PWMHIP_ERR_UNSUPPORTED;
  auto* rs = new (std::nothrow) pwmhip_rs();
  if (!rs) return PWMHIP_ERR_OOM;
  const int st = pwm::guarded([&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    const uint64_t L = ws->levels;
    rs->levels = L;
    rs->n = ws->text_size;
    rs->is_tree = ws->is_tree;
    std::vector<LevelRS> desc(L);
    size_t bytes = 0;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    for (uint64_t l = 0; l < L; ++l) {
      LevelRS& v = desc[l];
      v.nbits = ws->level_bits[l];
      v.nsb = (v.nbits >> 16) + 1;
      v.nblk = (v.nbits >> 9) + 1;
      v.ns1 = v.nbits / kSample + 1;
      v.ns0 = v.nbits / kSample + 1;
      bytes += al((v.nsb + 1) * 8) + al(v.nblk * 2) + al(v.ns1 * 4) + al(v.ns0 * 4);
    }
    bytes += al(L * sizeof(LevelRS)) + al(L * 8);
    // Huffman-shaped: code tables and the sorted complete codes
    const uint64_t S = huff ? ws->n_codes : 0;
    std::vector<uint64_t> hlen(S), hword(S), leaf_key;
    std::vector<uint32_t> leaf_sym;
    if (huff) {
      std::vector<std::pair<uint64_t, uint32_t>> leaves;
      for (uint64_t c = 0; c < S; ++c) {
        hlen[c] = ws->code_len[c];
        hword[c] = ws->code_word[c];
        if (hlen[c]) leaves.emplace_back((hlen[c] << 48) | hword[c], uint32_t(c));
      }
      std::sort(leaves.begin(), leaves.end());
      for (auto& e : leaves) {
        leaf_key.push_back(e.first);
        leaf_sym.push_back(e.second);
      }
      bytes += al(sizeof(HuffQuery)) + al(S * 8) * 2 + al(leaf_key.size() * 8 + 8) + al(leaf_sym.size() * 4 + 4);
    }
    PWM_HIP_TRY(hipMalloc(&rs->mem, std::max<size_t>(bytes, 256)));
    char* p = static_cast<char*>(rs->mem);
    for (uint64_t l = 0; l < L; ++l) {
      LevelRS& v = desc[l];
      v.words = ws->d_words + ws->level_offset[l];
      v.sb = reinterpret_cast<uint64_t*>(p); p += al((v.nsb + 1) * 8);
      v.blk = reinterpret_cast<uint16_t*>(p); p += al(v.nblk * 2);
      v.s1 = reinterpret_cast<uint32_t*>(p); p += al(v.ns1 * 4);
      v.s0 = reinterpret_cast<uint32_t*>(p); p += al(v.ns0 * 4);
      hipLaunchKernelGGL(k_rs_blocks, dim3(uint32_t(v.nsb)), dim3(128), 0, ctx->stream, v);
      hipLaunchKernelGGL(k_rs_sb_scan, dim3(1), dim3(1024), 0, ctx->stream, v);
    }
    rs->d_desc = p;
    p += al(L * sizeof(LevelRS));
    uint64_t* d_zeros = reinterpret_cast<uint64_t*>(p);
    p += al(L * 8);
    if (huff) {
      HuffQuery hq{};
      rs->d_huff = reinterpret_cast<HuffQuery*>(p); p += al(sizeof(HuffQuery));
      uint64_t* d_len = reinterpret_cast<uint64_t*>(p); p += al(S * 8);
      uint64_t* d_word = reinterpret_cast<uint64_t*>(p); p += al(S * 8);
      uint64_t* d_key = reinterpret_cast<uint64_t*>(p); p += al(leaf_key.size() * 8 + 8);
      uint32_t* d_sym = reinterpret_cast<uint32_t*>(p); p += al(leaf_sym.size() * 4 + 4);
      hq.len = d_len;
      hq.word = d_word;
      hq.nsym = S;
      hq.leaf_key = d_key;
      hq.leaf_sym = d_sym;
      hq.nleaf = uint32_t(leaf_key.size());
      if (S) {
        PWM_HIP_TRY(hipMemcpyAsync(d_len, hlen.data(), S * 8, hipMemcpyHostToDevice, ctx->stream));
        PWM_HIP_TRY(hipMemcpyAsync(d_word, hword.data(), S * 8, hipMemcpyHostToDevice, ctx->stream));
      }
      if (!leaf_key.empty()) {
        PWM_HIP_TRY(hipMemcpyAsync(d_key, leaf_key.data(), leaf_key.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        PWM_HIP_TRY(hipMemcpyAsync(d_sym, leaf_sym.data(), leaf_sym.size() * 4, hipMemcpyHostToDevice, ctx->stream));
      }
      PWM_HIP_TRY(hipMemcpyAsync(rs->d_huff, &hq, sizeof(hq), hipMemcpyHostToDevice, ctx->stream));
      PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));  // hq and the vectors are host stack objects
    }
    // totals -> ones per level
    std::vector<uint64_t> ones(L);
    for (uint64_t l = 0; l < L; ++l)
      PWM_HIP_TRY(hipMemcpyAsync(&ones[l], desc[l].sb + desc[l].nsb, 8, hipMemcpyDeviceToHost, ctx->stream));
    PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (uint64_t l = 0; l < L; ++l) {
      desc[l].ones = ones[l];
      rs->true_zeros[l] = desc[l].nbits - ones[l];
      hipLaunchKernelGGL(k_rs_samples, dim3(uint32_t((desc[l].nsb + 255) / 256)), dim3(256), 0, ctx->stream,
                         desc[l]);
    }
    PWM_HIP_TRY(hipGetLastError());
    PWM_HIP_TRY(hipMemcpyAsync(rs->d_desc, desc.data(), L * sizeof(LevelRS), hipMemcpyHostToDevice, ctx->stream));
    PWM_HIP_TRY(hipMemcpyAsync(d_zeros, rs->true_zeros, L * 8, hipMemcpyHostToDevice, ctx->stream));
    PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  });
  if (st != PWMHIP_OK) {
    if (rs->mem) (void)hipFree(rs->mem);
    delete rs;
    return st;
  }
  *out = rs;
  return PWMHIP_OK;
}

int pwmhip_rs_free(pwmhip_ctx* ctx, pwmhip_rs* rs) {
  if (!ctx || !rs) return PWMHIP_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  if (rs->mem) (void)hipFree(rs->mem);
  delete rs;
  return PWMHIP_OK;
}

}  // extern "C"

namespace {

// Stages host arrays into the context's query arena, runs `launch`, copies
// results back.
template <typename F>
int run_query(pwmhip_ctx* ctx, uint64_t n, uint32_t flags, const void* const* ins, const size_t* isz, int nin,
              uint64_t* result, F&& launch) {
  return pwm::guarded([&] {
    PWM_HIP_TRY(hipSetDevice(ctx->device));
    if (n == 0) return;
    const bool on_dev = (flags & PWMHIP_FLAG_TEXT_ON_DEVICE) != 0;
    std::vector<const void*> d_in(nin);
    uint64_t* d_res = result;
    if (!on_dev) {
      size_t need = n * 8 + 256;
      for (int k = 0; k < nin; ++k) need += ((isz[k] * n + 255) & ~size_t(255));
      ctx->rs_scratch.ensure(need);
      ctx->rs_scratch.reset();
      for (int k = 0; k < nin; ++k) {
        void* d = ctx->rs_scratch.take<char>(isz[k] * n);
        if (ins[k]) PWM_HIP_TRY(hipMemcpyAsync(d, ins[k], isz[k] * n, hipMemcpyHostToDevice, ctx->stream));
        d_in[k] = ins[k] ? d : nullptr;
      }
      d_res = ctx->rs_scratch.take<uint64_t>(n);
    } else {
      for (int k = 0; k < nin; ++k) d_in[k] = ins[k];
    }
    launch(d_in.data(), d_res);
    PWM_HIP_TRY(hipGetLastError());
    if (!on_dev) PWM_HIP_TRY(hipMemcpyAsync(result, d_res, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    PWM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  });
}

}  // namespace

extern "C" {

int pwmhip_bv_query(pwmhip_ctx* ctx, const pwmhip_rs* rs, const uint32_t* level, const uint32_t* op,
                    const uint64_t* arg, uint64_t n, uint64_t* result, uint32_t flags) {
  if (!ctx || !rs || (n && (!level || !op || !arg || !result))) return PWMHIP_ERR_INVALID;
  const void* ins[3] = {level, op, arg};
  const size_t isz[3] = {4, 4, 8};
  return run_query(ctx, n, flags, ins, isz, 3, result, [&](const void* const* d, uint64_t* r) {
    QueryParams p{};
    p.lv = static_cast<const LevelRS*>(rs->d_desc);
    p.L = uint32_t(rs->levels);
    p.level = static_cast<const uint32_t*>(d[0]);
    p.op = static_cast<const uint32_t*>(d[1]);
    p.a = static_cast<const uint64_t*>(d[2]);
    p.res = r;
    p.q = n;
    hipLaunchKernelGGL(k_bv_query, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, ctx->stream, p);
  });
}

int pwmhip_query(pwmhip_ctx* ctx, const pwmhip_ws* ws, const pwmhip_rs* rs, const uint32_t* op,
                 const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t* result, uint32_t flags) {
  if (!ctx || !ws || !rs || (n && (!op || !a || !result))) return PWMHIP_ERR_INVALID;
  if (ws->is_huffman == 1) return PWMHIP_ERR_UNSUPPORTED;
  const void* ins[3] = {op, a, b};
  const size_t isz[3] = {4, 8, 8};
  return run_query(ctx, n, flags, ins, isz, 3, result, [&](const void* const* d, uint64_t* r) {
    QueryParams p{};
    p.lv = static_cast<const LevelRS*>(rs->d_desc);
    p.L = uint32_t(rs->levels);
    p.tree = rs->is_tree;
    p.n = rs->n;
    p.zeros = reinterpret_cast<const uint64_t*>(static_cast<const char*>(rs->d_desc) +
                                                ((rs->levels * sizeof(LevelRS) + 255) & ~size_t(255)));
    p.huff = rs->d_huff;
    p.op = static_cast<const uint32_t*>(d[0]);
    p.a = static_cast<const uint64_t*>(d[1]);
    p.b = d[2] ? static_cast<const uint64_t*>(d[2]) : static_cast<const uint64_t*>(d[1]);
    p.res = r;
    p.q = n;
    hipLaunchKernelGGL(k_sym_query, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, ctx->stream, p);
  });
}

}  // extern "C"
