// oracle/oracle.cpp -- CPU restatement of kurpicz/pwm's construction and
// query semantics.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (pwm_amd/, include/)
// links, loads or calls this file; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, as the checker.
//
// Every function cites the reference file:line whose behaviour it restates.
// Each function is pinned (tests/test_oracle.py) by:
//   * golden vectors produced by the compiled reference (tests/golden/), and
//   * closed-form / brute-force properties (decode round trip, popcount ==
//     n - zeros, rank/select by brute force, the bit_copy known-answer tests).
//
// Build: oracle/Makefile  ->  oracle/liboracle.so  (plain g++, no GPU).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <queue>
#include <random>
#include <unordered_map>
#include <vector>

namespace {

inline uint64_t load_sym(const void* text, uint32_t width, uint64_t i) {
  switch (width) {
    case 1: return static_cast<const uint8_t*>(text)[i];
    case 2: return static_cast<const uint16_t*>(text)[i];
    case 4: return static_cast<const uint32_t*>(text)[i];
    default: return static_cast<const uint64_t*>(text)[i];
  }
}

inline uint64_t words_for(uint64_t bits) { return (bits + 63) >> 6; }

// MSB-first packing of a bit sequence, zero padded tail
// (reference: construction/building_blocks.hpp:18-50, util/common.hpp:60-68).
inline void pack_bits(const std::vector<uint8_t>& bits, uint64_t* out) {
  const uint64_t w = words_for(bits.size());
  for (uint64_t k = 0; k < w; ++k) out[k] = 0;
  for (uint64_t j = 0; j < bits.size(); ++j)
    if (bits[j]) out[j >> 6] |= 1ULL << (63 - (j & 63));
}

inline bool get_bit(const uint64_t* words, uint64_t j) {
  return (words[j >> 6] >> (63 - (j & 63))) & 1ULL;
}

// ---------------------------------------------------------------------------
// Canonical Huffman codes (reference: huffman/huff_codes.hpp:141-285,
// util/histogram.hpp:39-76, huffman/huff_level_sizes_builder.hpp:13-46).
// The tie order of std::priority_queue / std::sort decides the exact code
// words, so the same standard algorithms are applied to the same input
// sequence here.
// ---------------------------------------------------------------------------
struct huff_result {
  std::vector<uint64_t> len;    // per symbol 0..max_symbol
  std::vector<uint64_t> word;   // per symbol
  std::vector<uint64_t> level_sizes;
  uint64_t levels = 0;
  bool empty = true;
};

struct freq_item {
  uint64_t occ;
  std::vector<uint64_t> syms;
  bool operator>(const freq_item& o) const { return occ > o.occ; }
};

huff_result huff_codes(const void* text, uint64_t n, uint32_t width, bool is_tree) {
  huff_result r;
  // histogram entries in the reference's order: ascending symbol for bytes,
  // unordered_map iteration order (insertion in text order) otherwise.
  std::vector<std::pair<uint64_t, uint64_t>> hist;
  uint64_t max_sym = 0;
  if (width == 1) {
    std::vector<uint64_t> cnt(256, 0);
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t c = load_sym(text, 1, i);
      max_sym = std::max(max_sym, c);
      ++cnt[c];
    }
    for (uint64_t c = 0; c < 256; ++c)
      if (cnt[c]) hist.emplace_back(c, cnt[c]);
  } else {
    auto build = [&](auto zero) {
      using A = decltype(zero);
      std::unordered_map<A, uint64_t> m;
      for (uint64_t i = 0; i < n; ++i) {
        A c = static_cast<A>(load_sym(text, width, i));
        max_sym = std::max<uint64_t>(max_sym, c);
        auto it = m.find(c);
        if (it == m.end()) m.emplace(c, 1);
        else ++(it->second);
      }
      for (const auto& e : m) hist.emplace_back(uint64_t(e.first), e.second);
    };
    if (width == 2) build(uint16_t(0));
    else if (width == 4) build(uint32_t(0));
    else build(uint64_t(0));
  }
  if (hist.empty()) return r;
  r.empty = false;

  std::priority_queue<freq_item, std::vector<freq_item>, std::greater<freq_item>> pq;
  r.len.assign(max_sym + 1, 0);
  r.word.assign(max_sym + 1, 0);
  for (auto& h : hist) pq.emplace(freq_item{h.second, std::vector<uint64_t>{h.first}});
  if (pq.size() == 1) ++r.len[pq.top().syms.front()];
  while (pq.size() > 1) {
    auto a = pq.top(); pq.pop();
    auto b = pq.top(); pq.pop();
    std::move(b.syms.begin(), b.syms.end(), std::back_inserter(a.syms));
    for (auto c : a.syms) ++r.len[c];
    pq.emplace(freq_item{a.occ + b.occ, a.syms});
  }
  std::vector<uint64_t> order(r.len.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(),
            [&](uint64_t a, uint64_t b) { return r.len[a] < r.len[b]; });
  r.levels = r.len[order.back()];
  std::vector<uint64_t> freq(max_sym + 1, 0);
  for (auto& h : hist) freq[h.first] = h.second;
  r.level_sizes.assign(r.levels, 0);

  uint64_t nr = 0;
  while (nr < order.size() && r.len[order[nr]] == 0) ++nr;
  if (!is_tree) {
    uint64_t cur = 1;
    std::vector<uint64_t> cws{0, 1};
    while (nr < order.size()) {
      if (r.len[order[nr]] > cur) {
        for (uint64_t i = cur; i < r.len[order[nr]]; ++i) {
          std::vector<uint64_t> nw;
          nw.reserve(cws.size() * 2);
          for (auto cw : cws) nw.push_back(cw << 1);
          for (auto cw : cws) nw.push_back((cw << 1) + 1);
          cws.swap(nw);
        }
        cur = r.len[order[nr]];
      }
      while (nr < order.size() && r.len[order[nr]] == cur) {
        uint64_t s = order[nr++];
        r.word[s] = cws.back();
        cws.pop_back();
        r.level_sizes[r.len[s] - 1] += freq[s];
      }
    }
  } else {
    uint64_t cw = 0;
    auto gen = [&] {
      uint64_t s = order[nr];
      uint64_t l = r.len[s];
      uint64_t mask = (l >= 64) ? ~0ULL : ((1ULL << l) - 1);
      r.word[s] = (~cw) & mask;
      r.level_sizes[l - 1] += freq[s];
    };
    gen();
    for (++nr; nr < order.size(); ++nr) {
      cw = (cw + 1) << (r.len[order[nr]] - r.len[order[nr - 1]]);
      gen();
    }
  }
  for (uint64_t i = r.levels - 1; i > 0; --i) r.level_sizes[i - 1] += r.level_sizes[i];
  return r;
}

inline uint64_t code_bit(uint64_t len, uint64_t word, uint64_t l) {
  return (word >> (len - 1 - l)) & 1ULL;  // huff_codes.hpp:58-61
}
inline uint64_t code_prefix(uint64_t len, uint64_t word, uint64_t l) {
  return word >> (len - l);  // huff_codes.hpp:51-54
}

}  // namespace

extern "C" {

int oracle_version() { return 3; }

// ---------------------------------------------------------------------------
// wx_naive: top-down construction of every level
// (reference: wx_naive.hpp:44-75 for trees, :108-151 for matrices).
// out: levels * ceil(n/64) words, level l at out + l*W.  zeros: levels
// entries (matrix only; true count of 0-bits per level).
// ---------------------------------------------------------------------------
int oracle_wx_naive(const void* text, uint64_t n, uint32_t width, uint32_t levels,
                    int is_tree, uint64_t* out, uint64_t* zeros) {
  const uint64_t W = words_for(n);
  std::vector<uint64_t> cur(n);
  for (uint64_t i = 0; i < n; ++i) cur[i] = load_sym(text, width, i);
  std::vector<uint8_t> bits(n);
  for (uint32_t l = 0; l < levels; ++l) {
    const uint32_t shift = levels - 1 - l;
    uint64_t z = 0;
    for (uint64_t j = 0; j < n; ++j) {
      bits[j] = (cur[j] >> shift) & 1ULL;
      z += bits[j] == 0;
    }
    pack_bits(bits, out + uint64_t(l) * W);
    if (!is_tree && zeros) zeros[l] = z;
    if (l + 1 == levels) break;
    std::vector<uint64_t> next;
    next.reserve(n);
    if (is_tree) {
      // stable bucket sort by the (l+1)-bit prefix, numeric node order
      std::vector<std::vector<uint64_t>> buckets(1ULL << (l + 1));
      for (uint64_t c : cur) buckets[c >> shift].push_back(c);
      for (auto& b : buckets) next.insert(next.end(), b.begin(), b.end());
    } else {
      // stable partition by the current bit, zeros first
      for (uint64_t c : cur) if (!((c >> shift) & 1)) next.push_back(c);
      for (uint64_t c : cur) if ((c >> shift) & 1) next.push_back(c);
    }
    cur.swap(next);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Huffman codes: per-symbol code length / word for symbols 0..max_symbol,
// the number of levels and level sizes (huff_codes.hpp:141-285).
// Two-call protocol: call with code_len == nullptr to get *n_symbols.
// ---------------------------------------------------------------------------
int oracle_huff_codes(const void* text, uint64_t n, uint32_t width, int is_tree,
                      uint64_t* n_symbols, uint64_t* levels, uint64_t* code_len,
                      uint64_t* code_word, uint64_t* level_sizes) {
  huff_result r = huff_codes(text, n, width, is_tree != 0);
  *n_symbols = r.len.size();
  *levels = r.levels;
  if (code_len) {
    std::copy(r.len.begin(), r.len.end(), code_len);
    std::copy(r.word.begin(), r.word.end(), code_word);
    std::copy(r.level_sizes.begin(), r.level_sizes.end(), level_sizes);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// huff_naive (reference: huffman/huff_naive.hpp:18-126): level l holds the
// occurrences with code_length > l; WM order is the stable partition by the
// previous code bit with finished symbols dropped, WT order is numeric by
// the l-bit code prefix.  out: levels concatenated, level l starting at word
// offsets[l] (ceil(level_sizes[l]/64) words each).  zeros: true counts.
// ---------------------------------------------------------------------------
int oracle_huff_naive(const void* text, uint64_t n, uint32_t width, int is_tree,
                      uint64_t* out, uint64_t* zeros) {
  huff_result r = huff_codes(text, n, width, is_tree != 0);
  if (r.empty || n == 0) return 0;
  std::vector<uint64_t> cur(n);
  for (uint64_t i = 0; i < n; ++i) cur[i] = load_sym(text, width, i);
  uint64_t off = 0;
  for (uint64_t l = 0; l < r.levels; ++l) {
    std::vector<uint8_t> bits(cur.size());
    uint64_t z = 0;
    for (uint64_t j = 0; j < cur.size(); ++j) {
      bits[j] = code_bit(r.len[cur[j]], r.word[cur[j]], l);
      z += bits[j] == 0;
    }
    pack_bits(bits, out + off);
    off += words_for(cur.size());
    if (zeros) zeros[l] = is_tree ? 0 : z;
    std::vector<uint64_t> next;
    if (is_tree) {
      std::vector<std::vector<uint64_t>> buckets(1ULL << (l + 1));
      for (uint64_t c : cur)
        if (r.len[c] > l + 1) buckets[code_prefix(r.len[c], r.word[c], l + 1)].push_back(c);
      for (auto& b : buckets) next.insert(next.end(), b.begin(), b.end());
    } else {
      for (uint64_t c : cur)
        if (r.len[c] > l + 1 && !code_bit(r.len[c], r.word[c], l)) next.push_back(c);
      for (uint64_t c : cur)
        if (r.len[c] > l + 1 && code_bit(r.len[c], r.word[c], l)) next.push_back(c);
    }
    cur.swap(next);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// huff_pc (reference: huffman/huff_pc.hpp:19-83 with
// huff_building_blocks.hpp:104-222 and ctx_huffman.hpp:80-102).  Restated
// with the reference's border bookkeeping, including its quirks:
//   * borders are assigned by key *position* i (borders[i] = borders[i-1] +
//     hist[i-1]) for i < #keys, where the keys are the pre-seeded l-bit code
//     prefixes of all codes with length >= l (huff_building_blocks.hpp:186-222);
//   * the border array persists across levels (levels < 8 share one array,
//     levels >= 8 another), so unassigned entries keep earlier values;
//   * zeros[l-1] = borders[1] (matrix), zeros[levels-1] stays 0;
//   * at level 1 symbols of code length 1 still advance their border.
// Returns -2 if an occurrence's position falls outside its level (the
// reference reads and writes past the level there: undefined behaviour).
// ---------------------------------------------------------------------------
int oracle_huff_pc(const void* text, uint64_t n, uint32_t width, int is_tree,
                   uint64_t* out, uint64_t* zeros) {
  huff_result r = huff_codes(text, n, width, is_tree != 0);
  if (r.empty || n == 0) return 0;
  const uint64_t L = r.levels;
  std::vector<uint64_t> off(L + 1, 0);
  for (uint64_t l = 0; l < L; ++l) off[l + 1] = off[l] + words_for(r.level_sizes[l]);
  std::fill(out, out + off[L], 0);
  std::vector<uint64_t> sym(n);
  for (uint64_t i = 0; i < n; ++i) sym[i] = load_sym(text, width, i);

  // level 0 in text order
  {
    std::vector<uint8_t> bits(n);
    for (uint64_t i = 0; i < n; ++i) bits[i] = code_bit(r.len[sym[i]], r.word[sym[i]], 0);
    pack_bits(bits, out);
  }
  // per-level prefix histograms, pre-seeded (ctx_huffman.hpp:95-101) and
  // filled by the first scan (huff_building_blocks.hpp:113-126)
  std::vector<std::unordered_map<uint64_t, uint64_t>> hist(L + 1);
  for (uint64_t i = 1; i < L; ++i)
    for (uint64_t s = 0; s < r.len.size(); ++s)
      if (r.len[s] >= i) hist[i][code_prefix(r.len[s], r.word[s], i)] = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t s = sym[i];
    for (uint64_t l = 1; l <= r.len[s]; ++l) hist[l][code_prefix(r.len[s], r.word[s], l)]++;
  }
  std::unordered_map<uint64_t, uint64_t> upper, lower;
  if (zeros) std::fill(zeros, zeros + L, 0);
  for (uint64_t l = 1; l < L && r.level_sizes[l] > 0; ++l) {
    auto& B = (l < 8) ? upper : lower;
    const uint64_t keys = hist[l].size();
    B[0] = 0;
    for (uint64_t i = 1; i < keys; ++i) B[i] = B[i - 1] + hist[l][i - 1];
    if (!is_tree && zeros) zeros[l - 1] = B[1];
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t s = sym[i];
      const uint64_t len = r.len[s];
      if (len > l) {
        const uint64_t p = code_prefix(len, r.word[s], l);
        const uint64_t pos = B[p]++;
        if (pos >= r.level_sizes[l]) return -2;  // outside the level: undefined
        if (code_bit(len, r.word[s], l)) out[off[l] + (pos >> 6)] |= 1ULL << (63 - (pos & 63));
      } else if (l == 1 && len == 1) {
        B[code_prefix(len, r.word[s], 1)]++;
      }
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Bit-vector primitives.
// rank1(i) = ones in [0, i)           (queries/bin_rank_popcnt.hpp:98-128)
// select1(k) = 0-based position of the k-th one, k >= 1, select(0) = 0
//                                     (queries/bin_select_popcnt.hpp:115-120)
// Brute force; used to pin the GPU rank/select tables.
// ---------------------------------------------------------------------------
uint64_t oracle_rank1(const uint64_t* words, uint64_t i) {
  uint64_t r = 0;
  for (uint64_t k = 0; k < (i >> 6); ++k) r += __builtin_popcountll(words[k]);
  if (i & 63) r += __builtin_popcountll(words[i >> 6] >> (64 - (i & 63)));
  return r;
}

uint64_t oracle_select(const uint64_t* words, uint64_t nbits, int value, uint64_t k) {
  if (k == 0) return 0;
  uint64_t seen = 0;
  for (uint64_t j = 0; j < nbits; ++j) {
    if (get_bit(words, j) == (value != 0) && ++seen == k) return j;
  }
  return nbits;  // fewer than k occurrences
}

// Batched brute-force symbol queries over the text (the meaning of
// query_support.hpp:37-69 for access / rank / select):
// op 0: access(a)     -> T[a]
// op 1: rank(c=a, i=b)-> #{j < b : T[j] == a}
// op 2: select(c=a, k=b) -> 0-based position of the b-th a (b >= 1)
// Uses per-symbol occurrence lists; O(n + q log n).
int oracle_symbol_queries(const void* text, uint64_t n, uint32_t width, uint64_t q,
                          const uint32_t* ops, const uint64_t* a, const uint64_t* b,
                          uint64_t* res) {
  std::unordered_map<uint64_t, std::vector<uint64_t>> occ;
  for (uint64_t i = 0; i < n; ++i) occ[load_sym(text, width, i)].push_back(i);
  for (uint64_t t = 0; t < q; ++t) {
    if (ops[t] == 0) {
      res[t] = a[t] < n ? load_sym(text, width, a[t]) : ~0ULL;
    } else {
      auto it = occ.find(a[t]);
      if (ops[t] == 1) {
        if (it == occ.end()) { res[t] = 0; continue; }
        res[t] = std::lower_bound(it->second.begin(), it->second.end(), b[t]) - it->second.begin();
      } else {
        if (b[t] == 0) { res[t] = 0; continue; }
        if (it == occ.end() || b[t] > it->second.size()) { res[t] = ~0ULL; continue; }
        res[t] = it->second[b[t] - 1];
      }
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// copy_bits (reference: construction/merge.hpp:19-190): copy `len` bits
// from src[src_off..] to dst[dst_off..], MSB-first in 64-bit words.  Bits of
// partially covered destination words are OR-ed in, fully covered words are
// overwritten.
// ---------------------------------------------------------------------------
void oracle_copy_bits64(uint64_t* dst, const uint64_t* src, uint64_t dst_off,
                        uint64_t src_off, uint64_t len) {
  const uint64_t end = dst_off + len;
  while (dst_off < end) {
    if ((dst_off & 63) == 0 && end - dst_off >= 64) {
      uint64_t w = 0;
      for (int i = 0; i < 64; ++i) w = (w << 1) | (get_bit(src, src_off + i) ? 1 : 0);
      dst[dst_off >> 6] = w;
      dst_off += 64;
      src_off += 64;
    } else {
      if (get_bit(src, src_off)) dst[dst_off >> 6] |= 1ULL << (63 - (dst_off & 63));
      ++dst_off;
      ++src_off;
    }
  }
}

// Byte-word variant used by the reference's own known-answer tests
// (test/bit_copy.cpp:45-127 use copy_bits<uint8_t>).
void oracle_copy_bits8(uint8_t* dst, const uint8_t* src, uint64_t dst_off,
                       uint64_t src_off, uint64_t len) {
  for (uint64_t i = 0; i < len; ++i) {
    const uint64_t s = src_off + i, d = dst_off + i;
    if ((src[s >> 3] >> (7 - (s & 7))) & 1) dst[d >> 3] |= uint8_t(1u << (7 - (d & 7)));
  }
}

// The bit vector of the reference's select test
// (test/bin_rank_select_tests.cpp:39-62): n_words words drawn in order from
// std::uniform_int_distribution<uint64_t>(0) over std::mt19937(seed).  Both
// are the standard library's, so the same library gives the same words.
void oracle_mt19937_words(uint64_t seed, uint64_t n_words, uint64_t* out) {
  std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
  std::uniform_int_distribution<uint64_t> dis(0);
  for (uint64_t i = 0; i < n_words; ++i) out[i] = dis(gen);
}

// reduce_alphabet (util/alphabet_util.hpp:18-45): symbols renumbered in
// first-occurrence order; returns levels_for_max_char(distinct - 1)
// (util/alphabet_util.hpp:99-106).  The reference's byte table has 255
// entries (byte 255 is undefined there); all 256 values are handled here.
uint64_t oracle_reduce_alphabet(void* text, uint64_t n, uint32_t width) {
  std::unordered_map<uint64_t, uint64_t> code;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t c = load_sym(text, width, i);
    if (!code.count(c)) {
      const uint64_t k = code.size();
      code.emplace(c, k);
    }
  }
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t v = code[load_sym(text, width, i)];
    switch (width) {
      case 1: static_cast<uint8_t*>(text)[i] = uint8_t(v); break;
      case 2: static_cast<uint16_t*>(text)[i] = uint16_t(v); break;
      case 4: static_cast<uint32_t*>(text)[i] = uint32_t(v); break;
      default: static_cast<uint64_t*>(text)[i] = v; break;
    }
  }
  uint64_t max_char = code.empty() ? 0 : code.size() - 1, levels = 0;
  while (max_char) {
    max_char >>= 1;
    ++levels;
  }
  return levels;
}

}  // extern "C"
