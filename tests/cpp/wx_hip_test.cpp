// C++ adapter test: every wx_hip_* class builds the reference's round-trip
// corpus (test/util/util.hpp:90-289, a subset) and seeded random texts; the
// level words and zeros must equal the oracle's (oracle_wx_naive, test
// infrastructure) word for word, and query_support answers must match a
// brute-force scan of the text.  Exit status 0 = all passed.
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pwm_hip_wx.hpp"

extern "C" int oracle_wx_naive(const void* text, uint64_t n, uint32_t width, uint32_t levels, int is_tree,
                               uint64_t* out, uint64_t* zeros);

static int failures = 0;
#define EXPECT(c, ...)                      \
  do {                                      \
    if (!(c)) {                             \
      ++failures;                           \
      std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);             \
      std::printf("\n");                    \
    }                                       \
  } while (0)

template <typename Algo, typename A>
void check_against_oracle(const std::vector<A>& text, uint64_t levels, bool zeros_last_zero, const char* name) {
  auto ws = Algo::compute(text.data(), text.size(), levels);
  const uint64_t n = text.size();
  if (n == 0) {
    EXPECT(ws.levels() == 0, "%s: empty text gives %llu levels", name, (unsigned long long)ws.levels());
    return;
  }
  const uint64_t W = (n + 63) / 64;
  std::vector<uint64_t> ref(W * levels), z(levels);
  oracle_wx_naive(text.data(), n, sizeof(A), uint32_t(levels), Algo::is_tree ? 1 : 0, ref.data(), z.data());
  EXPECT(ws.levels() == levels, "%s: levels %llu", name, (unsigned long long)ws.levels());
  for (uint64_t l = 0; l < levels && l < ws.levels(); ++l) {
    EXPECT(ws.level_bit_size(l) == n, "%s: level %llu size", name, (unsigned long long)l);
    EXPECT(std::memcmp(ws.bvs()[l].data(), ref.data() + l * W, W * 8) == 0, "%s n=%llu level %llu differs", name,
           (unsigned long long)n, (unsigned long long)l);
  }
  if (!Algo::is_tree) {
    auto zz = ws.zeros();
    EXPECT(zz.size() == levels, "%s: zeros size", name);
    for (uint64_t l = 0; l + 1 < levels && l < zz.size(); ++l) EXPECT(zz[l] == z[l], "%s: zeros[%llu]", name, (unsigned long long)l);
    if (zz.size() == levels)
      EXPECT(zz[levels - 1] == (zeros_last_zero ? 0 : z[levels - 1]), "%s: zeros[last]", name);
  }
}

template <bool tree>
void corpus_round() {
  const std::vector<std::string> corpus = {
      std::string("\x00\x01\x06\x07\x01\x05\x04\x02\x06\x03", 10), "abcdefg", "abcdebcdeabc", "a", "", "aaaaaaaaa",
      "banana", "ananas", "abcdefgh#defgh_abcde", "foobar", "abc abc  abc", "ABBCBCABA", "abbbbbbbbbbcbbbbbbbbbb"};
  for (const auto& s : corpus) {
    std::vector<uint8_t> t(s.begin(), s.end());
    check_against_oracle<pwm_hip::wx_hip_ppc<uint8_t, tree>>(t, 8, true, "ppc");
    check_against_oracle<pwm_hip::wx_hip_pps<uint8_t, tree>>(t, 8, false, "pps");
    check_against_oracle<pwm_hip::wx_hip_dd_pc<uint8_t, tree>>(t, 8, false, "dd_pc");
    check_against_oracle<pwm_hip::wx_hip_pc_ss<uint8_t, tree>>(t, 8, true, "pc_ss");
  }
}

int main() {
  static_assert(pwm_hip::wx_hip_ppc<uint8_t, false>::is_parallel && !pwm_hip::wx_hip_pc<uint8_t, false>::is_parallel, "traits");
  static_assert(!pwm_hip::wx_hip_ppc<uint8_t, false>::is_tree, "traits");
  static_assert(pwm_hip::wx_hip_pps<uint8_t, true>::is_tree, "traits");
  static_assert(pwm_hip::wx_hip_huff_pc<uint8_t, false>::is_huffman_shaped, "traits");
  static_assert(pwm_hip::wx_hip_ppc<uint32_t, false>::word_width == 4, "traits");
  corpus_round<false>();
  corpus_round<true>();

  std::mt19937_64 rng(44227);
  for (uint64_t n : {1ull, 4095ull, 4096ull, 4097ull, 100003ull}) {
    std::vector<uint16_t> t(n);
    for (auto& x : t) x = uint16_t(rng() % 1000);
    check_against_oracle<pwm_hip::wx_hip_pc<uint16_t, false>>(t, 10, false, "pc16");
    check_against_oracle<pwm_hip::wx_hip_ps<uint16_t, true>>(t, 10, false, "ps16");
    std::vector<uint32_t> u(n);
    for (auto& x : u) x = uint32_t(rng() % (1u << 20));
    check_against_oracle<pwm_hip::wx_hip_pps<uint32_t, false>>(u, 20, false, "pps32");
  }

  // query_support over a device-resident structure against a text scan
  std::vector<uint8_t> text(50000);
  for (auto& x : text) x = uint8_t(rng() % 5);
  for (bool tree : {false, true}) {
    auto ws = tree ? pwm_hip::wx_hip_ppc<uint8_t, true>::compute(text.data(), text.size(), 3)
                   : pwm_hip::wx_hip_ppc<uint8_t, false>::compute(text.data(), text.size(), 3);
    pwm_hip::query_support qs(ws);
    std::vector<uint32_t> ops;
    std::vector<uint64_t> a, b, want;
    for (int q = 0; q < 3000; ++q) {
      const uint32_t op = q % 3;
      const uint64_t i = rng() % text.size();
      const uint64_t c = text[rng() % text.size()];
      ops.push_back(op);
      if (op == 0) {
        a.push_back(i);
        b.push_back(0);
        want.push_back(text[i]);
      } else if (op == 1) {
        a.push_back(c);
        b.push_back(i);
        uint64_t r = 0;
        for (uint64_t j = 0; j < i; ++j) r += text[j] == c;
        want.push_back(r);
      } else {
        uint64_t occ = 0;
        for (auto x : text) occ += x == c;
        const uint64_t k = 1 + rng() % occ;
        a.push_back(c);
        b.push_back(k);
        uint64_t seen = 0, pos = 0;
        for (uint64_t j = 0; j < text.size(); ++j)
          if (text[j] == c && ++seen == k) {
            pos = j;
            break;
          }
        want.push_back(pos);
      }
    }
    auto got = qs.batch(ops, a, b);
    for (size_t q = 0; q < got.size(); ++q)
      EXPECT(got[q] == want[q], "query %zu op %u: %llu vs %llu", q, ops[q], (unsigned long long)got[q],
             (unsigned long long)want[q]);
  }

  // batched builds equal single builds and stay valid in any release order
  {
    std::vector<std::vector<uint8_t>> texts;
    for (uint64_t n : {0ull, 77ull, 4096ull, 9000ull, 123457ull}) {
      texts.emplace_back(n);
      for (auto& x : texts.back()) x = uint8_t(rng() % 5);
    }
    std::vector<const uint8_t*> ptrs;
    std::vector<uint64_t> sizes;
    for (auto& t : texts) {
      ptrs.push_back(t.data());
      sizes.push_back(t.size());
    }
    auto batch = pwm_hip::compute_batch<uint8_t>(ptrs, sizes, 3, PWMHIP_ALGO_PPC, false);
    // every structure against the oracle, words and zeros (ppc: zeros[2] = 0)
    auto check_one = [&](const pwm_hip::wavelet_structure& ws, size_t i) {
      const uint64_t n = texts[i].size();
      if (n == 0) {
        EXPECT(ws.levels() == 0, "batch[%zu]: empty text", i);
        return;
      }
      const uint64_t W = (n + 63) / 64;
      std::vector<uint64_t> ref(W * 3), z(3);
      oracle_wx_naive(texts[i].data(), n, 1, 3, 0, ref.data(), z.data());
      EXPECT(ws.levels() == 3, "batch[%zu] levels", i);
      for (uint64_t l = 0; l < 3 && l < ws.levels(); ++l)
        EXPECT(std::memcmp(ws.bvs()[l].data(), ref.data() + l * W, W * 8) == 0, "batch[%zu] level %llu", i,
               (unsigned long long)l);
      const auto zz = ws.zeros();
      EXPECT(zz.size() == 3 && zz[0] == z[0] && zz[1] == z[1] && zz[2] == 0, "batch[%zu] zeros", i);
    };
    for (size_t i = 0; i < batch.size(); ++i) check_one(batch[i], i);
    auto last = std::move(batch.back());
    batch.clear();  // the block stays alive through `last`
    check_one(last, texts.size() - 1);
  }

  // errors surface as exceptions with the C-ABI status
  try {
    std::vector<uint8_t> t = {1, 2, 3};
    pwm_hip::wx_hip_ppc<uint8_t, false>::compute(t.data(), t.size(), 0);
    EXPECT(false, "levels == 0 accepted");
  } catch (const pwm_hip::error& e) {
    EXPECT(e.status() == PWMHIP_ERR_INVALID, "status %d", e.status());
  }

  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
