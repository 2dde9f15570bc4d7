// pwm_amd/csrc/gen.cpp -- seeded synthetic texts for the benchmarks
// (host side, multi-threaded).  Texts are generated in independent blocks of
// 2^20 symbols, each with its own generator state derived from (seed, block),
// so results do not depend on the number of threads.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/pwm_hip.h"

namespace {

constexpr uint64_t kBlock = 1ull << 20;

inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct xoshiro256ss {
  uint64_t s[4];
  explicit xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double unit() { return double(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

template <typename F>
void parallel_blocks(uint64_t n, F&& fn) {
  const uint64_t nb = (n + kBlock - 1) / kBlock;
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  nt = unsigned(std::min<uint64_t>(nt, nb));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (uint64_t b = t; b < nb; b += nt) fn(b, b * kBlock, std::min(n, (b + 1) * kBlock));
    });
  for (auto& x : th) x.join();
}

inline void store(void* out, uint32_t w, uint64_t i, uint64_t v) {
  switch (w) {
    case 1: static_cast<uint8_t*>(out)[i] = uint8_t(v); break;
    case 2: static_cast<uint16_t*>(out)[i] = uint16_t(v); break;
    case 4: static_cast<uint32_t*>(out)[i] = uint32_t(v); break;
    default: static_cast<uint64_t*>(out)[i] = v; break;
  }
}

// Order-2 Markov chain over A,C,G,T (codes 0..3) with Dirichlet(1) rows,
// N (code 4) runs starting with probability 1e-5 per position, geometric
// length with mean 500.
void gen_dna(uint64_t seed, uint64_t n, uint32_t w, void* out) {
  xoshiro256ss g(seed);
  uint32_t thr[16][3];  // cumulative thresholds of the first three bases
  for (int ctx = 0; ctx < 16; ++ctx) {
    double e[4], sum = 0;
    for (double& v : e) {
      v = -std::log(1.0 - g.unit());
      sum += v;
    }
    double acc = 0;
    for (int k = 0; k < 3; ++k) {
      acc += e[k] / sum;
      thr[ctx][k] = uint32_t(std::min(acc, 1.0) * 4294967295.0);
    }
  }
  const uint64_t n_start = uint64_t(1e-5 * 18446744073709551616.0);
  const double log_q = std::log(1.0 - 1.0 / 500.0);
  parallel_blocks(n, [&](uint64_t b, uint64_t lo, uint64_t hi) {
    uint64_t st = seed ^ (0xD1B54A32D192ED03ull * (b + 1));
    xoshiro256ss r(splitmix64(st));
    uint32_t p2 = uint32_t(r.next() & 3), p1 = uint32_t(r.next() & 3);
    uint64_t i = lo;
    while (i < hi) {
      const uint64_t u = r.next();
      if (u < n_start) {
        const double x = 1.0 - r.unit();
        uint64_t len = 1 + uint64_t(std::floor(std::log(x) / log_q));
        const uint64_t end = std::min(hi, i + len);
        for (; i < end; ++i) store(out, w, i, 4);
        continue;
      }
      const uint32_t v = uint32_t(u >> 32);
      const uint32_t* t = thr[p2 * 4 + p1];
      const uint32_t c = v < t[0] ? 0 : (v < t[1] ? 1 : (v < t[2] ? 2 : 3));
      store(out, w, i++, c);
      p2 = p1;
      p1 = c;
    }
  });
}

void gen_uniform(uint64_t seed, uint64_t n, uint32_t w, uint32_t levels, void* out) {
  const uint64_t mask = levels >= 64 ? ~0ull : ((1ull << levels) - 1);
  parallel_blocks(n, [&](uint64_t b, uint64_t lo, uint64_t hi) {
    uint64_t st = seed ^ (0xA0761D6478BD642Full * (b + 1));
    if (w == 1 && levels == 8) {
      uint8_t* o = static_cast<uint8_t*>(out);
      for (uint64_t i = lo; i < hi; i += 8) {
        uint64_t v = splitmix64(st);
        for (uint64_t k = 0; k < 8 && i + k < hi; ++k) o[i + k] = uint8_t(v >> (8 * k));
      }
      return;
    }
    for (uint64_t i = lo; i < hi; ++i) store(out, w, i, splitmix64(st) & mask);
  });
}

// Zipf(s = 1.1) over ranks 1..K (K = 2^levels) through an integer alias table,
// ranks mapped to symbols by a seeded Fisher-Yates permutation.
void gen_zipf(uint64_t seed, uint64_t n, uint32_t w, uint32_t levels, void* out) {
  const uint64_t K = 1ull << levels;
  std::vector<double> p(K);
  double sum = 0;
  for (uint64_t k = 0; k < K; ++k) sum += (p[k] = std::pow(double(k + 1), -1.1));
  std::vector<uint32_t> alias(K), prob(K);
  std::vector<double> q(K);
  std::vector<uint64_t> small, large;
  for (uint64_t k = 0; k < K; ++k) {
    q[k] = p[k] / sum * double(K);
    (q[k] < 1.0 ? small : large).push_back(k);
  }
  while (!small.empty() && !large.empty()) {
    const uint64_t s = small.back(), l = large.back();
    small.pop_back();
    prob[s] = uint32_t(std::min(q[s], 1.0) * 4294967295.0);
    alias[s] = uint32_t(l);
    q[l] -= 1.0 - q[s];
    if (q[l] < 1.0) {
      large.pop_back();
      small.push_back(l);
    }
  }
  for (uint64_t k : large) prob[k] = 0xFFFFFFFFu, alias[k] = uint32_t(k);
  for (uint64_t k : small) prob[k] = 0xFFFFFFFFu, alias[k] = uint32_t(k);
  std::vector<uint32_t> perm(K);
  for (uint64_t k = 0; k < K; ++k) perm[k] = uint32_t(k);
  xoshiro256ss g(seed);
  for (uint64_t k = K - 1; k > 0; --k) std::swap(perm[k], perm[g.next() % (k + 1)]);
  parallel_blocks(n, [&](uint64_t b, uint64_t lo, uint64_t hi) {
    uint64_t st = seed ^ (0xE7037ED1A0B428DBull * (b + 1));
    for (uint64_t i = lo; i < hi; ++i) {
      const uint64_t u = splitmix64(st);
      const uint64_t k = (u & 0xFFFFFFFFull) % K;
      const uint32_t c = uint32_t(u >> 32) <= prob[k] ? uint32_t(k) : alias[k];
      store(out, w, i, perm[c]);
    }
  });
}

}  // namespace

extern "C" int pwmhip_generate(uint32_t kind, uint64_t seed, uint64_t size, uint32_t word_width,
                               uint32_t levels, void* host_out) {
  if (size && !host_out) return PWMHIP_ERR_INVALID;
  if (word_width != 1 && word_width != 2 && word_width != 4 && word_width != 8) return PWMHIP_ERR_INVALID;
  try {
    switch (kind) {
      case 0:
        if (levels == 0 || levels > 8 * word_width) return PWMHIP_ERR_INVALID;
        gen_uniform(seed, size, word_width, levels, host_out);
        return PWMHIP_OK;
      case 1: gen_dna(seed, size, word_width, host_out); return PWMHIP_OK;
      case 2:
        if (levels == 0 || levels > 24 || levels > 8 * word_width) return PWMHIP_ERR_INVALID;
        gen_zipf(seed, size, word_width, levels, host_out);
        return PWMHIP_OK;
      default: return PWMHIP_ERR_INVALID;
    }
  } catch (...) {
    return PWMHIP_ERR_OOM;
  }
}
