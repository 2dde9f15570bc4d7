// oracle/ref_shim.cpp -- a C-ABI shim that instantiates kurpicz/pwm's own
// header-only construction algorithms, compiled from the reference sources
// where they lie (oracle/Makefile, target `ref`, -I$(REF)/include).
//
// TEST INFRASTRUCTURE ONLY: it produces golden vectors (tests/golden/) and is
// the CPU baseline in bench.py ("kind": "reference").  The output goes to
// oracle/_ref/ and is never part of the product.
//
// The shim contains no reference code: it only calls
//   wx_*<A, is_tree>::compute(text, size, levels)          (include/wx_*.hpp)
//   query support bin_rank_popcnt / bin_select_popcnt      (include/queries/)
//   copy_bits<uint8_t>                                      (construction/merge.hpp)

#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "wx_dd_pc.hpp"
#include "wx_dd_pc_ss.hpp"
#include "wx_dd_ps.hpp"
#include "wx_huff_dd_pc.hpp"
#include "wx_huff_naive.hpp"
#include "wx_huff_pc.hpp"
#include "wx_huff_pc_ss.hpp"
#include "wx_huff_ps.hpp"
#include "wx_naive.hpp"
#include "wx_pc.hpp"
#include "wx_pc_ss.hpp"
#include "wx_ppc.hpp"
#include "wx_ppc_ss.hpp"
#include "wx_pps.hpp"
#include "wx_ps.hpp"
#include "queries/bin_rank_popcnt.hpp"
#include "queries/bin_select_popcnt.hpp"
#include "construction/merge.hpp"

namespace {

struct handle {
  wavelet_structure ws;
  int width;
  bool tree;
  bool huff;
  explicit handle(wavelet_structure&& w) : ws(std::move(w)) {}
};

template <template <typename, bool> class Algo, typename A>
wavelet_structure run_algo(bool tree, const void* text, uint64_t n, uint64_t levels) {
  const A* t = static_cast<const A*>(text);
  if (tree) return Algo<A, true>::compute(t, n, levels);
  return Algo<A, false>::compute(t, n, levels);
}

template <template <typename, bool> class Algo>
wavelet_structure run_width(int width, bool tree, const void* text, uint64_t n,
                            uint64_t levels) {
  switch (width) {
    case 1: return run_algo<Algo, uint8_t>(tree, text, n, levels);
    case 2: return run_algo<Algo, uint16_t>(tree, text, n, levels);
    case 4: return run_algo<Algo, uint32_t>(tree, text, n, levels);
    default: return run_algo<Algo, uint64_t>(tree, text, n, levels);
  }
}

template <typename A, bool T>
int codes_of(const wavelet_structure& ws, uint64_t* len, uint64_t* word, uint64_t cap) {
  const auto& cp = ws.codes<A, T>().code_pairs();
  if (len == nullptr) return int(cp.size());
  for (uint64_t i = 0; i < cp.size() && i < cap; ++i) {
    len[i] = cp[i].code_length();
    word[i] = cp[i].code_word();
  }
  return int(cp.size());
}

}  // namespace

extern "C" {

// Returns an opaque handle (an empty structure for an unknown algorithm
// name), or nullptr if the reference threw (its allocations can).
// *seconds receives the wall time of compute() alone (steady_clock).
void* ref_build(const char* algo, int width, int is_tree, const void* text,
                uint64_t n, uint64_t levels, double* seconds) try {
  const std::string a(algo);
  const bool tree = is_tree != 0;
  auto t0 = std::chrono::steady_clock::now();
  wavelet_structure ws = [&]() -> wavelet_structure {
    if (a == "pc") return run_width<wx_pc>(width, tree, text, n, levels);
    if (a == "ps") return run_width<wx_ps>(width, tree, text, n, levels);
    if (a == "ppc") return run_width<wx_ppc>(width, tree, text, n, levels);
    if (a == "pps") return run_width<wx_pps>(width, tree, text, n, levels);
    if (a == "pc_ss") return run_width<wx_pc_ss>(width, tree, text, n, levels);
    if (a == "ppc_ss") return run_width<wx_ppc_ss>(width, tree, text, n, levels);
    if (a == "dd_pc") return run_width<wx_dd_pc>(width, tree, text, n, levels);
    if (a == "dd_ps") return run_width<wx_dd_ps>(width, tree, text, n, levels);
    if (a == "dd_pc_ss") return run_width<wx_dd_pc_ss>(width, tree, text, n, levels);
    if (a == "naive") return run_width<wx_naive>(width, tree, text, n, levels);
    if (a == "huff_pc") return run_width<wx_huff_pc>(width, tree, text, n, levels);
    if (a == "huff_ps") return run_width<wx_huff_ps>(width, tree, text, n, levels);
    if (a == "huff_pc_ss") return run_width<wx_huff_pc_ss>(width, tree, text, n, levels);
    if (a == "huff_naive") return run_width<wx_huff_naive>(width, tree, text, n, levels);
    if (a == "huff_dd_pc") return run_width<wx_huff_dd_pc>(width, tree, text, n, levels);
    return wavelet_structure(wavelet_structure_tree());
  }();
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  auto* h = new handle(std::move(ws));
  h->width = width;
  h->tree = tree;
  h->huff = a.rfind("huff", 0) == 0;
  return h;
} catch (...) {
  return nullptr;
}

void ref_free(void* h) { delete static_cast<handle*>(h); }

uint64_t ref_levels(void* h) { return static_cast<handle*>(h)->ws.levels(); }

uint64_t ref_level_bits(void* h, uint64_t l) {
  return static_cast<handle*>(h)->ws.bvs().level_bit_size(l);
}

void ref_level_words(void* h, uint64_t l, uint64_t* out) {
  auto span = static_cast<handle*>(h)->ws.bvs()[l];
  std::memcpy(out, span.data(), span.size() * sizeof(uint64_t));
}

uint64_t ref_zeros(void* h, uint64_t* out) {
  const auto& z = static_cast<handle*>(h)->ws.zeros();
  if (out) std::memcpy(out, z.data(), z.size() * sizeof(uint64_t));
  return z.size();
}

// Huffman code pairs per symbol (count returned; arrays filled if non-null).
int ref_codes(void* hp, uint64_t* len, uint64_t* word, uint64_t cap) {
  auto* h = static_cast<handle*>(hp);
  if (!h->huff) return -1;
  switch (h->width) {
    case 1: return h->tree ? codes_of<uint8_t, true>(h->ws, len, word, cap)
                           : codes_of<uint8_t, false>(h->ws, len, word, cap);
    case 2: return h->tree ? codes_of<uint16_t, true>(h->ws, len, word, cap)
                           : codes_of<uint16_t, false>(h->ws, len, word, cap);
    case 4: return h->tree ? codes_of<uint32_t, true>(h->ws, len, word, cap)
                           : codes_of<uint32_t, false>(h->ws, len, word, cap);
    default: return h->tree ? codes_of<uint64_t, true>(h->ws, len, word, cap)
                            : codes_of<uint64_t, false>(h->ws, len, word, cap);
  }
}

// rank1 over a bit vector through bin_rank_popcnt, for positions `pos`.
void ref_rank1(const uint64_t* words, uint64_t n_words, const uint64_t* pos,
               uint64_t q, uint64_t* out) {
  bin_rank_popcnt rs(span<uint64_t const>(words, n_words));
  for (uint64_t i = 0; i < q; ++i) out[i] = rs.rank1(pos[i]);
}

// select through bin_select_popcnt<value>, for ranks `k`.
void ref_select(const uint64_t* words, uint64_t n_words, int value,
                const uint64_t* k, uint64_t q, uint64_t* out) {
  bin_rank_popcnt rs(span<uint64_t const>(words, n_words));
  if (value) {
    bin_select_popcnt<1> ss(rs);
    for (uint64_t i = 0; i < q; ++i) out[i] = ss.select(k[i]);
  } else {
    bin_select_popcnt<0> ss(rs);
    for (uint64_t i = 0; i < q; ++i) out[i] = ss.select(k[i]);
  }
}

void ref_copy_bits8(uint8_t* dst, const uint8_t* src, uint64_t dst_off,
                    uint64_t src_off, uint64_t len) {
  copy_bits<uint8_t>(dst, src, dst_off, src_off, len);
}

}  // extern "C"
