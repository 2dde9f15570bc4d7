"""Pins the CPU oracle (oracle/oracle.cpp) -- CPU only.

* against the golden vectors written by the compiled reference
  (tests/golden/make_golden.py);
* against properties the mathematics fixes: decoding every structure gives
  the text back, popcount(level l) == n - zeros[l], rank/select by brute
  force, the reference's own bit_copy known-answer tests
  (test/bit_copy.cpp:45-127);
* against the compiled reference directly where oracle/_ref was built.
"""
import numpy as np
import pytest

import bitutil
import golden_data
import oracle_lib as O

ZEROS_LAST_IS_ZERO = {"ppc", "ppc_ss", "pc_ss", "dd_pc_ss"}


@pytest.fixture(scope="module")
def plain():
    return golden_data.plain_cases()


@pytest.fixture(scope="module")
def huff():
    return golden_data.huff_cases()


def test_oracle_matches_reference_golden(plain):
    assert len(plain) > 150
    for c in plain:
        t, L, tree = c["text"], c["levels"], c["tree"]
        o = O.naive(t, L, tree)
        assert len(o.levels) == len(c["bvs"]), c["key"]
        for x, y in zip(o.levels, c["bvs"]):
            assert np.array_equal(x, y), c["key"]
        for a, z in c["zeros"].items():
            if tree or t.size == 0:
                assert z.size == 0 or (a == "pps" and t.size < 64)
                continue
            if a == "pps" and t.size < 64:
                continue  # the reference's pps is undefined below 64 symbols (pps.hpp:133)
            assert np.array_equal(z[:-1], o.zeros[:-1]), (c["key"], a)
            if a in ZEROS_LAST_IS_ZERO:
                assert int(z[-1]) == 0, (c["key"], a)
            elif a != "pps":  # pps' last entry is racy with >1 thread
                assert int(z[-1]) == int(o.zeros[-1]), (c["key"], a)


def test_oracle_decodes_back(plain):
    for c in plain[::3]:
        t, L, tree = c["text"], c["levels"], c["tree"]
        if t.size == 0:
            continue
        o = O.naive(t, L, tree)
        if tree:
            d = bitutil.decode_wt(o.levels, t.size, L)
        else:
            d = bitutil.decode_wm(o.levels, o.zeros, t.size, L)
            for l in range(L):
                ones = int(bitutil.unpack(o.levels[l], t.size).sum())
                assert ones == t.size - int(o.zeros[l])
        assert np.array_equal(d, t.astype(np.uint64)), c["key"]


def test_decoder_catches_planted_mistake(plain):
    # A transposed level (a plausible construction mistake) must not decode.
    c = next(c for c in plain if not c["tree"] and c["levels"] >= 3 and c["text"].size > 1000)
    t, L = c["text"], c["levels"]
    o = O.naive(t, L, False)
    bad = list(o.levels)
    bad[1], bad[2] = bad[2], bad[1]
    assert not np.array_equal(bitutil.decode_wm(bad, o.zeros, t.size, L), t.astype(np.uint64))


def test_huffman_oracle_matches_reference_golden(huff):
    assert len(huff) > 50
    for c in huff:
        t, tree = c["text"], c["tree"]
        for name, fn in (("huff_pc", O.huff_pc), ("huff_naive", O.huff_naive)):
            g = c["algos"][name]
            o = fn(t, tree)
            assert o.bits == g["bits"], (c["key"], name)
            for x, y in zip(o.levels, g["bvs"]):
                assert np.array_equal(x, y), (c["key"], name)
            assert np.array_equal(o.zeros, g["zeros"]), (c["key"], name)
            assert np.array_equal(o.codes[0], g["codes"][0]) and np.array_equal(o.codes[1], g["codes"][1])
        # huff_ps and huff_pc_ss produce huff_pc's levels in the reference
        for other in ("huff_ps", "huff_pc_ss"):
            g2 = c["algos"][other]
            assert g2["bits"] == c["algos"]["huff_pc"]["bits"]


def test_huff_dd_pc_one_thread_is_huff_pc(huff):
    """The reference's wx_huff_dd_pc at one OpenMP thread (golden_huff_dd.npz)
    equals huff_pc on every golden text: levels, sizes, zeros and codes --
    the oracle's huff_pc restatement therefore pins HUFF_DD_PC too."""
    for c in huff:
        g, d = c["algos"]["huff_pc"], c["algos"]["huff_dd_pc"]
        assert d["bits"] == g["bits"], c["key"]
        for x, y in zip(d["bvs"], g["bvs"]):
            assert np.array_equal(x, y), c["key"]
        assert np.array_equal(d["zeros"], g["zeros"]), c["key"]
        assert np.array_equal(d["codes"][0], g["codes"][0]) and np.array_equal(d["codes"][1], g["codes"][1])
        o = O.huff_pc(c["text"], c["tree"])
        for x, y in zip(o.levels, d["bvs"]):
            assert np.array_equal(x, y), c["key"]
        assert np.array_equal(o.zeros, d["zeros"]), c["key"]


def test_huffman_naive_decodes_back(huff):
    for c in huff[::4]:
        t, tree = c["text"], c["tree"]
        o = O.huff_naive(t, tree)
        d = bitutil.decode_huff(o.levels, o.bits, o.codes, tree, t.size)
        assert np.array_equal(d, t.astype(np.int64)), c["key"]


def test_huffman_codes_are_prefix_free_and_optimal_lengths():
    rng = np.random.default_rng(3)
    for _ in range(20):
        sig = int(rng.integers(2, 200))
        t = rng.choice(sig, 4000, p=rng.dirichlet(np.ones(sig))).astype(np.uint8)
        for tree in (0, 1):
            ln, wd, L, ls = O.huff_codes(t, tree)
            used = [(int(a), int(b)) for a, b in zip(ln, wd) if a]
            strs = sorted(format(w, "0%db" % l) for l, w in used)
            for a, b in zip(strs, strs[1:]):
                assert not b.startswith(a)
            # Kraft equality of a full binary code tree
            assert abs(sum(2.0 ** -l for l, _ in used) - 1.0) < 1e-12
            # Huffman cost equals the classic heap merge cost
            cnt = np.bincount(t, minlength=ln.size)
            cost = sum(int(cnt[s]) * int(ln[s]) for s in range(ln.size))
            import heapq
            h = [int(x) for x in cnt if x]
            heapq.heapify(h)
            tot = 0
            while len(h) > 1:
                a = heapq.heappop(h) + heapq.heappop(h)
                tot += a
                heapq.heappush(h, a)
            assert cost == tot
            assert int(ls[0]) == t.size


def test_rank_select_brute_force():
    rng = np.random.default_rng(44227)
    nb = 5000
    bits = rng.integers(0, 2, nb).astype(np.uint8)
    words = bitutil.pack(bits)
    lib = O.oracle()
    pre = np.concatenate([[0], np.cumsum(bits)])
    for i in list(range(0, 200)) + list(rng.integers(0, nb + 1, 200)):
        assert lib.oracle_rank1(words.ctypes.data, int(i)) == pre[i]
    ones = np.flatnonzero(bits)
    zeros = np.flatnonzero(bits == 0)
    for k in range(1, 100):
        assert lib.oracle_select(words.ctypes.data, nb, 1, k) == ones[k - 1]
        assert lib.oracle_select(words.ctypes.data, nb, 0, k) == zeros[k - 1]
    assert lib.oracle_select(words.ctypes.data, nb, 1, 0) == 0


def test_symbol_queries_brute_force():
    rng = np.random.default_rng(5)
    t = rng.integers(0, 7, 3000).astype(np.uint16)
    q = 500
    ops = rng.integers(0, 3, q).astype(np.uint32)
    a = np.where(ops == 0, rng.integers(0, t.size, q), rng.integers(0, 7, q)).astype(np.uint64)
    b = np.where(ops == 1, rng.integers(0, t.size + 1, q), rng.integers(1, 400, q)).astype(np.uint64)
    r = O.symbol_queries(t, ops, a, b)
    for k in range(q):
        if ops[k] == 0:
            assert r[k] == t[a[k]]
        elif ops[k] == 1:
            assert r[k] == int((t[:b[k]] == a[k]).sum())
        else:
            occ = np.flatnonzero(t == a[k])
            assert r[k] == (occ[b[k] - 1] if b[k] <= occ.size else np.iinfo(np.uint64).max)


# test/bit_copy.cpp:45-127 -- the reference's own known-answer tests
KATS = [
    ([0b00000000, 0b01101011, 0b00000000, 0b00010100, 0b10001101, 0b10011101, 0b00101000],
     [((3, 9, 7), [0b00011010, 0b11000000, 0, 0, 0, 0, 0]),
      ((0, 27, 3), [0b10111010, 0b11000000, 0, 0, 0, 0, 0]),
      ((10, 27, 26), [0b10111010, 0b11101001, 0b00011011, 0b00111010, 0b01010000, 0, 0])]),
    ([0xff] * 7,
     [((3, 9, 25), [0b00011111, 0xff, 0xff, 0b11110000, 0, 0, 0]),
      ((34, 34, 16), [0b00011111, 0xff, 0xff, 0b11110000, 0b00111111, 0xff, 0b11000000])]),
    ([0xff] * 7, [((32, 32, 16), [0, 0, 0, 0, 0xff, 0xff, 0])]),
]


def test_copy_bits_known_answers():
    lib = O.oracle()
    for src, steps in KATS:
        s = np.array(src, np.uint8)
        d = np.zeros(7, np.uint8)
        for (doff, soff, ln), exp in steps:
            lib.oracle_copy_bits8(d.ctypes.data, s.ctypes.data, doff, soff, ln)
            assert list(d) == exp


def test_copy_bits64_matches_bitwise():
    lib = O.oracle()
    rng = np.random.default_rng(9)
    src_bits = rng.integers(0, 2, 2000).astype(np.uint8)
    src = np.concatenate([bitutil.pack(src_bits), np.zeros(2, np.uint64)])
    for _ in range(50):
        so, do, ln = int(rng.integers(0, 900)), int(rng.integers(0, 900)), int(rng.integers(0, 900))
        dst = np.zeros(40, np.uint64)
        lib.oracle_copy_bits64(dst.ctypes.data, src.ctypes.data, do, so, ln)
        got = bitutil.unpack(dst, 40 * 64)
        exp = np.zeros(40 * 64, np.uint8)
        exp[do:do + ln] = src_bits[so:so + ln]
        assert np.array_equal(got, exp)


@pytest.mark.skipif(not O.have_ref(), reason="compiled reference not built")
def test_oracle_against_compiled_reference_random():
    rng = np.random.default_rng(11)
    for _ in range(6):
        L = int(rng.integers(1, 9))
        t = rng.integers(0, 1 << L, int(rng.integers(64, 5000))).astype(np.uint8)
        for tree in (0, 1):
            o = O.naive(t, L, tree)
            r, _ = O.ref_build("ppc", t, L, tree)
            for x, y in zip(o.levels, r.levels):
                assert np.array_equal(x, y)
    # rank1 via the reference's bin_rank_popcnt on sizes not divisible by 2048
    bits = rng.integers(0, 2, 100000).astype(np.uint8)
    w = bitutil.pack(bits)
    pos = rng.integers(0, 100000, 300).astype(np.uint64)
    out = np.zeros(300, np.uint64)
    O.ref().ref_rank1(w.ctypes.data, w.size, pos.ctypes.data, 300, out.ctypes.data)
    lib = O.oracle()
    assert all(out[i] == lib.oracle_rank1(w.ctypes.data, int(pos[i])) for i in range(300))


def test_reduce_alphabet_first_occurrence():
    """Closed form: code = rank of the symbol's first occurrence
    (np.unique's return_index ordered), levels = bit length of distinct-1."""
    rng = np.random.default_rng(11)
    for dt, hi in ((np.uint8, 256), (np.uint16, 40000), (np.uint32, 2**32), (np.uint64, 2**63)):
        t = rng.integers(0, hi, 5000, dtype=np.uint64).astype(dt)
        t[:3] = t[10]  # repeats at the start
        r, levels = O.reduce_alphabet(t)
        vals, first = np.unique(t, return_index=True)
        order = np.argsort(first)
        rank = np.empty(vals.size, np.int64)
        rank[order] = np.arange(vals.size)
        want = rank[np.searchsorted(vals, t)]
        assert np.array_equal(r.astype(np.int64), want)
        assert levels == int(vals.size - 1).bit_length()
    r, levels = O.reduce_alphabet(np.array([255, 0, 255], np.uint8))  # byte 255 handled
    assert list(r) == [0, 1, 0] and levels == 1
    r, levels = O.reduce_alphabet(np.array([7, 7], np.uint8))
    assert list(r) == [0, 0] and levels == 0


def test_select_test_words_match_an_independent_mt19937():
    """The words of the reference's select test (bin_rank_select_tests.cpp:
    42-48) from the oracle equal numpy's own MT19937 seeded the same way
    (init_genrand), two 32-bit outputs per word, high half first."""
    bg = np.random.MT19937()
    bg._legacy_seeding(44227)
    raw = bg.random_raw(2 * 4096).astype(np.uint64)
    want = (raw[0::2] << np.uint64(32)) | raw[1::2]
    assert np.array_equal(O.mt19937_words(44227, 4096), want)
