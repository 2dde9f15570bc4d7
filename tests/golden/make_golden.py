"""Writes the golden fixtures in tests/golden/ from the compiled reference
(oracle/_ref/libpwmref.so, built from kurpicz/pwm's own headers by
`make -C oracle ref`).  Run in the build container:

    OMP_NUM_THREADS=4 python tests/golden/make_golden.py
    OMP_NUM_THREADS=1 python tests/golden/make_golden.py --huff-dd

Fixtures are data only: inputs plus the reference's outputs (level words,
zeros, Huffman codes).  The corner-case strings below are the test data of the
reference's round-trip corpus (test/util/util.hpp:90-289), a subset.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

PLAIN_ALGOS = ["pc", "ps", "ppc", "pps", "pc_ss", "ppc_ss", "dd_pc", "dd_ps", "dd_pc_ss", "naive"]
HUFF_ALGOS = ["huff_pc", "huff_naive", "huff_ps", "huff_pc_ss"]

CORPUS = [
    bytes([0, 1, 6, 7, 1, 5, 4, 2, 6, 3]),
    b"abcdefg", b"abcdebcdeabc", b"a", b"", b"aaaaaaaaa", b"banana", b"ananas",
    b"abcdefgh#defgh_abcde", b"abcdebcdeabcd", b"foobar", b"abcabcabcabc", b"abc abc  abc",
    b"abaaabbababb",
    (b"asdfasctjkcbweasbebvtiwetwcnbwbbqnqxernqze" b"qcrnzxbneqebwcbqwicbqcbtnqweqxcbwuexcbzqwe"
     b"dassdasdfzdfgfsdfsdgfducezctzqwebctuiqwiiq"),
    "ประเทศไทย中华Việt Nam".encode("utf-8"),
    ("Лорэм атоморюм ут хаж, эа граэки емпыдит ёудёкабет мэль, декам дежпютатионй про ты. "
     "Нэ ёужто жэмпэр жкрибэнтур векж, незл коррюмпит.").encode("utf-8"),
    b"struct Foo { uint8_t bar }", b"ABBCBCABA", b"abcabca", b"abbbbbbbbbbcbbbbbbbbbb",
    bytes(range(256)),
]


def levels_for_max_char(m):
    """util/alphabet_util.hpp:99-106"""
    lv = 0
    while m:
        m >>= 1
        lv += 1
    return lv


def cases():
    rng = np.random.default_rng(20240517)
    out = []
    for s in CORPUS:
        t = np.frombuffer(s, np.uint8).copy()
        out.append(("corpus", t, 8))
        if t.size and t.max() > 0:
            out.append(("corpus_reduced_levels", t, levels_for_max_char(int(t.max()))))
    sizes = [1, 2, 63, 64, 65, 127, 128, 129, 1000, 4095, 4096, 4097, 16383, 16384, 16385, 40000]
    for i, n in enumerate(sizes):
        for w, lmax in ((1, 8), (2, 12), (4, 20)):
            L = int(rng.integers(1, lmax + 1))
            dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[w]
            out.append(("uniform", rng.integers(0, 1 << L, n).astype(dt), L))
    # skewed, few-symbol texts (DNA-like)
    for n in (5000, 70001):
        out.append(("dna", rng.choice(5, n, p=[.3, .2, .2, .29, .01]).astype(np.uint8), 3))
    out.append(("zipfish", (rng.zipf(1.3, 30000) % 4096).astype(np.uint32), 12))
    return out


def huff_dd_main():
    """golden_huff_dd.npz: wx_huff_dd_pc (wx_huff_dd_pc.hpp:17-32, huff_dd.hpp:
    29-160) on every text of golden_huff.npz.  Its output depends on the
    OpenMP thread count (one domain per thread); the one-thread result is the
    one the library reproduces, so this must run with OMP_NUM_THREADS=1."""
    assert os.environ.get("OMP_NUM_THREADS") == "1", "run with OMP_NUM_THREADS=1"
    z = np.load(os.path.join(HERE, "golden_huff.npz"))
    out = {}
    for k in sorted(x for x in z.files if x.endswith("_text")):
        key = k[: -len("_text")]
        t = z[k]
        r, _ = O.ref_build("huff_dd_pc", t, 0, key.endswith("_wt"))
        out[key + "_huff_dd_pc_words"] = np.concatenate(r.levels) if r.levels else np.zeros(0, np.uint64)
        out[key + "_huff_dd_pc_bits"] = np.array(r.bits, np.uint64)
        out[key + "_huff_dd_pc_zeros"] = np.array(r.zeros, np.uint64)
        out[key + "_huff_dd_pc_codelen"] = r.codes[0]
        out[key + "_huff_dd_pc_codeword"] = r.codes[1]
    np.savez_compressed(os.path.join(HERE, "golden_huff_dd.npz"), **out)
    print("huff_dd_pc cases:", len(out) // 5)


def main():
    assert O.have_ref(), "build the reference shim first: make -C oracle ref"
    plain = {}
    for ci, (kind, t, L) in enumerate(cases()):
        for tree in (0, 1):
            nv = O.naive(t, L, tree)
            words = np.concatenate(nv.levels) if nv.levels else np.zeros(0, np.uint64)
            zs = []
            for a in PLAIN_ALGOS:
                if t.size < 64 and a in ("pps",):
                    zs.append(np.full(L, np.iinfo(np.uint64).max, np.uint64))  # reference UB (pps.hpp:133)
                    continue
                r, _ = O.ref_build(a, t, L, tree)
                assert len(r.levels) == (L if t.size else 0), (a, kind, L)
                for x, y in zip(r.levels, nv.levels):
                    assert np.array_equal(x, y), ("level bits differ from naive", a, kind, L, tree)
                z = np.array(r.zeros, np.uint64)
                if not tree and t.size:
                    assert np.array_equal(z[:-1], nv.zeros[:-1]), (a, kind)
                zs.append(z if z.size else np.zeros(0, np.uint64))
            key = "c%03d_%s" % (ci, "wt" if tree else "wm")
            plain[key + "_text"] = t
            plain[key + "_levels"] = np.array([L], np.uint64)
            plain[key + "_words"] = words
            for a, z in zip(PLAIN_ALGOS, zs):
                plain[key + "_zeros_" + a] = z
    np.savez_compressed(os.path.join(HERE, "golden_plain.npz"), **plain)

    huff = {}
    rng = np.random.default_rng(77)
    hcases = [t for (_, t, _) in cases() if t.size > 0 and t.dtype == np.uint8]
    for n in (3000, 20000):
        for _ in range(3):
            hcases.append(rng.choice(5, n, p=rng.dirichlet(np.ones(5))).astype(np.uint8))
    hcases.append(np.full(100, 7, np.uint8))
    hcases.append(rng.choice(40, 5000, p=rng.dirichlet(np.ones(40) * .3)).astype(np.uint16))
    kept = 0
    for hi, t in enumerate(hcases):
        for tree in (0, 1):
            try:
                O.huff_pc(t, tree)  # well-defined in the reference?
            except RuntimeError:
                continue  # the reference writes out of bounds here (UB)
            key = "h%03d_%s" % (hi, "wt" if tree else "wm")
            huff[key + "_text"] = t
            for a in HUFF_ALGOS:
                r, _ = O.ref_build(a, t, 0, tree)
                huff[key + "_" + a + "_words"] = (np.concatenate(r.levels) if r.levels
                                                  else np.zeros(0, np.uint64))
                huff[key + "_" + a + "_bits"] = np.array(r.bits, np.uint64)
                huff[key + "_" + a + "_zeros"] = np.array(r.zeros, np.uint64)
                huff[key + "_" + a + "_codelen"] = r.codes[0]
                huff[key + "_" + a + "_codeword"] = r.codes[1]
            kept += 1
    np.savez_compressed(os.path.join(HERE, "golden_huff.npz"), **huff)
    print("plain cases:", len(plain) // (3 + len(PLAIN_ALGOS)), "huffman cases:", kept)


if __name__ == "__main__":
    if "--huff-dd" in sys.argv:
        huff_dd_main()
    else:
        main()
