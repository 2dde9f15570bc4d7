"""ctypes access to the CPU checker (oracle/liboracle.so) and, where it was
built, the compiled reference (oracle/_ref/libpwmref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpwmref.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(vp)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_wx_naive.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_int, vp, vp]
        lib.oracle_huff_codes.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                                          vp, vp, vp, vp, vp]
        lib.oracle_huff_naive.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, vp, vp]
        lib.oracle_huff_pc.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, vp, vp]
        lib.oracle_rank1.argtypes = [vp, ctypes.c_uint64]
        lib.oracle_rank1.restype = ctypes.c_uint64
        lib.oracle_select.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64]
        lib.oracle_select.restype = ctypes.c_uint64
        lib.oracle_symbol_queries.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                              vp, vp, vp, vp]
        lib.oracle_copy_bits64.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.oracle_copy_bits8.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.oracle_reduce_alphabet.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32]
        lib.oracle_reduce_alphabet.restype = ctypes.c_uint64
        lib.oracle_mt19937_words.argtypes = [ctypes.c_uint64, ctypes.c_uint64, vp]
        _oracle = lib
    return _oracle


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.ref_build.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, vp, ctypes.c_uint64,
                                  ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        lib.ref_build.restype = vp
        lib.ref_free.argtypes = [vp]
        lib.ref_levels.argtypes = [vp]
        lib.ref_levels.restype = ctypes.c_uint64
        lib.ref_level_bits.argtypes = [vp, ctypes.c_uint64]
        lib.ref_level_bits.restype = ctypes.c_uint64
        lib.ref_level_words.argtypes = [vp, ctypes.c_uint64, vp]
        lib.ref_zeros.argtypes = [vp, vp]
        lib.ref_zeros.restype = ctypes.c_uint64
        lib.ref_codes.argtypes = [vp, vp, vp, ctypes.c_uint64]
        lib.ref_rank1.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint64, vp]
        lib.ref_select.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, vp, ctypes.c_uint64, vp]
        lib.ref_copy_bits8.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        _ref = lib
    return _ref


def words_for(bits):
    return (bits + 63) // 64


class Structure:
    """Plain host-side result: per-level uint64 word arrays, level bit sizes,
    zeros and (Huffman) codes."""

    def __init__(self, levels, bits, zeros, codes=None):
        self.levels = levels  # list of np.uint64 arrays
        self.bits = bits      # list of level bit sizes
        self.zeros = zeros    # np.uint64 array (matrix) or empty
        self.codes = codes    # (len, word) arrays or None


def naive(text, levels, is_tree):
    text = np.ascontiguousarray(text)
    n = text.size
    W = words_for(n)
    out = np.zeros(max(1, levels * W), dtype=np.uint64)
    zeros = np.zeros(max(1, levels), dtype=np.uint64)
    if n:
        oracle().oracle_wx_naive(_ptr(text), n, text.itemsize, levels, int(is_tree), _ptr(out), _ptr(zeros))
    lv = [out[l * W:(l + 1) * W].copy() for l in range(levels)] if n else []
    return Structure(lv, [n] * len(lv), np.array([], np.uint64) if (is_tree or not n) else zeros[:levels].copy())


def huff_codes(text, is_tree):
    text = np.ascontiguousarray(text)
    ns = ctypes.c_uint64(0)
    lv = ctypes.c_uint64(0)
    lib = oracle()
    lib.oracle_huff_codes(_ptr(text), text.size, text.itemsize, int(is_tree), ctypes.byref(ns),
                          ctypes.byref(lv), None, None, None)
    ln = np.zeros(max(1, ns.value), np.uint64)
    wd = np.zeros(max(1, ns.value), np.uint64)
    ls = np.zeros(max(1, lv.value), np.uint64)
    lib.oracle_huff_codes(_ptr(text), text.size, text.itemsize, int(is_tree), ctypes.byref(ns),
                          ctypes.byref(lv), _ptr(ln), _ptr(wd), _ptr(ls))
    return ln[:ns.value], wd[:ns.value], int(lv.value), ls[:lv.value]


def _huff(fn, text, is_tree):
    text = np.ascontiguousarray(text)
    ln, wd, L, ls = huff_codes(text, is_tree)
    offs = [0]
    for s in ls:
        offs.append(offs[-1] + words_for(int(s)))
    out = np.zeros(max(1, offs[-1]), np.uint64)
    zeros = np.zeros(max(1, L), np.uint64)
    rc = fn(_ptr(text), text.size, text.itemsize, int(is_tree), _ptr(out), _ptr(zeros))
    if rc != 0:
        raise RuntimeError("oracle huffman construction returned %d" % rc)
    lv = [out[offs[l]:offs[l + 1]].copy() for l in range(L)] if text.size else []
    return Structure(lv, [int(s) for s in ls] if text.size else [],
                     np.array([], np.uint64) if is_tree else zeros[:L].copy(), (ln, wd))


def huff_pc(text, is_tree):
    return _huff(oracle().oracle_huff_pc, text, is_tree)


def huff_naive(text, is_tree):
    return _huff(oracle().oracle_huff_naive, text, is_tree)


def symbol_queries(text, ops, a, b):
    text = np.ascontiguousarray(text)
    ops = np.ascontiguousarray(ops, np.uint32)
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    res = np.zeros(max(1, ops.size), np.uint64)
    oracle().oracle_symbol_queries(_ptr(text), text.size, text.itemsize, ops.size, _ptr(ops), _ptr(a),
                                   _ptr(b), _ptr(res))
    return res[:ops.size]


def ref_build(algo, text, levels, is_tree):
    """Run the compiled reference's wx_<algo>::compute; returns (Structure, seconds)."""
    text = np.ascontiguousarray(text)
    lib = ref()
    secs = ctypes.c_double(0)
    h = lib.ref_build(algo.encode(), text.itemsize, int(is_tree), _ptr(text), text.size, levels,
                      ctypes.byref(secs))
    if not h:
        raise RuntimeError("compiled reference wx_%s::compute threw" % algo)
    try:
        L = lib.ref_levels(h)
        lv, bits = [], []
        for l in range(L):
            nb = lib.ref_level_bits(h, l)
            w = np.zeros(max(1, words_for(nb)), np.uint64)
            lib.ref_level_words(h, l, _ptr(w))
            lv.append(w[:words_for(nb)])
            bits.append(int(nb))
        nz = lib.ref_zeros(h, None)
        zeros = np.zeros(max(1, nz), np.uint64)
        lib.ref_zeros(h, _ptr(zeros))
        codes = None
        if algo.startswith("huff"):
            nc = lib.ref_codes(h, None, None, 0)
            ln = np.zeros(max(1, nc), np.uint64)
            wd = np.zeros(max(1, nc), np.uint64)
            lib.ref_codes(h, _ptr(ln), _ptr(wd), nc)
            codes = (ln[:nc], wd[:nc])
        return Structure(lv, bits, zeros[:nz], codes), secs.value
    finally:
        lib.ref_free(h)


def reduce_alphabet(text):
    """(reduced copy, levels): first-occurrence renumbering (alphabet_util.hpp:18-45)."""
    t = np.array(text, copy=True)
    levels = oracle().oracle_reduce_alphabet(_ptr(t), t.size, t.itemsize)
    return t, int(levels)


def mt19937_words(seed, n_words):
    """Words of the reference's select test: uniform_int_distribution<uint64_t>
    over std::mt19937(seed) (test/bin_rank_select_tests.cpp:42-48)."""
    out = np.zeros(max(1, n_words), np.uint64)
    oracle().oracle_mt19937_words(seed, n_words, _ptr(out))
    return out[:n_words]
