"""Writes tests/golden/fullsize_digests.json: per-level SHA-256 digests, zero
counts, level sizes and Huffman codes of the reference's own algorithms on the
full-size benchmark texts.

Runs in the build container only (it needs oracle/_ref/libpwmref.so, which is
compiled from the reference's headers by `make -C oracle ref`).  The texts come
from the library's seeded host generator (pwmhip_generate, host code only; no
GPU is used), so the GPU tests regenerate the same bytes and compare digests.

    python tests/golden/make_fullsize_digests.py [--only NAME ...]

Cases (reference algorithm in parentheses; wx_naive is not usable here: its
positions are uint32_t, wx_naive.hpp:45,110):
  dna4g      4 GiB DNA-like uint8 text, 3 levels: WM and WT via wx_ppc and
             wx_pps (wx_ppc.hpp:36-58, wx_pps.hpp:37-68); the two must agree
             on every level; WM Huffman via wx_huff_pc (wx_huff_pc.hpp:45-80)
  zipf256m   2^28 uint32 Zipf text, 20 levels: WM via wx_pps and wx_dd_pc
             (wx_dd_pc.hpp:42-119)
  uniform64m 64 MiB i.i.d. uniform bytes, 8 levels: WM and WT via wx_ppc and
             wx_pps
The reference's zeros[levels-1] differ per algorithm (wx_ppc never sets it);
each algorithm's own vector is stored.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

OUT = os.path.join(HERE, "fullsize_digests.json")

# name: (generator kind, seed, symbols, width, levels, [(algo, tree), ...])
CASES = {
    "dna4g": ("dna", 0xC2, 1 << 32, 1, 3,
              [("ppc", False), ("pps", False), ("ppc", True), ("pps", True), ("huff_pc", False)]),
    "zipf256m": ("zipf", 0xC3, 1 << 28, 4, 20, [("pps", False), ("dd_pc", False)]),
    "uniform64m": ("uniform", 0xC1, 1 << 26, 1, 8, [("ppc", False), ("pps", False), ("ppc", True), ("pps", True)]),
}


def sha(a):
    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()


def level_digests(h, lib, oracle_lib):
    L = int(lib.ref_levels(h))
    bits, digests = [], []
    for l in range(L):
        nb = int(lib.ref_level_bits(h, l))
        w = np.zeros(max(1, oracle_lib.words_for(nb)), np.uint64)
        lib.ref_level_words(h, l, oracle_lib._ptr(w))
        digests.append(sha(w[:oracle_lib.words_for(nb)]))
        bits.append(nb)
        del w
    return bits, digests


def run_case(name, oracle_lib, pwm_amd):
    kind, seed, n, width, levels, runs = CASES[name]
    t0 = time.time()
    text = pwm_amd.generate(kind, seed, n, width, levels)
    print("%s: generated %d symbols in %.1f s" % (name, n, time.time() - t0), flush=True)
    entry = {"generator": kind, "seed": seed, "symbols": n, "width": width, "levels": levels,
             "text_sha256": sha(text), "builds": {}}
    if width == 1:
        hist = np.bincount(text, minlength=256)
    else:
        hist = None
    if hist is not None:
        entry["histogram"] = [int(x) for x in hist[: int(np.flatnonzero(hist).max()) + 1]]
    lib = oracle_lib.ref()
    import ctypes
    for algo, tree in runs:
        secs = ctypes.c_double(0)
        t0 = time.time()
        h = lib.ref_build(algo.encode(), width, int(tree), oracle_lib._ptr(text), n,
                          0 if algo.startswith("huff") else levels, ctypes.byref(secs))
        try:
            bits, digests = level_digests(h, lib, oracle_lib)
            nz = int(lib.ref_zeros(h, None))
            z = np.zeros(max(1, nz), np.uint64)
            lib.ref_zeros(h, oracle_lib._ptr(z))
            rec = {"level_bits": bits, "level_sha256": digests, "zeros": [int(x) for x in z[:nz]],
                   "compute_s": round(secs.value, 2)}
            if algo.startswith("huff"):
                nc = lib.ref_codes(h, None, None, 0)
                ln = np.zeros(max(1, nc), np.uint64)
                wd = np.zeros(max(1, nc), np.uint64)
                lib.ref_codes(h, oracle_lib._ptr(ln), oracle_lib._ptr(wd), nc)
                rec["code_len"] = [int(x) for x in ln[:nc]]
                rec["code_word"] = [int(x) for x in wd[:nc]]
        finally:
            lib.ref_free(h)
        key = ("wt_" if tree else "wm_") + algo
        entry["builds"][key] = rec
        print("%s %s: %.1f s (compute %.1f s), zeros %s" % (name, key, time.time() - t0, secs.value,
                                                             rec["zeros"][:4]), flush=True)
    # all plain algorithms of one shape give the same level bits
    for shape in ("wm_", "wt_"):
        keys = [k for k in entry["builds"] if k.startswith(shape) and "huff" not in k]
        for k in keys[1:]:
            assert entry["builds"][k]["level_sha256"] == entry["builds"][keys[0]]["level_sha256"], (name, k)
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    import oracle_lib
    import pwm_amd
    if not oracle_lib.have_ref():
        sys.exit("oracle/_ref/libpwmref.so is missing: run `make -C oracle ref` in the build container")
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    for name in (args.only or list(CASES)):
        out[name] = run_case(name, oracle_lib, pwm_amd)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
