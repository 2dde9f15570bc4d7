"""Loads the committed golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PLAIN_ALGOS = ["pc", "ps", "ppc", "pps", "pc_ss", "ppc_ss", "dd_pc", "dd_ps", "dd_pc_ss", "naive"]
HUFF_ALGOS = ["huff_pc", "huff_naive", "huff_ps", "huff_pc_ss"]


def plain_cases():
    z = np.load(os.path.join(HERE, "golden_plain.npz"))
    keys = sorted({k.rsplit("_text", 1)[0] for k in z.files if k.endswith("_text")})
    out = []
    for k in keys:
        t = z[k + "_text"]
        L = int(z[k + "_levels"][0])
        words = z[k + "_words"]
        W = (t.size + 63) // 64
        lv = [words[l * W:(l + 1) * W] for l in range(L)] if t.size else []
        zeros = {a: z[k + "_zeros_" + a] for a in PLAIN_ALGOS}
        out.append(dict(key=k, text=t, levels=L, tree=k.endswith("_wt"), bvs=lv, zeros=zeros))
    return out


def huff_cases():
    """Huffman cases; algos holds huff_pc, huff_naive, huff_ps, huff_pc_ss and
    huff_dd_pc (the last written by the reference at one OpenMP thread)."""
    z = np.load(os.path.join(HERE, "golden_huff.npz"))
    zd = np.load(os.path.join(HERE, "golden_huff_dd.npz"))
    keys = sorted({k.rsplit("_text", 1)[0] for k in z.files if k.endswith("_text")})
    out = []
    for k in keys:
        t = z[k + "_text"]
        d = dict(key=k, text=t, tree=k.endswith("_wt"), algos={})
        for a, src in [(a, z) for a in HUFF_ALGOS] + [("huff_dd_pc", zd)]:
            z_ = src
            bits = [int(b) for b in z_[k + "_" + a + "_bits"]]
            words = z_[k + "_" + a + "_words"]
            lv, off = [], 0
            for b in bits:
                nw = (b + 63) // 64
                lv.append(words[off:off + nw])
                off += nw
            d["algos"][a] = dict(bvs=lv, bits=bits, zeros=z_[k + "_" + a + "_zeros"],
                                 codes=(z_[k + "_" + a + "_codelen"], z_[k + "_" + a + "_codeword"]))
        out.append(d)
    return out
