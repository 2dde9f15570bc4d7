"""Independent numpy helpers for the tests: unpack MSB-first level words and
decode wavelet matrices / trees back to the text (the meaning of the
reference's util/decode.hpp:18-97 and huffman/huff_decode.hpp:18-119)."""
import numpy as np


def unpack(words, nbits):
    """MSB-first uint64 words -> uint8 array of nbits bits."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    b = np.unpackbits(words.byteswap().view(np.uint8))
    return b[:nbits]


def pack(bits):
    bits = np.asarray(bits, np.uint8)
    n = bits.size
    pad = (-n) % 64
    b = np.concatenate([bits, np.zeros(pad, np.uint8)])
    return np.packbits(b).view(np.uint64).byteswap() if n else np.zeros(0, np.uint64)


def _rank1_prefix(bits):
    r = np.zeros(bits.size + 1, np.int64)
    np.cumsum(bits, out=r[1:])
    return r


def decode_wm(bvs, zeros, n, levels):
    """Wavelet matrix -> symbols (zeros[l] = zero count of level l)."""
    pos = np.arange(n, dtype=np.int64)
    sym = np.zeros(n, np.uint64)
    for l in range(levels):
        bits = unpack(bvs[l], n)
        r1 = _rank1_prefix(bits)
        b = bits[pos].astype(np.int64)
        sym = (sym << np.uint64(1)) | b.astype(np.uint64)
        z = n - int(r1[-1])
        pos = np.where(b == 1, z + r1[pos], pos - r1[pos])
    return sym


def decode_wt(bvs, n, levels):
    """Wavelet tree (levels in numeric node order) -> symbols."""
    pos = np.arange(n, dtype=np.int64)
    s = np.zeros(n, np.int64)
    e = np.full(n, n, np.int64)
    sym = np.zeros(n, np.uint64)
    for l in range(levels):
        bits = unpack(bvs[l], n)
        r1 = _rank1_prefix(bits)
        i = s + pos
        b = bits[i].astype(np.int64)
        sym = (sym << np.uint64(1)) | b.astype(np.uint64)
        ones_before = r1[i] - r1[s]
        zeros_node = (e - s) - (r1[e] - r1[s])
        new_s = np.where(b == 1, s + zeros_node, s)
        new_e = np.where(b == 1, e, s + zeros_node)
        pos = np.where(b == 1, ones_before, pos - ones_before)
        s, e = new_s, new_e
    return sym


def decode_huff(bvs, bits_per_level, codes, is_tree, n):
    """Huffman-shaped matrix / tree -> symbols using only the structure's code
    table: level l holds the occurrences whose code is longer than l; matrix
    order = continuing zeros then continuing ones of the previous level, tree
    order = numeric order of the (l+1)-bit code prefix."""
    clen, cword = codes
    complete = {}
    for s, (ln, wd) in enumerate(zip(clen, cword)):
        if ln:
            complete[(int(ln), int(wd))] = s
    cur = np.arange(n, dtype=np.int64)
    prefix = np.zeros(n, np.int64)
    out = np.full(n, -1, np.int64)
    for l in range(len(bvs)):
        b = unpack(bvs[l], bits_per_level[l]).astype(np.int64)
        if b.size != cur.size:
            raise ValueError("level %d holds %d bits, expected %d" % (l, b.size, cur.size))
        prefix[cur] = (prefix[cur] << 1) | b
        sym = np.array([complete.get((l + 1, int(p)), -1) for p in prefix[cur]], np.int64)
        done = sym >= 0
        out[cur[done]] = sym[done]
        keep = ~done
        if is_tree:
            c = cur[keep]
            cur = c[np.argsort(prefix[c], kind="stable")]
        else:
            cur = np.concatenate([cur[keep & (b == 0)], cur[keep & (b == 1)]])
    if cur.size:
        raise ValueError("%d occurrences never reached a complete code" % cur.size)
    return out
