"""Full-size parity: the benchmark workloads built on the GPU, compared with the
reference's own algorithms on the same texts through per-level SHA-256
digests, zero counts, level sizes and Huffman codes
(tests/golden/fullsize_digests.json, written by make_fullsize_digests.py from
the compiled reference: wx_ppc / wx_pps / wx_huff_pc on the 4 GiB DNA text,
wx_pps / wx_dd_pc on the 2^28 x uint32 Zipf text, wx_ppc / wx_pps on the
64 MiB uniform text).  The 4 GiB levels are 2^32 bits long, where wx_naive is
not usable (uint32_t positions, wx_naive.hpp:45,110).  Queries on the DNA
matrix are checked by brute force over the text up to position 2^32."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

pwm_amd = pytest.importorskip("pwm_amd")
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "fullsize_digests.json")) as f:
    DIG = json.load(f)


def sha(a):
    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()


def _text(name):
    c = DIG[name]
    print("generating %s: %d symbols" % (name, c["symbols"]), flush=True)
    t = pwm_amd.generate(c["generator"], c["seed"], c["symbols"], c["width"], c["levels"])
    assert sha(t) == c["text_sha256"], "generator output differs from the text the digests were made on"
    return t


def _check(ws, rec, what):
    L = len(rec["level_bits"])
    assert ws.levels() == L, what
    assert [ws.level_bit_size(l) for l in range(L)] == rec["level_bits"], what
    for l in range(L):
        nw = (ws.level_bit_size(l) + 63) // 64
        a = np.empty(nw, np.uint64)
        pwm_amd._check(pwm_amd.lib().pwmhip_ws_download(ws._ctx.h, ctypes.byref(ws._ws), l,
                                                        a.ctypes.data_as(ctypes.c_void_p)), "download")
        assert sha(a) == rec["level_sha256"][l], (what, "level", l)
        del a
    assert [int(z) for z in ws.zeros()] == rec["zeros"], (what, list(ws.zeros()), rec["zeros"])


@pytest.fixture(scope="module")
def dna4g():
    t = _text("dna4g")
    d = torch.from_numpy(t).to("cuda:0")
    torch.cuda.synchronize()
    yield t, d
    del d
    torch.cuda.empty_cache()


@pytest.mark.gpu
@pytest.mark.parametrize("shape,algo", [("wm", "ppc"), ("wm", "pps"), ("wt", "ppc"), ("wt", "pps")])
def test_dna4g_plain_fullsize(dna4g, shape, algo):
    _, d = dna4g
    rec = DIG["dna4g"]["builds"]["%s_%s" % (shape, algo)]
    ws = pwm_amd.build(d, 3, algo, shape == "wt")
    _check(ws, rec, (shape, algo))
    ws.free()


@pytest.mark.gpu
def test_dna4g_huffman_fullsize(dna4g):
    _, d = dna4g
    rec = DIG["dna4g"]["builds"]["wm_huff_pc"]
    ws = pwm_amd.build(d, 0, "huff_pc", False)
    ln, wd = ws.codes()
    assert [int(x) for x in ln] == rec["code_len"] and [int(x) for x in wd] == rec["code_word"]
    _check(ws, rec, "wm_huff_pc")
    ws.free()


@pytest.mark.gpu
def test_dna4g_queries_up_to_2_32(dna4g):
    """access / rank / select on the 4 GiB matrix (n = 2^32: positions up to
    2^32 - 1, rank up to 2^32, where the reference's bin_rank_popcnt stops
    being valid, bin_rank_popcnt.hpp:54-86), against the text
    (query_support.hpp:37-69 semantics)."""
    t, d = dna4g
    n = t.size
    ws = pwm_amd.build(d, 3, "ppc", False)
    qs = pwm_amd.QuerySupport(ws)
    rng = np.random.default_rng(0xC2)
    pos = np.concatenate([rng.integers(0, n, 4000), np.arange(n - 3000, n),
                          np.arange((1 << 31) - 70, (1 << 31) + 70), [0]]).astype(np.uint64)
    pos = np.unique(pos)
    got = qs.query(np.zeros(pos.size, np.uint32), pos)
    assert np.array_equal(got, t[pos].astype(np.uint64))
    # rank(c, i) for every symbol at sorted positions (and at n): counts
    # between consecutive positions, one pass over the text
    rpos = np.concatenate([pos, [n]]).astype(np.uint64)
    counts = np.zeros((rpos.size, 5), np.uint64)
    acc = np.zeros(5, np.uint64)
    prev = 0
    for k, p in enumerate(rpos.tolist()):
        acc += np.bincount(t[prev:p], minlength=5)[:5].astype(np.uint64)
        counts[k] = acc
        prev = p
    for c in range(5):
        r = qs.query(np.ones(rpos.size, np.uint32), np.full(rpos.size, c, np.uint64), rpos)
        assert np.array_equal(r, counts[:, c]), c
    # select(T[p], rank(T[p], p) + 1) == p, and one past the last occurrence
    sym = t[pos].astype(np.int64)
    k = counts[np.arange(pos.size), sym] + np.uint64(1)
    s = qs.query(np.full(pos.size, 2, np.uint32), sym.astype(np.uint64), k)
    assert np.array_equal(s, pos)
    last = qs.query(np.full(5, 2, np.uint32), np.arange(5, dtype=np.uint64), counts[-1] + np.uint64(1))
    assert np.all(last == np.uint64(2 ** 64 - 1))
    qs.free()
    ws.free()


@pytest.mark.gpu
def test_zipf256m_fullsize():
    t = _text("zipf256m")
    d = torch.from_numpy(t.view(np.int32)).to("cuda:0")  # (same bytes, 4-byte symbols)
    for algo in ("pps", "dd_pc"):
        ws = pwm_amd.build(d, 20, algo, False)
        _check(ws, DIG["zipf256m"]["builds"]["wm_" + algo], algo)
        ws.free()
    del d


@pytest.mark.gpu
def test_uniform64m_fullsize():
    t = _text("uniform64m")
    d = torch.from_numpy(t).to("cuda:0")
    for shape in ("wm", "wt"):
        for algo in ("ppc", "pps"):
            ws = pwm_amd.build(d, 8, algo, shape == "wt")
            _check(ws, DIG["uniform64m"]["builds"]["%s_%s" % (shape, algo)], (shape, algo))
            ws.free()
    del d
