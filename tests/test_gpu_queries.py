"""GPU rank / select directories and batched access / rank / select queries,
against brute force (rank/select on bit vectors, symbol queries over the text)
and the reference's own query tests' semantics
(test/query_support_tests.cpp:20-85, test/bin_rank_select_tests.cpp:39-62)."""
import numpy as np
import pytest

import bitutil
import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")


def _bv_brute(bits, ops, args):
    pre = np.concatenate([[0], np.cumsum(bits)]).astype(np.int64)
    ones = np.flatnonzero(bits)
    zeros = np.flatnonzero(bits == 0)
    out = []
    for o, x in zip(ops, args):
        x = int(x)
        if o == 0:
            out.append(pre[min(x, bits.size)])
        elif o == 1:
            y = min(x, bits.size)
            out.append(y - pre[y])
        elif o == 2:
            out.append(0 if x == 0 else (ones[x - 1] if x <= ones.size else 2**64 - 1))
        else:
            out.append(0 if x == 0 else (zeros[x - 1] if x <= zeros.size else 2**64 - 1))
    return np.array(out, dtype=np.uint64)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 100, 511, 512, 513, 65535, 65536, 65537, 3 * 65536 + 777, 1 << 21])
def test_bitvector_rank_select(n):
    rng = np.random.default_rng(44227 + n)
    t = rng.integers(0, 2, n).astype(np.uint8)  # 1 level: level 0 == the text bits
    ws = pwm_amd.build(t, 1, "ppc")
    qs = pwm_amd.QuerySupport(ws)
    bits = bitutil.unpack(ws.bvs()[0], n)
    assert np.array_equal(bits, t)
    q = 4000
    ops = rng.integers(0, 4, q).astype(np.uint32)
    args = np.where(ops < 2, rng.integers(0, n + 1, q), rng.integers(0, n + 2, q)).astype(np.uint64)
    args[:8] = [0, 1, n, n + 1, 0, 1, 2, 3]
    ops[:8] = [2, 2, 0, 0, 3, 3, 1, 1]
    res = qs.bv_query(np.zeros(q, np.uint32), ops, args)
    assert np.array_equal(res, _bv_brute(t, ops, args))


@pytest.mark.gpu
def test_sparse_and_dense_select():
    rng = np.random.default_rng(3)
    for p in (0.0005, 0.9995):
        n = 1 << 20
        t = (rng.random(n) < p).astype(np.uint8)
        ws = pwm_amd.build(t, 1, "pc")
        qs = pwm_amd.QuerySupport(ws)
        ones = int(t.sum())
        ops = np.array([2] * 50 + [3] * 50, np.uint32)
        args = np.concatenate([rng.integers(1, ones + 1, 50), rng.integers(1, n - ones + 1, 50)]).astype(np.uint64)
        assert np.array_equal(qs.bv_query(np.zeros(100, np.uint32), ops, args), _bv_brute(t, ops, args))


@pytest.mark.gpu
@pytest.mark.parametrize("tree", [False, True])
@pytest.mark.parametrize("w,L,n", [(1, 8, 5000), (1, 3, 200003), (2, 11, 70000), (4, 20, 300000)])
def test_symbol_queries(tree, w, L, n):
    rng = np.random.default_rng(L * 100 + n)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[w]
    k = min(1 << L, 300)
    t = rng.choice(k, n, p=rng.dirichlet(np.ones(k) * 0.5)).astype(dt)
    for algo in ("ppc", "pps"):  # ppc leaves zeros[L-1] = 0: queries must not depend on it
        ws = pwm_amd.build(t, L, algo, tree)
        qs = pwm_amd.QuerySupport(ws)
        q = 20000
        ops = rng.integers(0, 3, q).astype(np.uint32)
        syms = t[rng.integers(0, n, q)].astype(np.uint64)
        a = np.where(ops == 0, rng.integers(0, n + 1, q), syms).astype(np.uint64)
        b = np.where(ops == 1, rng.integers(0, n + 1, q), rng.integers(0, 50, q)).astype(np.uint64)
        got = qs.query(ops, a, b)
        exp = O.symbol_queries(t, ops, a, b)
        bad = np.flatnonzero(got != exp)
        assert bad.size == 0, (algo, bad[:5], got[bad[:5]], exp[bad[:5]], ops[bad[:5]], a[bad[:5]], b[bad[:5]])
        qs.free()
        ws.free()


@pytest.mark.gpu
def test_reference_query_test_all_symbols():
    """query_support_tests.cpp:20-85: access and rank of all 256 symbols at
    every position of a short text, matrix and tree."""
    t = np.frombuffer("ประเทศไทย中华Việt Nam abracadabra".encode("utf-8"), np.uint8).copy()
    for tree in (False, True):
        ws = pwm_amd.build(t, 8, "ppc", tree)
        qs = pwm_amd.QuerySupport(ws)
        n = t.size
        acc = qs.query(np.zeros(n, np.uint32), np.arange(n))
        assert np.array_equal(acc, t.astype(np.uint64))
        syms = np.repeat(np.arange(256), n + 1)
        idx = np.tile(np.arange(n + 1), 256)
        got = qs.query(np.ones(syms.size, np.uint32), syms, idx)
        pre = np.zeros((256, n + 1), np.int64)
        for c in range(256):
            pre[c, 1:] = np.cumsum(t == c)
        assert np.array_equal(got, pre[syms, idx].astype(np.uint64))


def _huff_queries(rng, t, q):
    n = t.size
    syms = np.unique(t)
    op = rng.integers(0, 3, q).astype(np.uint32)
    pick = syms[rng.integers(0, syms.size, q)].astype(np.uint64)
    absent = rng.integers(0, 8, q) == 0  # some symbols that do not occur
    pick = np.where(absent, rng.integers(0, 1 << (8 * t.itemsize), q, dtype=np.uint64) % (int(t.max()) + 40), pick)
    a = np.where(op == 0, rng.integers(0, n + 3, q).astype(np.uint64), pick).astype(np.uint64)
    b = np.where(op == 1, rng.integers(0, n + 2, q), rng.integers(0, max(2, n // max(1, syms.size) * 2), q))
    return op, a, b.astype(np.uint64)


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["huff_naive", "huff_pc_ss", "huff_pc"])
def test_huffman_symbol_queries_brute_force(algo):
    """access / rank / select on Huffman-shaped matrices and trees (the
    reference has no query support for them) against brute force over the
    text; huff_pc structures whose node borders differ from huff_naive's
    report UNSUPPORTED instead."""
    rng = np.random.default_rng(99)
    cases = [(np.uint8, 4, 5000), (np.uint8, 30, 20000), (np.uint8, 256, 40000), (np.uint16, 900, 30000)]
    answered = 0
    for dt, sig, n in cases:
        for conc in (0.3, 3.0):
            t = rng.choice(sig, n, p=rng.dirichlet(np.ones(sig) * conc)).astype(dt)
            for tree in (False, True):
                try:
                    ws = pwm_amd.build(t, 0, algo, tree)
                except pwm_amd.PwmError as e:
                    assert algo == "huff_pc" and e.status == pwm_amd.ERR_UNSUPPORTED
                    continue
                try:
                    qs = pwm_amd.QuerySupport(ws)
                except pwm_amd.PwmError as e:
                    assert algo == "huff_pc" and e.status == pwm_amd.ERR_UNSUPPORTED, (algo, sig, tree)
                    ws.free()
                    continue
                op, a, b = _huff_queries(rng, t, 20000)
                got = qs.query(op, a, b)
                want = O.symbol_queries(t, op, a, b)
                assert np.array_equal(got, want), (algo, sig, n, conc, tree, np.flatnonzero(got != want)[:5])
                answered += 1
                qs.free()
                ws.free()
    assert answered >= (8 if algo != "huff_pc" else 1)


@pytest.mark.gpu
def test_reference_select_test_at_its_size():
    """test/bin_rank_select_tests.cpp:39-62 at its own size: 2^27 bits drawn
    from std::mt19937(44227); for every position i below 2^27/100 the
    reference asserts select1(ones so far) == i at a one and
    select0(i - ones + 1) == i at a zero.  The same loop runs here, then
    random rank / select over the whole vector against brute force."""
    import torch
    nbits = 1 << 27
    words = O.mt19937_words(44227, nbits // 64)
    bits = np.unpackbits(words.byteswap().view(np.uint8))  # MSB-first bit order
    ws = pwm_amd.build(torch.from_numpy(bits).to("cuda:0"), 1, "ppc")
    assert np.array_equal(ws.bvs()[0], words)
    qs = pwm_amd.QuerySupport(ws)
    m = nbits // 100
    pre = bits[:m].astype(np.int64)
    ones = np.flatnonzero(pre)
    zeros = np.flatnonzero(pre == 0)
    ops = np.concatenate([np.full(ones.size, 2), np.full(zeros.size, 3)]).astype(np.uint32)
    args = np.concatenate([np.arange(1, ones.size + 1), np.arange(1, zeros.size + 1)]).astype(np.uint64)
    got = qs.bv_query(np.zeros(ops.size, np.uint32), ops, args)
    assert np.array_equal(got, np.concatenate([ones, zeros]).astype(np.uint64))
    # the whole vector: random rank0/1 and select0/1
    cum = np.concatenate([[0], np.cumsum(bits, dtype=np.int64)])
    all_ones = np.flatnonzero(bits)
    all_zeros = np.flatnonzero(bits == 0)
    rng = np.random.default_rng(44227)
    q = 1 << 20
    i = rng.integers(0, nbits + 1, q)
    k1 = rng.integers(1, all_ones.size + 1, q)
    k0 = rng.integers(1, all_zeros.size + 1, q)
    ops = np.repeat(np.arange(4, dtype=np.uint32), q)
    args = np.concatenate([i, i, k1, k0]).astype(np.uint64)
    want = np.concatenate([cum[i], i - cum[i], all_ones[k1 - 1], all_zeros[k0 - 1]]).astype(np.uint64)
    got = qs.bv_query(np.zeros(ops.size, np.uint32), ops, args)
    assert np.array_equal(got, want)
    qs.free()
    ws.free()


@pytest.mark.gpu
def test_rank_select_level_past_2_32():
    """A level of 2^32 + 4097 bits (a one-level build of a periodic 0/1 text
    made on the device): rank0/1 and select0/1 at positions on both sides of
    2^32 against the closed forms of a periodic bit string (period P bits
    with C ones: rank1(i) = (i // P) C + ones in the first i % P bits;
    select1(k) = ((k-1) // C) P + position of the ((k-1) % C + 1)-th one)."""
    import torch
    n = (1 << 32) + 4097
    P = 100003
    rng = np.random.default_rng(2 ** 32 + 4097)
    base = rng.integers(0, 2, P).astype(np.uint8)
    d = torch.from_numpy(base).to("cuda:0").repeat((n + P - 1) // P)[:n]
    ws = pwm_amd.build(d, 1, "ppc")
    assert ws.level_bit_size(0) == n
    # level words at sampled indices against the period
    nw = (n + 63) // 64
    lvl = ws.bvs()[0]
    widx = np.unique(np.concatenate([rng.integers(0, nw, 2000), np.arange((1 << 26) - 4, (1 << 26) + 4),
                                     [nw - 1]]))
    pos = widx[:, None] * 64 + np.arange(64)[None, :]
    expb = np.where(pos < n, base[pos % P], 0).astype(np.uint8)
    expw = np.packbits(expb, axis=1).view(">u8").ravel().astype(np.uint64)
    assert np.array_equal(lvl[widx], expw)
    del d
    qs = pwm_amd.QuerySupport(ws)
    C = int(base.sum())
    pre = np.concatenate([[0], np.cumsum(base, dtype=np.int64)])
    one_pos = np.flatnonzero(base)
    zero_pos = np.flatnonzero(base == 0)
    total1 = (n // P) * C + int(pre[n % P])
    total0 = n - total1
    i = np.unique(np.concatenate([rng.integers(0, n + 1, 20000), rng.integers(1 << 32, n + 1, 5000),
                                  np.arange((1 << 32) - 600, (1 << 32) + 600), [0, n - 1, n]])).astype(np.int64)
    r1 = (i // P) * C + pre[i % P]
    k1 = np.unique(np.concatenate([rng.integers(1, total1 + 1, 20000), [1, total1]])).astype(np.int64)
    s1 = ((k1 - 1) // C) * P + one_pos[(k1 - 1) % C]
    Z = P - C
    k0 = np.unique(np.concatenate([rng.integers(1, total0 + 1, 20000), [1, total0]])).astype(np.int64)
    s0 = ((k0 - 1) // Z) * P + zero_pos[(k0 - 1) % Z]
    ops = np.concatenate([np.zeros(i.size), np.ones(i.size), np.full(k1.size, 2), np.full(k0.size, 3)])
    args = np.concatenate([i, i, k1, k0]).astype(np.uint64)
    want = np.concatenate([r1, i - r1, s1, s0]).astype(np.uint64)
    got = qs.bv_query(np.zeros(ops.size, np.uint32), ops.astype(np.uint32), args)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad[:5], ops[bad[:5]], args[bad[:5]], got[bad[:5]], want[bad[:5]])
    qs.free()
    ws.free()
