"""Batched construction (pwmhip_build_batch): every structure of a batch is
bit-identical to the single-text build of its text and to the oracle --
uneven and empty texts, chunk-boundary sizes, every plain shape, trees and
matrices, host and device texts."""
import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _same(a, b, key):
    assert a.levels() == b.levels(), key
    assert a.text_size() == b.text_size(), key
    for l, (x, y) in enumerate(zip(a.bvs(), b.bvs())):
        assert np.array_equal(x, y), (key, l)
    assert np.array_equal(a.zeros(), b.zeros()), key


def _texts(rng, sizes, w, L):
    return [rng.integers(0, 1 << L, n, dtype=np.uint64).astype(DT[w]) for n in sizes]


EDGE_SIZES = [0, 1, 63, 64, 65, 4095, 4096, 4097, 0, 8191, 12289, 3, 100000, 1, 0, 77777]


@pytest.mark.gpu
@pytest.mark.parametrize("w,L", [(1, 1), (1, 2), (1, 3), (1, 5), (1, 8), (2, 4), (2, 7), (4, 8), (8, 6)])
def test_batch_equals_single(w, L):
    rng = np.random.default_rng(31 * w + L)
    texts = _texts(rng, EDGE_SIZES, w, L)
    for tree in (False, True):
        for algo in ("ppc", "pps"):
            got = pwm_amd.build_batch(texts, L, algo, tree, check_range=True)
            assert len(got) == len(texts)
            for i, t in enumerate(texts):
                ref = pwm_amd.build(t, L, algo, tree)
                _same(got[i], ref, (w, L, tree, algo, i, t.size))
                if t.size and t.size < 20000:
                    o = O.naive(t, L, tree)
                    for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
                        assert np.array_equal(x, y), (w, L, tree, i, l)
                ref.free()


@pytest.mark.gpu
def test_batch_many_uneven_dna_texts():
    """2000 DNA texts, log-normal lengths (median 4 KiB) -- the shape of the
    many-texts workload at a size the check finishes quickly."""
    rng = np.random.default_rng(2024)
    sizes = np.clip(rng.lognormal(np.log(4096), 1.5, 2000), 1, 1 << 20).astype(np.int64)
    sizes[::97] = 0
    big = pwm_amd.generate("dna", 0xBA7C, int(sizes.sum()) + 1)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    texts = [big[offs[i]:offs[i + 1]] for i in range(len(sizes))]
    got = pwm_amd.build_batch(texts, 3, "ppc", False)
    # every text against the oracle: levels and zeros (ppc leaves zeros[L-1] at 0)
    for i, t in enumerate(texts):
        assert got[i].levels() == (3 if t.size else 0), i
        if not t.size:
            continue
        o = O.naive(t, 3, False)
        for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
            assert np.array_equal(x, y), (i, l)
        z = o.zeros.copy()
        z[-1] = 0
        assert np.array_equal(got[i].zeros(), z), i
    for i in list(range(0, 2000, 37)) + [int(np.argmax(sizes))]:
        ref = pwm_amd.build(texts[i], 3, "ppc", False)
        _same(got[i], ref, i)
        ref.free()


@pytest.mark.gpu
def test_batch_dna_lognormal_texts_vs_oracle():
    """The many-texts workload's length law (lognormal, median 64 KiB, sigma
    1.5, clipped to [1 KiB, 64 MiB]) over the DNA-like source, 250 texts, each
    structure against the oracle; also with graph replay and autotuning."""
    rng = np.random.default_rng(0xC4)
    sizes = np.clip(rng.lognormal(np.log(65536), 1.5, 250), 1024, 64 << 20).astype(np.int64)
    big = pwm_amd.generate("dna", 0xC4, int(sizes.sum()))
    offs = np.concatenate([[0], np.cumsum(sizes)])
    texts = [big[offs[i]:offs[i + 1]] for i in range(len(sizes))]
    ctx = pwm_amd.Context(0)
    try:
        prev = None
        for rep in range(3):
            got = pwm_amd.build_batch(texts, 3, "ppc", False, ctx=ctx, graph=rep > 0, autotune=rep > 0, reuse=prev)
            for i, t in enumerate(texts):
                o = O.naive(t, 3, False)
                for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
                    assert np.array_equal(x, y), (rep, i, l)
                z = o.zeros.copy()
                z[-1] = 0
                assert np.array_equal(got[i].zeros(), z), (rep, i)
            prev = got
        del got, prev
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [None, "8", "16"])
def test_batch_unit_sizes_vs_oracle(waves, monkeypatch):
    """Batches over units of 8 and of 16 chunks: few long texts (16 by
    default), many short ones (8 by default), and either size forced with
    PWMHIP_SUPER_WAVES; every structure against the oracle, matrix and tree."""
    if waves is not None:
        monkeypatch.setenv("PWMHIP_SUPER_WAVES", waves)
    rng = np.random.default_rng(0xB16 + (0 if waves is None else int(waves)))
    shapes = [
        np.array([70001, 900000, 4096 * 16, 123457, 600000, 4096 * 8 + 5, 700001, 65536 * 3], np.int64),
        np.clip(rng.lognormal(np.log(40000), 1.0, 120), 32768, 400000).astype(np.int64),
    ]
    for si, sizes in enumerate(shapes):
        big = pwm_amd.generate("dna", 0xB17 + si, int(sizes.sum()))
        offs = np.concatenate([[0], np.cumsum(sizes)])
        texts = [big[offs[i]:offs[i + 1]] for i in range(len(sizes))]
        for tree in (False, True):
            got = pwm_amd.build_batch(texts, 3, "pps", tree)
            for i, t in enumerate(texts):
                o = O.naive(t, 3, tree)
                for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
                    assert np.array_equal(x, y), (waves, si, tree, i, l)
                if not tree:
                    assert np.array_equal(got[i].zeros(), o.zeros), (waves, si, i)
            del got


@pytest.mark.gpu
def test_reuse_of_batch_views_is_refused():
    """build(reuse=<a structure of a batch>) must not take over (or free) the
    batch's shared block: the new structure gets its own memory and every
    structure of the batch stays intact."""
    rng = np.random.default_rng(31)
    texts = _texts(rng, [5000, 70000, 4096 * 3 + 5, 200], 1, 3)
    res = pwm_amd.build_batch(texts, 3, "pps", False)
    new_texts = _texts(rng, [9000, 123457], 1, 3)
    a = pwm_amd.build(new_texts[0], 3, "pps", False, reuse=res[0])
    b = pwm_amd.build(new_texts[1], 3, "pps", False, reuse=res[1])
    for ws, t in ((a, new_texts[0]), (b, new_texts[1])):
        o = O.naive(t, 3, False)
        for x, y in zip(ws.bvs(), o.levels):
            assert np.array_equal(x, y)
    for i, t in enumerate(texts):
        res[i]._bvs = None  # download again
        o = O.naive(t, 3, False)
        for l, (x, y) in enumerate(zip(res[i].bvs(), o.levels)):
            assert np.array_equal(x, y), (i, l)
    a.free()
    b.free()
    del res


@pytest.mark.gpu
def test_batch_device_texts_and_lifetime():
    import torch
    rng = np.random.default_rng(9)
    texts = _texts(rng, [5000, 70000, 1, 4096 * 3], 1, 8)
    dev = [torch.from_numpy(t).to("cuda:0") for t in texts]
    got = pwm_amd.build_batch(dev, 8, "dd_pc", True)
    first = got[0]
    del got[1:]  # the others' memory stays valid while one structure lives
    o = O.naive(texts[0], 8, True)
    for x, y in zip(first.bvs(), o.levels):
        assert np.array_equal(x, y)
    first.free()


@pytest.mark.gpu
def test_batch_fallback_shapes_and_errors():
    rng = np.random.default_rng(4)
    # more than 8 levels and Huffman shapes build text by text
    texts = _texts(rng, [3000, 0, 9000], 2, 12)
    got = pwm_amd.build_batch(texts, 12, "pc", False)
    for i, t in enumerate(texts):
        ref = pwm_amd.build(t, 12, "pc", False)
        _same(got[i], ref, ("L12", i))
        if t.size:
            o = O.naive(t, 12, False)
            for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
                assert np.array_equal(x, y), ("L12", i, l)
            assert np.array_equal(got[i].zeros(), o.zeros)
    dna = [pwm_amd.generate("dna", 5 + i, 20000 + 17 * i) for i in range(3)]
    got = pwm_amd.build_batch(dna, 0, "huff_naive", True)
    for i, t in enumerate(dna):
        ref = pwm_amd.build(t, 0, "huff_naive", True)
        _same(got[i], ref, ("huff", i))
        assert np.array_equal(got[i].codes()[0], ref.codes()[0])
        o = O.huff_naive(t, True)
        assert [got[i].level_bit_size(l) for l in range(got[i].levels())] == o.bits
        for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
            assert np.array_equal(x, y), ("huff", i, l)
        assert np.array_equal(got[i].codes()[0], o.codes[0]) and np.array_equal(got[i].codes()[1], o.codes[1])
    with pytest.raises(pwm_amd.PwmError) as e:
        pwm_amd.build_batch([np.array([1, 2], np.uint8), np.array([9], np.uint8)], 3, "ppc", check_range=True)
    assert e.value.status == pwm_amd.ERR_SYMBOL_RANGE
    with pytest.raises(pwm_amd.PwmError):
        pwm_amd.build_batch([np.zeros(3, np.uint8), np.zeros(3, np.uint16)], 3, "ppc")


@pytest.mark.gpu
def test_batch_graph_replay_and_autotune():
    """Captured-graph replays and autotuned launch variants give the same
    structures as the eager launch: the same geometry is replayed several
    times (new text contents each time, eager builds in between), then a
    batch of another geometry is captured anew."""
    rng = np.random.default_rng(404)
    sizes = np.clip(rng.lognormal(np.log(3000), 1.4, 600), 1, 300000).astype(np.int64)
    sizes[::53] = 0
    ctx = pwm_amd.Context(0)
    try:
        prev = None
        for rep in range(4):
            texts = [rng.integers(0, 5, n).astype(np.uint8) for n in sizes]
            tree = rep == 3  # reps 1 and 2 replay the graph of rep 0
            got = pwm_amd.build_batch(texts, 3, "ppc", tree, ctx=ctx, graph=True, autotune=True, reuse=prev)
            eager = pwm_amd.build_batch(texts, 3, "ppc", tree, ctx=ctx)
            for i in list(range(0, len(texts), 29)) + [int(np.argmax(sizes))]:
                _same(got[i], eager[i], (rep, i))
            for i in range(len(texts)):
                if texts[i].size:
                    o = O.naive(texts[i], 3, tree)
                    for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
                        assert np.array_equal(x, y), (rep, i, l)
            del eager
            prev = got
        sizes2 = sizes[: len(sizes) // 2] + 7
        texts = [rng.integers(0, 5, n).astype(np.uint8) for n in sizes2]
        got = pwm_amd.build_batch(texts, 3, "ppc", False, ctx=ctx, graph=True, reuse=prev)
        for i in range(0, len(texts), 17):
            ref = pwm_amd.build(texts[i], 3, "ppc", False, ctx=ctx)
            _same(got[i], ref, ("second geometry", i))
            ref.free()
        del got
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("w,L", [(1, 2), (1, 3), (1, 4), (2, 3), (4, 4)])
def test_batch_long_texts_super_units(w, L):
    """Batches of long texts (mean above 32 Ki symbols: units of 16 chunks,
    the super-chunk build and stitch) equal single builds; empty and tiny
    texts mixed in; graph replay and autotuning on the same path."""
    rng = np.random.default_rng(500 + 10 * w + L)
    sizes = np.clip(rng.lognormal(np.log(60000), 1.2, 48), 1, 900_000).astype(np.int64)
    sizes[[3, 17]] = 0
    sizes[[5, 30]] = [1, 65537]
    sizes[40] = 65536 * 3
    texts = _texts(rng, sizes, w, L)
    assert sizes.sum() / sizes.size >= 32768
    for tree in (False, True):
        got = pwm_amd.build_batch(texts, L, "ppc", tree, check_range=True)
        for i, t in enumerate(texts):
            ref = pwm_amd.build(t, L, "ppc", tree)
            _same(got[i], ref, (w, L, tree, i, t.size))
            ref.free()
        for i in range(0, len(texts), 7):
            if 0 < texts[i].size < 20000:
                o = O.naive(texts[i], L, tree)
                for l, (x, y) in enumerate(zip(got[i].bvs(), o.levels)):
                    assert np.array_equal(x, y), (w, L, tree, i, l)
        del got
    ctx = pwm_amd.Context(0)
    try:
        prev = None
        for rep in range(3):
            got = pwm_amd.build_batch(texts, L, "pps", False, ctx=ctx, graph=True, autotune=True, reuse=prev)
            for i in range(0, len(texts), 5):
                ref = pwm_amd.build(texts[i], L, "pps", False, ctx=ctx)
                _same(got[i], ref, ("graph", rep, i))
                ref.free()
            prev = got
        del got, prev
    finally:
        ctx.close()
