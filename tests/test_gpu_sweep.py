"""Seeded randomized parity sweep through the C-ABI: plain builds of every
symbol width with 1 to 24 levels and odd sizes, host texts and unaligned
device views, matrices and trees, builds above four levels over super-chunk
units and over chunks, batches of mixed texts, and Huffman shapes -- each
against the oracle, word for word."""
import os

import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
SIGNED = {2: np.int16, 4: np.int32, 8: np.int64}


def _levels_equal(ws, levels, key):
    got = ws.bvs()
    assert len(got) == len(levels), key
    for l, (x, y) in enumerate(zip(got, levels)):
        if not np.array_equal(x, y):
            bad = np.flatnonzero(x != y)
            raise AssertionError("%s level %d: %d words differ, first at %d" % (key, l, bad.size, bad[0]))


def _symbols(rng, w, L, n):
    if rng.random() < 0.5:
        return rng.integers(0, 1 << L, n, dtype=np.uint64).astype(DT[w])
    k = int(min(1 << L, rng.integers(2, 40)))
    vals = rng.choice(1 << L, k, replace=False) if L < 62 else rng.integers(0, 1 << L, k, dtype=np.uint64)
    p = rng.dirichlet(np.ones(k) * 0.4)
    return np.asarray(vals, dtype=np.uint64)[rng.choice(k, n, p=p)].astype(DT[w])


def _device_view(t, rng):
    """A device copy of t, at an element offset of 0 to 3 from the allocation."""
    import torch
    off = int(rng.integers(0, 4))
    host = np.concatenate([np.zeros(off, t.dtype), t])
    dev = torch.from_numpy(host if t.itemsize == 1 else host.view(SIGNED[t.itemsize])).to("cuda:0")
    return dev[off:], dev


@pytest.mark.gpu
def test_random_plain_sweep():
    rng = np.random.default_rng(20261204)
    for case in range(96):
        w = int(rng.choice([1, 2, 4, 8]))
        L = int(rng.integers(1, min(24, 8 * w) + 1))
        n = int(rng.integers(1, 5000)) if rng.random() < 0.4 else int(rng.integers(5000, 400000))
        t = _symbols(rng, w, L, n)
        tree = bool(rng.random() < 0.5)
        algo = str(rng.choice(["ppc", "pps", "pc"]))
        units = str(rng.choice(["1", "0", ""])) if L > 4 else ""
        key = (case, w, L, n, tree, algo, units)
        if units:
            os.environ["PWMHIP_PASS_UNITS"] = units
        try:
            if rng.random() < 0.5:
                view, keep = _device_view(t, rng)
                ws = pwm_amd.build(view, L, algo, tree, check_range=True)
                del keep
            else:
                ws = pwm_amd.build(t, L, algo, tree, check_range=True)
        finally:
            os.environ.pop("PWMHIP_PASS_UNITS", None)
        o = O.naive(t, L, tree)
        _levels_equal(ws, o.levels, key)
        if not tree:
            z = ws.zeros()
            assert np.array_equal(z[:-1], o.zeros[:-1]), key
            if algo != "ppc":
                assert int(z[-1]) == int(o.zeros[-1]), key
        ws.free()


@pytest.mark.gpu
def test_random_batch_sweep():
    rng = np.random.default_rng(424242)
    for case in range(6):
        w = int(rng.choice([1, 2, 4]))
        L = int(rng.integers(1, min(12, 8 * w) + 1))
        T = int(rng.integers(5, 40))
        sizes = [int(rng.integers(0, 3000)) if rng.random() < 0.5 else int(rng.integers(3000, 150000))
                 for _ in range(T)]
        texts = [_symbols(rng, w, L, s) for s in sizes]
        tree = bool(rng.random() < 0.5)
        got = pwm_amd.build_batch(texts, L, "pps", tree, check_range=True)
        for i, t in enumerate(texts):
            if t.size == 0:
                assert got[i].text_size() == 0
                continue
            o = O.naive(t, L, tree)
            _levels_equal(got[i], o.levels, ("batch", case, i, w, L, t.size, tree))


@pytest.mark.gpu
def test_random_huffman_sweep():
    rng = np.random.default_rng(77)
    done = 0
    for case in range(16):
        w = int(rng.choice([1, 1, 2, 4]))
        k = int(rng.integers(2, 40))
        vals = rng.choice(min(1 << (8 * w), 4096), k, replace=False).astype(DT[w])
        n = int(rng.integers(1, 300000))
        t = vals[rng.choice(k, n, p=rng.dirichlet(np.ones(k) * 0.5))]
        tree = bool(rng.random() < 0.5)
        for algo, fn in (("huff_naive", O.huff_naive), ("huff_pc", O.huff_pc)):
            try:
                o = fn(t, tree)
            except RuntimeError:
                continue  # undefined in the reference
            ws = pwm_amd.build(t, 0, algo, tree)
            got = ws.bvs()
            assert len(got) == len(o.levels), (case, algo)
            for l, (x, y) in enumerate(zip(got, o.levels)):
                assert np.array_equal(x, y), (case, algo, l, w, k, n, tree)
            assert np.array_equal(ws.zeros(), o.zeros), (case, algo)
            ws.free()
            done += 1
    assert done >= 10
