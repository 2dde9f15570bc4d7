"""Streamed host texts (PWMHIP_FLAG_STREAM, the in-HBM analogue of
wx_ppc_ie, external_memory/wx_ppc_ie.hpp:21-65), with slices small enough
that the text spans many of them (ragged last slice included): every level
and zero count equals the oracle's (wx_naive restated) and the structure
built from a device-resident copy of the text."""
import os

import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")


@pytest.fixture
def small_slices(monkeypatch):
    monkeypatch.setenv("PWMHIP_STREAM_SLICE_SYMS", str(3 * 65536))  # three super-chunks per slice
    yield


def _oracle(ws, t, L, tree, algo, key):
    o = O.naive(t, L, tree)
    assert ws.levels() == len(o.levels), key
    for l, (x, y) in enumerate(zip(ws.bvs(), o.levels)):
        assert np.array_equal(x, y), (key, l)
    if not tree:
        z = o.zeros.copy()
        if algo == "ppc":
            z[-1] = 0  # wx_ppc never sets zeros[levels-1] (ppc.hpp:19-90)
        assert np.array_equal(ws.zeros(), z), key


def _same(a, b, key):
    assert a.levels() == b.levels(), key
    for l, (x, y) in enumerate(zip(a.bvs(), b.bvs())):
        assert np.array_equal(x, y), (key, l)
    assert np.array_equal(a.zeros(), b.zeros()), key


@pytest.mark.gpu
@pytest.mark.parametrize("w,L", [(1, 2), (1, 3), (1, 4), (2, 3), (4, 4)])
def test_stream_equals_resident(small_slices, w, L):
    rng = np.random.default_rng(10 * w + L)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[w]
    for n in (5, 65536 * 3, 65536 * 3 + 1, 1_000_003):
        t = rng.integers(0, 1 << L, n).astype(dt)
        for tree in (False, True):
            for algo in ("ppc", "pps"):
                a = pwm_amd.build(t, L, algo, tree, stream=True)
                b = pwm_amd.build(t, L, algo, tree)
                _same(a, b, (w, L, n, tree, algo))
                _oracle(a, t, L, tree, algo, (w, L, n, tree, algo))
                a.free()
                b.free()


@pytest.mark.gpu
def test_stream_dna_and_unsupported_shapes(small_slices):
    t = pwm_amd.generate("dna", 0x5E, 3_333_333)  # 17 slices, the last one ragged
    for tree in (False, True):
        a = pwm_amd.build(t, 3, "ppc", tree, stream=True)
        b = pwm_amd.build(t, 3, "ppc", tree)
        _same(a, b, ("dna", tree))
        _oracle(a, t, 3, tree, "ppc", ("dna", tree))
        a.free()
        b.free()
    for levels, algo in ((8, "ppc"), (1, "ppc"), (0, "huff_pc")):
        with pytest.raises(pwm_amd.PwmError) as e:
            pwm_amd.build(t, levels, algo, False, stream=True)
        assert e.value.status == pwm_amd.ERR_UNSUPPORTED, (levels, algo)
