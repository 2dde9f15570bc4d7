"""GPU bit-segment copy (pwmhip_copy_bits, the copy_bits<uint64_t> drop-in,
construction/merge.hpp:19-190) against the reference's own known answers
(test/bit_copy.cpp:45-127, re-packed from MSB-first bytes into MSB-first
64-bit words: byte k of the stream is bits 8k..8k+7) and against the
oracle's copy_bits on random ranges over random destinations."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from test_oracle import KATS

pwm_amd = pytest.importorskip("pwm_amd")
torch = pytest.importorskip("torch")


def _bytes_to_words(b, nwords):
    buf = np.zeros(nwords * 8, np.uint8)
    buf[: len(b)] = b
    return buf.view(">u8").astype(np.uint64)


def _words_to_bytes(w, nbytes):
    return w.astype(">u8").view(np.uint8)[:nbytes]


def _dev(words):
    return torch.from_numpy(words.view(np.int64).copy()).to("cuda:0")


def _copy(ctx, d_dst, d_src, doff, soff, ln):
    st = pwm_amd.lib().pwmhip_copy_bits(ctx.h, ctypes.c_void_p(d_dst.data_ptr()), ctypes.c_void_p(d_src.data_ptr()),
                                        doff, soff, ln)
    assert st == pwm_amd.OK, st


@pytest.mark.gpu
def test_copy_bits_reference_known_answers():
    ctx = pwm_amd.default_context()
    for src, steps in KATS:
        d_src = _dev(_bytes_to_words(src, 2))
        d_dst = _dev(np.zeros(2, np.uint64))
        for (doff, soff, ln), exp in steps:
            _copy(ctx, d_dst, d_src, doff, soff, ln)
            got = _words_to_bytes(d_dst.cpu().numpy().view(np.uint64), 7)
            assert list(got) == exp, ((doff, soff, ln), list(got), exp)


@pytest.mark.gpu
def test_copy_bits_random_vs_oracle():
    ctx = pwm_amd.default_context()
    rng = np.random.default_rng(0xB17C)
    lib = O.oracle()
    nsrc, ndst = 4096, 4096
    src = rng.integers(0, 2 ** 64, nsrc, dtype=np.uint64)
    d_src = _dev(src)
    cases = [(0, 0, 64), (0, 0, 0), (1, 0, 63), (63, 1, 2), (64, 64, 4096), (5, 70, 100000), (130, 3, 200)]
    for _ in range(150):
        ln = int(rng.choice([rng.integers(0, 200), rng.integers(0, 5000), rng.integers(0, 150000)]))
        cases.append((int(rng.integers(0, ndst * 64 - ln)), int(rng.integers(0, nsrc * 64 - ln)), ln))
    for doff, soff, ln in cases:
        dst = rng.integers(0, 2 ** 64, ndst, dtype=np.uint64)  # partially covered words keep their bits
        want = dst.copy()
        lib.oracle_copy_bits64(want.ctypes.data, src.ctypes.data, doff, soff, ln)
        d_dst = _dev(dst)
        _copy(ctx, d_dst, d_src, doff, soff, ln)
        got = d_dst.cpu().numpy().view(np.uint64)
        assert np.array_equal(got, want), (doff, soff, ln)
