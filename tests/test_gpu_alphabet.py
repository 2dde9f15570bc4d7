"""Device alphabet reduction (pwmhip_reduce_alphabet) against the oracle's
first-occurrence renumbering, then a build over the reduced text."""
import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")


@pytest.mark.gpu
@pytest.mark.parametrize("dt,hi,n", [(np.uint8, 256, 1 << 20), (np.uint16, 50000, 300001),
                                     (np.uint32, 2**32, 200003), (np.uint64, 2**64 - 1, 100000)])
def test_reduce_alphabet_matches_oracle(dt, hi, n):
    rng = np.random.default_rng(n)
    t = rng.integers(0, hi, n, dtype=np.uint64).astype(dt)
    if dt == np.uint64:
        t[5] = np.uint64(2**64 - 1)  # the all-ones symbol
    want, wl = O.reduce_alphabet(t)
    got, gl = pwm_amd.reduce_alphabet(t)
    assert gl == wl
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_reduce_alphabet_then_build():
    text = np.frombuffer(("ประเทศไทย中华Việt Nam".encode("utf-8") + bytes([255, 0, 255])) * 300, np.uint8).copy()
    red, levels = pwm_amd.reduce_alphabet(text)
    ref, rl = O.reduce_alphabet(text)
    assert levels == rl and np.array_equal(red, ref)
    ws = pwm_amd.build(red, levels, "ppc", check_range=True)
    o = O.naive(red, levels, False)
    for x, y in zip(ws.bvs(), o.levels):
        assert np.array_equal(x, y)
    assert pwm_amd.reduce_alphabet(np.zeros(0, np.uint8))[1] == 0
    assert pwm_amd.reduce_alphabet(np.full(10, 9, np.uint16))[1] == 0
