"""GPU parity of the plain construction (pwmhip_build through the C-ABI)
against the reference's golden vectors and the CPU oracle.  Bit-exact: every
level word, zeros."""
import numpy as np
import pytest

import golden_data
import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")

ZEROS_LAST_IS_ZERO = {"ppc", "ppc_ss", "pc_ss", "dd_pc_ss"}
DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _assert_same(ws, ref_levels, key):
    assert ws.levels() == len(ref_levels), key
    got = ws.bvs()
    for l, (x, y) in enumerate(zip(got, ref_levels)):
        assert x.size == y.size, (key, l)
        if not np.array_equal(x, y):
            bad = np.flatnonzero(x != y)
            raise AssertionError("%s level %d: %d words differ, first at %d" % (key, l, bad.size, bad[0]))


@pytest.mark.gpu
def test_golden_all_algorithms():
    cases = golden_data.plain_cases()
    for i, c in enumerate(cases):
        algos = golden_data.PLAIN_ALGOS if i % 7 == 0 else ["ppc", "pps", "dd_pc", "pc_ss"]
        for a in algos:
            ws = pwm_amd.build(c["text"], c["levels"], a, c["tree"])
            _assert_same(ws, c["bvs"], (c["key"], a))
            z = ws.zeros()
            gz = c["zeros"][a]
            if c["tree"] or c["text"].size == 0:
                assert z.size == 0
            elif a == "pps":
                naive = c["zeros"]["naive"]
                assert np.array_equal(z, naive), (c["key"], a)  # true count (the reference races here)
            else:
                assert np.array_equal(z, gz), (c["key"], a)
            ws.free()


def _random_case(rng, w, L, n, kind="uniform"):
    if kind == "uniform":
        return rng.integers(0, 1 << L, n, dtype=np.uint64).astype(DT[w])
    if kind == "skew":
        k = min(1 << L, 7)
        return rng.choice(k, n, p=rng.dirichlet(np.ones(k) * 0.3)).astype(DT[w])
    raise ValueError(kind)


@pytest.mark.gpu
@pytest.mark.parametrize("w,L", [(1, 1), (1, 2), (1, 3), (1, 5), (1, 8), (2, 9), (2, 12), (2, 16),
                                 (4, 8), (4, 13), (4, 20), (8, 3), (8, 17)])
def test_random_vs_oracle(w, L):
    rng = np.random.default_rng(1000 * w + L)
    sizes = [1, 63, 64, 65, 16383, 16384, 16385, 3 * 16384 + 17, 200003]
    for n in sizes:
        for kind in ("uniform", "skew"):
            t = _random_case(rng, w, L, n, kind)
            for tree in (False, True):
                o = O.naive(t, L, tree)
                ws = pwm_amd.build(t, L, "pc", tree, check_range=True)
                _assert_same(ws, o.levels, (w, L, n, kind, tree))
                if not tree:
                    assert np.array_equal(ws.zeros(), o.zeros)
                ws.free()


@pytest.mark.gpu
def test_device_text_and_reuse():
    import torch
    rng = np.random.default_rng(5)
    t = rng.integers(0, 256, 1 << 20).astype(np.uint8)
    dt = torch.from_numpy(t).to("cuda:0")
    o = O.naive(t, 8, False)
    ws = pwm_amd.build(dt, 8, "ppc")
    _assert_same(ws, o.levels, "device text")
    # second build into the same memory
    t2 = rng.integers(0, 256, 1 << 20).astype(np.uint8)
    o2 = O.naive(t2, 8, False)
    ws2 = pwm_amd.build(torch.from_numpy(t2).to("cuda:0"), 8, "ppc", reuse=ws)
    _assert_same(ws2, o2.levels, "reused output")
    assert int(ws2.zeros()[-1]) == 0  # ppc leaves zeros[L-1] at 0
    ws2.free()


@pytest.mark.gpu
def test_large_uniform_roundtrip():
    """A 2^26-symbol text (the 64 MiB config) against the oracle."""
    t = pwm_amd.generate("uniform", 0xC1, 1 << 26, 1, 8)
    for tree in (False, True):
        o = O.naive(t, 8, tree)
        ws = pwm_amd.build(t, 8, "pps", tree)
        _assert_same(ws, o.levels, ("64MiB", tree))
        if not tree:
            assert np.array_equal(ws.zeros(), o.zeros)
        ws.free()


@pytest.mark.gpu
def test_errors_and_empty():
    ws = pwm_amd.build(np.zeros(0, np.uint8), 8, "ppc")
    assert ws.levels() == 0 and ws.text_size() == 0
    with pytest.raises(pwm_amd.PwmError) as e:
        pwm_amd.build(np.array([1, 2, 9], np.uint8), 3, "ppc", check_range=True)
    assert e.value.status == pwm_amd.ERR_SYMBOL_RANGE
    with pytest.raises(pwm_amd.PwmError) as e:
        pwm_amd.build(np.array([1, 2, 3], np.uint8), 0, "ppc")
    assert e.value.status == pwm_amd.ERR_INVALID


@pytest.mark.gpu
def test_registry_names_and_traits():
    names = {a.name() for a in pwm_amd.algorithm_list()}
    for base in ("wm_ppc", "wt_ppc", "wm_pps", "wt_pps", "wm_dd_pc", "wm_huff_pc", "wt_huff_pc"):
        assert base in names
    a = pwm_amd.get_algorithm("wm_ppc")
    assert a.is_parallel() and not a.is_tree() and not a.is_huffman_shaped()
    t = np.frombuffer(b"abracadabra_banana", np.uint8)
    ws = a.compute(t, t.size, 8)
    o = O.naive(t, 8, False)
    _assert_same(ws, o.levels, "registry")


@pytest.mark.gpu
@pytest.mark.parametrize("w,L", [(4, 20), (2, 12), (1, 8)])
def test_zipf_sparse_buckets_vs_oracle(w, L):
    """Heavy-tailed symbols: deep levels have sparse buckets, so destination
    windows are covered by many one-chunk runs (the run-by-run window path)
    and passes reorder long stretches of equal symbols."""
    t = pwm_amd.generate("zipf", 0xC3 + L, 1 << 21, w, L)
    for tree in (False, True):
        o = O.naive(t, L, tree)
        ws = pwm_amd.build(t, L, "pps", tree, check_range=True)
        _assert_same(ws, o.levels, ("zipf", w, L, tree))
        if not tree:
            assert np.array_equal(ws.zeros(), o.zeros)
        ws.free()


@pytest.mark.gpu
@pytest.mark.parametrize("w,L", [(1, 5), (1, 8), (2, 9), (2, 16), (4, 13), (4, 20)])
def test_passes_over_units_vs_oracle(w, L):
    """Builds of more than 4 levels with every pass over 16-chunk (byte
    symbols) or 8-chunk (wider symbols) units -- the path large texts take --
    at sizes around the unit boundaries, uniform and heavy-tailed symbols."""
    import os
    rng = np.random.default_rng(7000 + 10 * w + L)
    os.environ["PWMHIP_PASS_UNITS"] = "1"
    try:
        for n in [1, 4097, 32768, 65535, 65536, 65537, 3 * 65536 + 4099, 300001]:
            for kind in ("uniform", "skew"):
                t = _random_case(rng, w, L, n, kind)
                for tree in (False, True):
                    o = O.naive(t, L, tree)
                    ws = pwm_amd.build(t, L, "pps", tree, check_range=True)
                    _assert_same(ws, o.levels, ("units", w, L, n, kind, tree))
                    if not tree:
                        assert np.array_equal(ws.zeros(), o.zeros)
                    ws.free()
        t = pwm_amd.generate("zipf", 0x5EED + L, 1 << 20, w, L)
        o = O.naive(t, L, False)
        ws = pwm_amd.build(t, L, "pps", False, check_range=True)
        _assert_same(ws, o.levels, ("units zipf", w, L))
        ws.free()
    finally:
        del os.environ["PWMHIP_PASS_UNITS"]


@pytest.mark.gpu
@pytest.mark.parametrize("w,L,ranked", [(8, 17, False), (2, 9, True), (8, 17, True)])
def test_passes_over_units_reorder_forms_vs_oracle(w, L, ranked):
    """Passes over units with the reorder that ranks 64 symbols per step
    (8-byte symbols always, others with PWMHIP_SCATTER_RANKED=1): its chunk
    positions come from the unit positions as in the counting sort's."""
    import os
    rng = np.random.default_rng(7100 + 10 * w + L + ranked)
    os.environ["PWMHIP_PASS_UNITS"] = "1"
    if ranked:
        os.environ["PWMHIP_SCATTER_RANKED"] = "1"
    try:
        for n in [4097, 65537, 3 * 65536 + 4099, 300001]:
            for kind in ("uniform", "skew"):
                t = _random_case(rng, w, L, n, kind)
                for tree in (False, True):
                    o = O.naive(t, L, tree)
                    ws = pwm_amd.build(t, L, "pps", tree, check_range=True)
                    _assert_same(ws, o.levels, ("units", w, L, ranked, n, kind, tree))
                    if not tree:
                        assert np.array_equal(ws.zeros(), o.zeros)
                    ws.free()
    finally:
        del os.environ["PWMHIP_PASS_UNITS"]
        os.environ.pop("PWMHIP_SCATTER_RANKED", None)
