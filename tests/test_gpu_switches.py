"""Every alternative kernel path that an environment switch selects, against
the oracle (wx_naive / huff_pc / huff_naive restated): plain matrices and
trees of 1-20 levels and several widths at sizes around the chunk and
super-chunk boundaries, multi-pass builds over chunks and over units, a
batch of uneven texts, and Huffman shapes.  The library reads the switches at
every call, so one process covers them all."""
import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}

SWITCHES = [
    {},
    {"PWMHIP_GENERIC_K1": "1"},
    {"PWMHIP_CHUNK_RUNS": "1"},
    {"PWMHIP_NARROW_GATHER": "1"},
    {"PWMHIP_WIDE_STEPS": "4"},
    {"PWMHIP_WIDE_STEPS": "32"},
    {"PWMHIP_SUPER_WAVES": "8"},
    {"PWMHIP_SCATTER_RANKED": "1", "PWMHIP_PASS_UNITS": "0"},
    {"PWMHIP_SCATTER_RANKED": "1", "PWMHIP_PASS_UNITS": "1"},
    {"PWMHIP_HUFF_CHUNKS": "1"},
    {"PWMHIP_HUFF_RUN_STITCH": "1"},
    {"PWMHIP_STITCH_ALL": "1"},
]


def _same(ws, o, key, zeros_last_zero=False):
    assert ws.levels() == len(o.levels), key
    for l, (x, y) in enumerate(zip(ws.bvs(), o.levels)):
        assert np.array_equal(x, y), (key, l)
    if o.zeros.size:
        z = o.zeros.copy()
        if zeros_last_zero:
            z[-1] = 0
        assert np.array_equal(ws.zeros(), z), key


@pytest.mark.gpu
@pytest.mark.parametrize("env", SWITCHES, ids=lambda e: ",".join("%s=%s" % kv for kv in e.items()) or "default")
def test_switch_paths_vs_oracle(monkeypatch, env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(0x5A17 + SWITCHES.index(env))
    # plain single builds: one pass (1-4 levels), several passes (5-20)
    for w, L in ((1, 1), (1, 2), (1, 3), (1, 4), (1, 8), (2, 11), (4, 20)):
        for n in (1, 4097, 65537, 3 * 65536 + 4099, 700001):
            k = min(1 << L, 9)
            t = (rng.choice(k, n, p=rng.dirichlet(np.ones(k) * 0.5)) if n % 2 else
                 rng.integers(0, 1 << L, n)).astype(DT[w])
            for tree in (False, True):
                ws = pwm_amd.build(t, L, "pps", tree, check_range=True)
                _same(ws, O.naive(t, L, tree), (env, w, L, n, tree))
                ws.free()
    # a DNA-like text of 2 Mi symbols, matrix and tree (super-chunk paths)
    t = pwm_amd.generate("dna", 0x5717, (1 << 21) + 333)
    for tree in (False, True):
        ws = pwm_amd.build(t, 3, "ppc", tree)
        _same(ws, O.naive(t, 3, tree), (env, "dna", tree), zeros_last_zero=True)
        ws.free()
    # a batch of uneven texts (chunk units and super-chunk units)
    for med in (3000, 90000):
        sizes = np.clip(rng.lognormal(np.log(med), 1.0, 40), 1, 1 << 20).astype(np.int64)
        texts = [rng.integers(0, 5, int(s)).astype(np.uint8) for s in sizes]
        res = pwm_amd.build_batch(texts, 3, "pps", graph=False)
        for i, (x, ws) in enumerate(zip(texts, res)):
            _same(ws, O.naive(x, 3, False), (env, "batch", med, i))
        del res
    # Huffman shapes: DNA-like bytes (super-chunk kernels) and a wider alphabet
    for t in (pwm_amd.generate("dna", 0x4FF, 1_000_003),
              rng.choice(40, 200_001, p=rng.dirichlet(np.ones(40) * 0.4)).astype(np.uint8)):
        for tree in (False, True):
            for algo, fn in (("huff_naive", O.huff_naive), ("huff_pc", O.huff_pc)):
                try:
                    o = fn(t, tree)
                except RuntimeError:
                    continue  # the reference's huff_pc writes outside a level here
                ws = pwm_amd.build(t, 0, algo, tree)
                assert [ws.level_bit_size(l) for l in range(ws.levels())] == o.bits
                _same(ws, o, (env, algo, t.size, tree))
                ws.free()
