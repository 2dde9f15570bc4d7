"""GPU Huffman-shaped construction (wx_huff_pc / wx_huff_naive drop-ins)
against the reference's golden vectors and the oracle, bit-exact: codes, level
sizes, every level word, zeros (huff_pc's border bookkeeping included)."""
import numpy as np
import pytest

import bitutil
import golden_data
import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")


def _same(ws, levels, bits, key):
    assert ws.levels() == len(levels), key
    assert [ws.level_bit_size(l) for l in range(ws.levels())] == list(bits), key
    for l, (x, y) in enumerate(zip(ws.bvs(), levels)):
        assert np.array_equal(x, y), (key, l)


@pytest.mark.gpu
def test_huffman_golden():
    for c in golden_data.huff_cases():
        t, tree = c["text"], c["tree"]
        for algo in ("huff_pc", "huff_naive", "huff_ps", "huff_pc_ss", "huff_dd_pc"):
            g = c["algos"][algo]
            ws = pwm_amd.build(t, 0, algo, tree)
            _same(ws, g["bvs"], g["bits"], (c["key"], algo))
            ln, wd = ws.codes()
            assert np.array_equal(ln, g["codes"][0]) and np.array_equal(wd, g["codes"][1]), c["key"]
            assert np.array_equal(ws.zeros(), g["zeros"]), (c["key"], algo, ws.zeros(), g["zeros"])
            ws.free()


@pytest.mark.gpu
def test_huffman_dna_large_vs_oracle():
    t = pwm_amd.generate("dna", 0xC2, 3_000_017)
    for tree in (False, True):
        for algo, fn in (("huff_pc", O.huff_pc), ("huff_naive", O.huff_naive)):
            o = fn(t, tree)
            ws = pwm_amd.build(t, 0, algo, tree)
            _same(ws, o.levels, o.bits, (algo, tree))
            assert np.array_equal(ws.zeros(), o.zeros)
            ws.free()


@pytest.mark.gpu
def test_huffman_naive_decodes_and_wide_symbols():
    rng = np.random.default_rng(17)
    for w, sig in ((1, 256), (2, 1000), (4, 3000)):
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[w]
        t = rng.choice(sig, 50000, p=rng.dirichlet(np.ones(sig) * 0.2)).astype(dt)
        for tree in (False, True):
            ws = pwm_amd.build(t, 0, "huff_naive", tree)
            o = O.huff_naive(t, tree)
            _same(ws, o.levels, o.bits, (w, tree))
            d = bitutil.decode_huff(ws.bvs(), [ws.level_bit_size(l) for l in range(ws.levels())],
                                    ws.codes(), tree, t.size)
            assert np.array_equal(d, t.astype(np.int64))
            ws.free()


@pytest.mark.gpu
def test_huffman_pc_undefined_cases_are_rejected():
    """Where the reference's huff_pc writes outside a level (undefined
    behaviour), the drop-in reports it instead of guessing."""
    rng = np.random.default_rng(7)
    rejected = 0
    for _ in range(30):
        sig = int(rng.integers(20, 256))
        t = rng.choice(sig, 3000, p=rng.dirichlet(np.ones(sig))).astype(np.uint8)
        try:
            O.huff_pc(t, False)
            defined = True
        except RuntimeError:
            defined = False
        if defined:
            ws = pwm_amd.build(t, 0, "huff_pc", False)
            o = O.huff_pc(t, False)
            _same(ws, o.levels, o.bits, "defined")
            ws.free()
        else:
            with pytest.raises(pwm_amd.PwmError) as e:
                pwm_amd.build(t, 0, "huff_pc", False)
            assert e.value.status == pwm_amd.ERR_UNSUPPORTED
            rejected += 1
    assert rejected > 0


@pytest.mark.gpu
@pytest.mark.parametrize("values", [(0, 1, 2), (0, 1, 2, 3, 4), (3, 5, 6, 7), (10, 200, 37, 99, 255, 64)])
def test_huffman_small_codes_vs_oracle(values):
    """Codes of 2..4 bits (the super-chunk path): sizes around chunk and
    super-chunk boundaries, symbols below 8 and above."""
    rng = np.random.default_rng(sum(values))
    vals = np.array(values, dtype=np.uint8)
    for n in (2, 4095, 4097, 65535, 65537, 200_003):
        p = rng.dirichlet(np.ones(len(vals)))
        t = vals[rng.choice(len(vals), n, p=p)]
        for tree in (False, True):
            for algo, fn in (("huff_pc", O.huff_pc), ("huff_naive", O.huff_naive)):
                try:
                    o = fn(t, tree)
                except RuntimeError:
                    continue  # undefined in the reference
                ws = pwm_amd.build(t, 0, algo, tree)
                _same(ws, o.levels, o.bits, (values, n, tree, algo))
                assert np.array_equal(ws.zeros(), o.zeros), (values, n, tree, algo)
                ws.free()


@pytest.mark.gpu
def test_huffman_histogram_small_alphabet_with_rare_outliers():
    """Byte texts almost all below 8 take the packed-counter histogram; rare
    larger bytes (below the sample's one-in-a-thousand bar, some past the
    1 MiB sample) are counted byte by byte.  Unaligned views too."""
    rng = np.random.default_rng(0xB17)
    n = 3 * (1 << 20) + 12345
    base = rng.integers(0, 5, n + 7, dtype=np.uint8)
    pos = rng.choice(n + 7, 40, replace=False)
    base[pos] = rng.integers(8, 256, pos.size, dtype=np.uint8)
    base[(1 << 20) + 999] = 7
    for off in (0, 3):
        t = base[off:off + n]
        o = O.huff_naive(t, False)
        ws = pwm_amd.build(t, 0, "huff_naive", False)
        _same(ws, o.levels, o.bits, ("outliers", off))
        assert np.array_equal(ws.zeros(), o.zeros), off
        ws.free()


@pytest.mark.gpu
def test_huffman_histogram_packed_counts_large():
    """Texts past one grid step of the small-alphabet histogram (16 MiB and
    more) are counted 16 bytes at a time: all below 4 (2-bit fields), some
    value 4-7 (3-bit fields) or a byte of 8 or more (byte by byte), the three
    kinds interleaved at 16-byte granularity; per-level sizes and zeros
    follow from the counts, so a miscount shows."""
    rng = np.random.default_rng(0xB18)
    n = (40 << 20) + 4321
    t = rng.integers(0, 4, n, dtype=np.uint8)
    for b in rng.choice(n // 65536, 40, replace=False):  # 3-bit stretches
        t[b * 65536:(b + 1) * 65536] = rng.integers(0, 8, 65536, dtype=np.uint8)
    g = rng.choice(n // 16, 5000, replace=False)  # lone 16-byte groups with a 4-7
    t[g * 16 + rng.integers(0, 16, g.size)] = rng.integers(4, 8, g.size, dtype=np.uint8)
    pos = (1 << 20) + rng.choice(n - (1 << 20), 300, replace=False)  # past the sample
    t[pos] = rng.integers(8, 256, pos.size, dtype=np.uint8)
    for off in (0, 5):
        v = t[off:]
        o = O.huff_naive(v, False)
        ws = pwm_amd.build(v, 0, "huff_naive", False)
        _same(ws, o.levels, o.bits, ("packed", off))
        assert np.array_equal(ws.zeros(), o.zeros), off
        ws.free()


@pytest.mark.gpu
def test_huffman_reused_output():
    """Huffman builds into an earlier structure's memory (large enough: taken
    over; too small: replaced), results unchanged."""
    rng = np.random.default_rng(0x5E)
    ws = None
    for n, k in ((300_000, 5), (20_000, 3), (600_000, 30)):
        vals = rng.choice(200, k, replace=False).astype(np.uint8)
        t = vals[rng.choice(k, n, p=rng.dirichlet(np.ones(k)))]
        o = O.huff_naive(t, False)
        ws = pwm_amd.build(t, 0, "huff_naive", False, reuse=ws)
        _same(ws, o.levels, o.bits, ("reuse", n, k))
        assert np.array_equal(ws.zeros(), o.zeros), (n, k)
    ws.free()
