"""The allocation-free build (pwmhip_build_noalloc) through the C-ABI against
pwmhip_build on the same device texts: every level word and zero count, two
calls in a row (the second with a smaller text, so leftovers of the first
call in the shared static scratch would show), matrices and trees, plus its
size limits."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")
torch = pytest.importorskip("torch")


def _out_words(n, levels):
    return levels * (((n + 63) // 64 + 2 + 31) // 32 * 32)


def _noalloc(ctx, d_text, n, w, levels, algo, flags, d_out, out_words):
    ws = pwm_amd._WS()
    st = pwm_amd.lib().pwmhip_build_noalloc(ctx.h, ctypes.c_void_p(d_text.data_ptr()), n, w, levels,
                                            pwm_amd.ALGOS[algo], flags, ctypes.c_void_p(d_out.data_ptr()),
                                            out_words, ctypes.byref(ws))
    return st, ws


def _levels_of(ws, d_out):
    host = d_out.cpu().numpy().view(np.uint64)
    return [host[ws.level_offset[l]:ws.level_offset[l] + (ws.level_bits[l] + 63) // 64] for l in range(ws.levels)]


@pytest.mark.gpu
def test_noalloc_equals_build_two_calls():
    ctx = pwm_amd.Context(0)
    try:
        rng = np.random.default_rng(0xA110C)
        cases = [
            # (text, levels, algo, tree): the second of each pair is smaller
            (pwm_amd.generate("dna", 0xD1, 5_000_011), 3, "ppc", False),
            (pwm_amd.generate("dna", 0xD2, 777_777), 3, "ppc", False),
            (rng.integers(0, 16, 3_000_001).astype(np.uint8), 4, "pps", True),
            (rng.integers(0, 4, 65_537).astype(np.uint16), 2, "dd_pc", True),
            (rng.integers(0, 8, 1_234_567).astype(np.uint32), 3, "naive", False),
            (rng.integers(0, 8, 4_097).astype(np.uint32), 3, "pc_ss", False),
        ]
        for t, L, algo, tree in cases:
            d = torch.from_numpy(t.view({1: np.int8, 2: np.int16, 4: np.int32}[t.itemsize])).to("cuda:0")
            flags = pwm_amd.FLAG_TEXT_ON_DEVICE | (pwm_amd.FLAG_TREE if tree else 0) | pwm_amd.FLAG_CHECK_RANGE
            want = pwm_amd.build(d, L, algo, tree, ctx=ctx, check_range=True)
            ow = _out_words(t.size, L)
            d_out = torch.full((ow,), -1, dtype=torch.int64, device="cuda:0")  # stale bits must not survive
            st, ws = _noalloc(ctx, d, t.size, t.itemsize, L, algo, flags, d_out, ow)
            assert st == pwm_amd.OK, (st, t.size, L, algo)
            assert ws.alloc_bytes == 0 and ws.levels == L and ws.d_words == d_out.data_ptr()
            got = _levels_of(ws, d_out)
            for l, (x, y) in enumerate(zip(got, want.bvs())):
                assert np.array_equal(x, y), (t.size, L, algo, tree, l)
            assert [ws.zeros[i] for i in range(ws.n_zeros)] == list(want.zeros()), (t.size, algo)
            o = O.naive(t, L, tree)
            for l, (x, y) in enumerate(zip(got, o.levels)):
                assert np.array_equal(x, y), ("oracle", t.size, L, algo, tree, l)
            want.free()
            del d, d_out
    finally:
        ctx.close()


@pytest.mark.gpu
def test_noalloc_limits():
    ctx = pwm_amd.Context(0)
    try:
        lib = pwm_amd.lib()
        t = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda:0")
        ow = _out_words(1 << 20, 3)
        out = torch.zeros(ow, dtype=torch.int64, device="cuda:0")
        # output capacity one word short
        st, _ = _noalloc(ctx, t, 1 << 20, 1, 3, "ppc", 0, out, ow - 1)
        assert st == pwm_amd.ERR_INVALID
        # shapes outside 2..4 levels / Huffman
        assert _noalloc(ctx, t, 1 << 20, 1, 1, "ppc", 0, out, ow)[0] == pwm_amd.ERR_UNSUPPORTED
        assert _noalloc(ctx, t, 1 << 20, 1, 5, "ppc", 0, out, ow)[0] == pwm_amd.ERR_UNSUPPORTED
        assert _noalloc(ctx, t, 1 << 20, 1, 3, "huff_pc", 0, out, ow)[0] == pwm_amd.ERR_UNSUPPORTED
        # temporaries beyond the static scratch: refused before any work (the
        # text and output are not touched, so a short text buffer is enough)
        n = 1 << 30
        big_ow = _out_words(n, 3)
        big = torch.empty(big_ow, dtype=torch.int64, device="cuda:0")
        st, _ = _noalloc(ctx, t, n, 1, 3, "ppc", 0, big, big_ow)
        assert st == pwm_amd.ERR_BUDGET
        del big
        # empty text: an empty structure
        st, ws = _noalloc(ctx, t, 0, 1, 3, "ppc", 0, out, ow)
        assert st == pwm_amd.OK and ws.levels == 0
        # still usable afterwards
        st, ws = _noalloc(ctx, t, 1 << 20, 1, 3, "ppc", 0, out, ow)
        assert st == pwm_amd.OK and [ws.zeros[i] for i in range(3)] == [1 << 20, 1 << 20, 0]
    finally:
        ctx.close()
