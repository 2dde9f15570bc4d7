"""The local build that writes bucket-0 runs in place (byte texts of 2 and 3
levels, the default for them) against the oracle (wx_naive restated), on the
shapes that stress it: bucket 0 empty, whole, or so rare that most units'
runs are shorter than a word; sizes at unit boundaries; and texts of many
units, so that every workgroup's running prefix crosses several rounds of
the unit queue.  Also the same texts with every word left to the stitch
(PWMHIP_STITCH_ALL=1), and a reused output."""
import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")

UNIT = 1 << 16  # symbols per unit of the local build


def _check(ws, t, L, tree, key):
    o = O.naive(t, L, tree)
    assert ws.levels() == L, key
    for l, (x, y) in enumerate(zip(ws.bvs(), o.levels)):
        if not np.array_equal(x, y):
            bad = np.flatnonzero(x != y)
            raise AssertionError("%s level %d: %d words differ, first at %d" % (key, l, bad.size, bad[0]))
    if not tree and t.size:
        assert np.array_equal(ws.zeros(), o.zeros), key  # pps: every level's true count


def _text(rng, kind, L, n):
    top = 1 << (L - 1)  # symbols >= top have the first level's bit set
    if kind == "zeros":
        return np.zeros(n, np.uint8)
    if kind == "no_bucket0":  # every symbol has its first bit set
        return rng.integers(top, 1 << L, n).astype(np.uint8)
    if kind.startswith("rare"):  # bucket 0 of every sub-level at rate p
        p = {"rare3": 1e-3, "rare5": 1e-5}[kind]
        t = rng.integers(top, 1 << L, n).astype(np.uint8)
        hit = rng.random(n) < p
        t[hit] = rng.integers(0, top, int(hit.sum())).astype(np.uint8)
        return t
    if kind == "mostly0":  # bucket 0 nearly everything, a few others
        t = np.zeros(n, np.uint8)
        hit = rng.random(n) < 1e-3
        t[hit] = rng.integers(1, 1 << L, int(hit.sum())).astype(np.uint8)
        return t
    if kind == "uniform":
        return rng.integers(0, 1 << L, n).astype(np.uint8)
    raise ValueError(kind)


@pytest.mark.gpu
@pytest.mark.parametrize("L", [2, 3])
def test_inplace_shapes_vs_oracle(L):
    rng = np.random.default_rng(0xD1 + L)
    sizes = [1, 63, 64, 65, UNIT - 1, UNIT, UNIT + 1, 5 * UNIT + 4097, 37 * UNIT + 65]
    for kind in ("zeros", "no_bucket0", "rare3", "rare5", "mostly0", "uniform"):
        for n in sizes:
            t = _text(rng, kind, L, n)
            for tree in (False, True):
                ws = pwm_amd.build(t, L, "pps", tree)
                _check(ws, t, L, tree, (kind, L, n, tree))
                ws.free()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["rare3", "uniform", "dna"])
def test_inplace_many_units_vs_oracle(kind):
    # 1100 units: more than four rounds of the queue on a 256-CU device
    rng = np.random.default_rng(0xD2)
    n = 1100 * UNIT + 12345
    t = pwm_amd.generate("dna", 0xD3, n) if kind == "dna" else _text(rng, kind, 3, n)
    ctx = pwm_amd.Context(0)
    ws = None
    for tree in (False, True):
        ws = pwm_amd.build(t, 3, "pps", tree, ctx=ctx, reuse=ws)
        _check(ws, t, 3, tree, (kind, tree))
    ws.free()
    ctx.close()


@pytest.mark.gpu
def test_inplace_equals_stitch_all(monkeypatch):
    rng = np.random.default_rng(0xD4)
    for L in (2, 3):
        for kind in ("rare3", "uniform", "mostly0"):
            t = _text(rng, kind, L, 9 * UNIT + 777)
            a = pwm_amd.build(t, L, "ppc")
            monkeypatch.setenv("PWMHIP_STITCH_ALL", "1")
            b = pwm_amd.build(t, L, "ppc")
            monkeypatch.delenv("PWMHIP_STITCH_ALL")
            for l, (x, y) in enumerate(zip(a.bvs(), b.bvs())):
                assert np.array_equal(x, y), (L, kind, l)
            assert np.array_equal(a.zeros(), b.zeros()), (L, kind)
            a.free()
            b.free()


@pytest.mark.gpu
def test_inplace_wait_gives_up(monkeypatch):
    """The bounded wait for earlier units' counts: with no unit publishing
    (PWMHIP_TEST_UNPUBLISHED=1, waits give up after 16 reloads) the build
    ends with the timeout status instead of hanging, and the next build on
    the same context is exact again."""
    t = pwm_amd.generate("dna", 0xD5, 40 * UNIT + 77)
    ctx = pwm_amd.Context(0)
    try:
        monkeypatch.setenv("PWMHIP_TEST_UNPUBLISHED", "1")
        with pytest.raises(pwm_amd.PwmError) as e:
            pwm_amd.build(t, 3, "ppc", False, ctx=ctx)
        assert e.value.status == pwm_amd.ERR_TIMEOUT
        monkeypatch.delenv("PWMHIP_TEST_UNPUBLISHED")
        ws = pwm_amd.build(t, 3, "pps", False, ctx=ctx)
        _check(ws, t, 3, False, "after the timeout")
        ws.free()
    finally:
        ctx.close()
