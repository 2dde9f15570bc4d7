"""Oversize and invalid requests through the C-ABI on a real device: each
returns a status and leaves the context usable.  Every case here fails on
the host before any kernel is launched (a device allocation that cannot be
satisfied, an admission check, a size check), so no kernel ever sees the
placeholder pointers."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")
torch = pytest.importorskip("torch")


def _still_works(ctx):
    t = np.arange(1000, dtype=np.uint8) % 5
    ws = pwm_amd.build(t, 3, "ppc", ctx=ctx)
    o = O.naive(t, 3, False)
    for x, y in zip(ws.bvs(), o.levels):
        assert np.array_equal(x, y)
    ws.free()


@pytest.mark.gpu
def test_oversize_requests_return_status():
    lib = pwm_amd.lib()
    ctx = pwm_amd.Context(0)
    try:
        dummy = torch.zeros(64, dtype=torch.uint8, device="cuda:0")
        # an output of 2^50 symbols x 3 levels cannot be allocated: the output
        # allocation fails before the text is read
        ws = pwm_amd._WS()
        st = lib.pwmhip_build(ctx.h, ctypes.c_void_p(dummy.data_ptr()), 1 << 50, 1, 3, 2,
                              pwm_amd.FLAG_TEXT_ON_DEVICE, ctypes.byref(ws))
        assert st in (pwm_amd.ERR_OOM, pwm_amd.ERR_HIP), st
        assert ws.d_words is None and ws.levels == 0
        _still_works(ctx)
        # a batch whose block cannot be allocated
        ptrs = (ctypes.c_void_p * 2)(dummy.data_ptr(), dummy.data_ptr())
        sizes = (ctypes.c_uint64 * 2)(1 << 49, 1 << 49)
        outs = (pwm_amd._WS * 2)()
        st = lib.pwmhip_build_batch(ctx.h, ptrs, sizes, 2, 1, 3, 2, pwm_amd.FLAG_TEXT_ON_DEVICE, outs)
        assert st in (pwm_amd.ERR_OOM, pwm_amd.ERR_HIP), st
        assert outs[0].d_words is None and outs[1].d_words is None
        _still_works(ctx)
        # rank/select over a level of 2^48 bits or more is refused (its
        # directory indexes superblocks with 32 bits)
        fake = pwm_amd._WS()
        fake.levels = 1
        fake.text_size = 1 << 48
        fake.level_bits[0] = 1 << 48
        fake.d_words = dummy.data_ptr()
        h = ctypes.c_void_p()
        assert lib.pwmhip_rs_build(ctx.h, ctypes.byref(fake), ctypes.byref(h)) == pwm_amd.ERR_UNSUPPORTED
        assert not h.value
        # query staging of 2^40 host queries cannot be allocated
        ws = pwm_amd.build(np.arange(5000, dtype=np.uint8) % 5, 3, "ppc", ctx=ctx)
        qs = pwm_amd.QuerySupport(ws)
        small = np.zeros(16, np.uint64)
        p = small.ctypes.data_as(ctypes.c_void_p)
        st = lib.pwmhip_query(ctx.h, ctypes.byref(ws._ws), qs.h, p, p, p, 1 << 40, p, 0)
        assert st in (pwm_amd.ERR_OOM, pwm_amd.ERR_HIP), st
        st = lib.pwmhip_bv_query(ctx.h, qs.h, p, p, p, 1 << 40, p, 0)
        assert st in (pwm_amd.ERR_OOM, pwm_amd.ERR_HIP), st
        qs.free()
        ws.free()
        _still_works(ctx)
        # a rejected build with REUSE_OUTPUT leaves the earlier result intact
        t = np.arange(3000, dtype=np.uint8) % 7
        ws = pwm_amd.build(t, 3, "ppc", ctx=ctx)
        st = lib.pwmhip_build(ctx.h, None, 10, 1, 0, 2, pwm_amd.FLAG_REUSE_OUTPUT, ctypes.byref(ws._ws))
        assert st == pwm_amd.ERR_INVALID
        ws._bvs = None
        o = O.naive(t, 3, False)
        for x, y in zip(ws.bvs(), o.levels):
            assert np.array_equal(x, y)
        ws.free()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_server_oversize_requests_return_status():
    srv = pwm_amd.Server(hbm_budget_bytes=8 << 30, workers=1)
    try:
        lib = pwm_amd.lib()
        rid = ctypes.c_uint64()
        small = np.zeros(16, np.uint8)
        # admission: the output alone exceeds the budget
        st = lib.pwmhip_server_build(srv.h, small.ctypes.data_as(ctypes.c_void_p), 1 << 40, 1, 3, 2, 0,
                                     ctypes.byref(rid))
        assert st == pwm_amd.ERR_BUDGET, st
        t = np.arange(4096, dtype=np.uint8) % 5
        idx = srv.build(t, 3, "ppc")
        srv.wait(idx)
        q = np.zeros(16, np.uint64)
        st = lib.pwmhip_server_query(srv.h, idx, q.ctypes.data_as(ctypes.c_void_p), q.ctypes.data_as(ctypes.c_void_p),
                                     None, 1 << 40, q.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rid))
        assert st == pwm_amd.ERR_BUDGET, st
        got = srv.levels(idx)
        o = O.naive(t, 3, False)
        for x, y in zip(got, o.levels):
            assert np.array_equal(x, y)
        srv.release(idx)
    finally:
        srv.close()
