import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pytest_collection_modifyitems(config, items):
    # GPU tests need a device; on a host without one they are skipped loudly
    # rather than routed anywhere else.
    have = None
    for item in items:
        if "gpu" in item.keywords:
            if have is None:
                try:
                    import pwm_amd
                    have = pwm_amd.device_count() > 0
                except Exception:
                    have = False
            if not have:
                item.add_marker(pytest.mark.skip(reason="no HIP device"))
