"""The C++ host side (include/pwm_hip_wx.hpp): the wx_hip_* adapters carry the
reference's algorithm traits and compute() signature; the compiled driver
(tests/cpp/wx_hip_test.cpp) checks their results against the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "wx_hip_test")


def test_adapter_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "pwm_hip_wx.hpp"\n'
                   'static_assert(pwm_hip::wx_hip_dd_pc<uint8_t, true>::is_tree, "");\n'
                   'int main() { return 0; }\n')
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                           "-fsyntax-only", str(src)])


@pytest.mark.gpu
def test_cpp_adapters_match_oracle():
    assert os.path.exists(DRIVER), "build the driver first (__graft_entry__.build())"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout
