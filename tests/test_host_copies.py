"""Host side of the staged host<->device transfers
(dxmclib_amd/csrc/hostcopy.hpp): the persistent helper threads that copy a
staging chunk in parts. Built as tests/cpp/bin/copy_helpers; it checks odd
lengths around the 4 MiB part threshold and the 8 MiB chunk, unaligned
addresses, repeated copies and copies without helper threads."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "copy_helpers")


def test_staged_copy_helpers():
    if not os.path.exists(EXE):
        pytest.skip("tests/cpp/bin/copy_helpers not built (python -m dxmclib_amd.build)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
