"""The drop-in C++ backend (csrc/host/bindings_hip.cpp: core/metal/bindings.h on the HIP C
ABI), linked into a driver that uses it as the reference's callers do (tests/
adapter_check.cpp): MTensor views of capacity buffers, train steps between syncs, host
writes after a sync, a device-side buffer copy, densify, opacity reset, the image
pipeline, and prior-guided steps, render, aux outputs and finalize_depth (a PriorStep with
every feature), each compared bit for bit with the C ABI called directly. The driver is built
with the reference's headers by oracle/Makefile (target adapter)."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "oracle", "_ref", "adapter_check")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="oracle/_ref/adapter_check not built (needs the reference's "
                                                       "headers: make -C oracle adapter)")
def test_dropin_adapter_runs_like_the_c_abi():
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    assert r.stdout.count(" ok") == 8
