"""The C ABI library loads without a GPU and exports every function that
include/msplat_hip.h declares, with the same names the ctypes binding uses. No compute
call is made here."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "msplat_hip.h")
LIB = os.path.join(ROOT, "msplat-street_amd", "msplat", "libmsplat_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(msplat_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    assert len(names) >= 30
    for must in ("msplat_hip_render", "msplat_hip_train_step", "msplat_hip_render_backward", "msplat_hip_densify",
                 "msplat_hip_radix_sort", "msplat_hip_knn3_mean_dist", "msplat_hip_image_metrics"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmsplat_hip.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmsplat_hip.so not built (run __graft_entry__.build())")
def test_ctypes_binding_covers_the_header():
    """every declared entry point has a ctypes signature in msplat._abi"""
    from msplat import _abi
    lib = _abi.lib()
    unbound = [n for n in declared() if getattr(lib, n).argtypes is None]
    assert not unbound, unbound
    assert lib.msplat_hip_stage_name(0) == b"proj_sh_fwd"
    # the stage mask's bit order is the library's stage numbering
    assert [lib.msplat_hip_stage_name(i).decode() for i in range(len(_abi.STAGE_NAMES))] == list(_abi.STAGE_NAMES)


REF = "/root/reference"


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "core", "metal", "bindings.h")),
                    reason="the reference's headers are not present")
def test_dropin_adapter_compiles_against_the_reference_interface():
    """csrc/host/bindings_hip.cpp implements core/metal/bindings.h: it compiles against
    the reference's own header and MTensor (syntax and types only, nothing is linked)."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    src = os.path.join(ROOT, "msplat-street_amd", "csrc", "host", "bindings_hip.cpp")
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(REF, "core", "metal"),
                        "-I", os.path.join(REF, "core", "include"), "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
