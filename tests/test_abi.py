"""CPU checks of the C ABI library: it loads and exports every declared entry
point; status strings and the no-device path work without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hpl_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_core_entry_points():
    names = declared_functions()
    for core in ("hpl_create", "hpl_destroy", "hpl_gefa", "hpl_gesl", "hpl_last_error"):
        assert core in names


def test_library_exports_every_declared_symbol():
    from hpcc_fpga_amd.linpack import LIB_PATH, load_library

    load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hpl_[a-z0-9_]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_status_strings_and_version():
    from hpcc_fpga_amd.linpack import load_library

    L = load_library()
    assert L.hpl_status_string(0) == b"success"
    assert L.hpl_status_string(4) == b"zero or non-finite pivot"
    assert b"gfx950" in L.hpl_version()


def test_no_device_fails_loudly():
    from hpcc_fpga_amd import Context, HplError
    from hpcc_fpga_amd.linpack import load_library

    L = load_library()
    n = C.c_int()
    # count devices without touching one
    try:
        import torch

        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(HplError) as e:
        Context(0)
    assert e.value.status == 9
    assert L.hpl_last_error(None) == b"null context"
    del n
