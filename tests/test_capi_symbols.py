"""The C-ABI library loads without a GPU and exports every entry point include/dpm.h
declares; the ctypes layer binds each of them."""
import ctypes
import os
import re

from _helpers import ROOT, pkg

HEADER = os.path.join(ROOT, "include", "dpm.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dpm_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "dpm_create" in names and "dpm_nve_run" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    P = pkg()
    lib = ctypes.CDLL(P.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_layer_binds_every_symbol():
    P = pkg()
    be = P.load()
    bound = {"dpm_" + k for k in be.fn}
    assert set(declared()) <= bound | set(), sorted(set(declared()) - bound)


def test_no_device_means_loud_failure():
    P = pkg()
    if P.device_count() > 0:
        return
    try:
        P.System(P.DISK, 16)
    except P.DpmError as e:
        assert "no HIP device" in str(e)
    else:
        raise AssertionError("System() must fail without a HIP device")
