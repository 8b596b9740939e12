"""The C-ABI library loads and exports every function include/kqed.h declares (no GPU needed)."""
import ctypes
import os
import re

import kqed_amd

HDR = os.path.join(kqed_amd.ROOT, "include", "kqed.h")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_exports_every_declared_function():
    lib = ctypes.CDLL(kqed_amd.LIB_PATH)
    names = declared_functions()
    assert "QED_kernel_L0" in names and "initialise" in names and "kqed_batch" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_initialise_reports_missing_tables(tmp_path):
    t = kqed_amd._Temps()
    rc = kqed_amd.lib().kqed_initialise_dir(ctypes.byref(t), str(tmp_path).encode())
    assert rc == 1
    assert b"FFxy_single_cksum.bin" in kqed_amd.lib().kqed_last_error()
    kqed_amd.lib().free_QED_temps(ctypes.byref(t))  # safe on a failed struct
