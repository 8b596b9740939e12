"""Table subsystem: readers, writers and checksums in the reference's file format (io.c:42-413)."""
import ctypes
import os
import shutil

import numpy as np
import pytest

import kqed_amd
from oracle_ctypes import Oracle


class HostTables(ctypes.Structure):
    _fields_ = [("nstpx", ctypes.c_int), ("nstpy", ctypes.c_int), ("nfmax", ctypes.c_int),
                ("XX", ctypes.POINTER(ctypes.c_double)), ("YY", ctypes.POINTER(ctypes.c_double)),
                ("nfx", ctypes.POINTER(ctypes.c_int)),
                ("Ffm", ctypes.POINTER(ctypes.c_float)), ("Ffp", ctypes.POINTER(ctypes.c_float)),
                ("TX", (ctypes.c_double * 810) * 23), ("TY", (ctypes.c_double * 100) * 14)]


def _L():
    L = kqed_amd.lib()
    P = ctypes.POINTER(HostTables)
    L.kqed_read_ff.argtypes = [ctypes.c_char_p, P]
    L.kqed_read_taylorx.argtypes = [ctypes.c_char_p, P]
    L.kqed_read_taylory.argtypes = [ctypes.c_char_p, P]
    L.kqed_write_ff.argtypes = [ctypes.c_char_p, P]
    L.kqed_write_taylor.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.kqed_free_host_tables.argtypes = [P]
    return L


def test_taylor_writer_reproduces_shipped_files(small_dir, tmp_path):
    """Reading the shipped Taylor tables and writing them back gives the same bytes (checksums
    included), so reader, writer and the CRC-32C rotate-xor scheme agree with the reference's."""
    L = _L()
    h = HostTables()
    assert L.kqed_read_taylorx(os.path.join(small_dir, "taylorx_cksum.bin").encode(), ctypes.byref(h)) == 0
    assert L.kqed_read_taylory(os.path.join(small_dir, "taylory_cksum.bin").encode(), ctypes.byref(h)) == 0
    for name, arr, r, c in (("taylorx_cksum.bin", h.TX, 23, 810), ("taylory_cksum.bin", h.TY, 14, 100)):
        out = tmp_path / name
        assert L.kqed_write_taylor(str(out).encode(), ctypes.addressof(arr), r, c) == 0
        assert out.read_bytes() == open(os.path.join(small_dir, name), "rb").read()


def test_ff_roundtrip_and_oracle_agree(small_dir, tmp_path, oracle_small):
    L = _L()
    h = HostTables()
    src = os.path.join(small_dir, "FFxy_single_cksum.bin")
    assert L.kqed_read_ff(src.encode(), ctypes.byref(h)) == 0
    assert h.nstpx == 100 and h.nstpy == 218
    XX = np.ctypeslib.as_array(h.XX, (h.nstpx,))
    assert np.array_equal(XX, oracle_small.XX)
    assert np.array_equal(np.ctypeslib.as_array(h.nfx, (h.nstpx,)), oracle_small.nfx)
    out = tmp_path / "FFxy_single_cksum.bin"
    assert L.kqed_write_ff(str(out).encode(), ctypes.byref(h)) == 0
    assert out.read_bytes() == open(src, "rb").read()
    L.kqed_free_host_tables(ctypes.byref(h))


def test_reference_x_grid_layout(oracle_small):
    """x nodes 3 Taylor-y steps apart up to node 14, then 9 apart, ending at 6.54."""
    h = 6.54 / 810
    k = np.array([3 * (i + 1) if i <= 14 else 45 + 9 * (i - 14) for i in range(100)])
    assert np.allclose(oracle_small.XX, k * h, rtol=0, atol=1e-15)
    assert abs(oracle_small.XX[-1] - 6.54) < 1e-12
    assert np.all(np.diff(oracle_small.nfx) >= 0)


@pytest.mark.parametrize("damage", ["payload", "checksum", "truncate", "magic"])
def test_damaged_form_factor_file_is_rejected(small_dir, tmp_path, damage):
    d = tmp_path / "tab"
    d.mkdir()
    for f in ("taylorx_cksum.bin", "taylory_cksum.bin"):
        shutil.copy(os.path.join(small_dir, f), d / f)
    b = bytearray(open(os.path.join(small_dir, "FFxy_single_cksum.bin"), "rb").read())
    want = {"payload": 4, "checksum": 4, "truncate": 5, "magic": 2}[damage]
    if damage == "payload":
        b[len(b) // 2] ^= 0x40
    elif damage == "checksum":
        b[-1] ^= 0x01
    elif damage == "truncate":
        b = b[: len(b) - 1000]
    else:
        b[0] ^= 0xFF
    (d / "FFxy_single_cksum.bin").write_bytes(bytes(b))
    L = _L()
    h = HostTables()
    assert L.kqed_read_ff(str(d / "FFxy_single_cksum.bin").encode(), ctypes.byref(h)) == want
    L.kqed_free_host_tables(ctypes.byref(h))
    with pytest.raises(RuntimeError):
        Oracle(str(d))


def test_missing_taylor_file(small_dir, tmp_path):
    L = _L()
    h = HostTables()
    assert L.kqed_read_taylorx(str(tmp_path / "nope.bin").encode(), ctypes.byref(h)) == 1


class TableInfo(ctypes.Structure):
    _fields_ = [("nstpx", ctypes.c_int), ("nstpy", ctypes.c_int), ("nffa", ctypes.c_int), ("nfmin", ctypes.c_int),
                ("nfmax", ctypes.c_int), ("x_first", ctypes.c_double), ("x_last", ctypes.c_double),
                ("y_first", ctypes.c_double), ("y_last", ctypes.c_double), ("x_increasing", ctypes.c_int),
                ("y_increasing", ctypes.c_int), ("nonfinite", ctypes.c_size_t)]


def table_info(d):
    L = kqed_amd.lib()
    L.kqed_tables_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(TableInfo)]
    info = TableInfo()
    rc = L.kqed_tables_info(d.encode(), ctypes.byref(info))
    return rc, info


def test_public_info_and_directory_round_trip(small_dir, tmp_path):
    """kqed_tables_info reports the shape; kqed_tables_rewrite reproduces every file byte for byte
    (the shipped Taylor files are the reference's own output)."""
    rc, info = table_info(small_dir)
    p = kqed_amd.TABLES["small"]
    assert rc == 0
    assert (info.nstpx, info.nstpy, info.nffa, info.nfmin, info.nfmax) == (p["nstpx"], p["nstpy"], 14, p["nf0"],
                                                                             max(p["nf0"], p["nf1"] - 1))
    assert info.x_increasing == 1 and info.y_increasing == 1 and info.nonfinite == 0
    assert info.x_first > 0 and info.x_last > info.x_first and info.y_last == pytest.approx(p["yh"] * p["nstpy"])
    L = kqed_amd.lib()
    L.kqed_tables_rewrite.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    assert L.kqed_tables_rewrite(small_dir.encode(), str(tmp_path).encode()) == 0
    for f in ("FFxy_single_cksum.bin", "taylorx_cksum.bin", "taylory_cksum.bin"):
        assert (tmp_path / f).read_bytes() == open(os.path.join(small_dir, f), "rb").read(), f


def test_public_info_rejects_a_damaged_table(small_dir, tmp_path):
    for f in ("FFxy_single_cksum.bin", "taylorx_cksum.bin", "taylory_cksum.bin"):
        shutil.copyfile(os.path.join(small_dir, f), tmp_path / f)
    b = bytearray((tmp_path / "taylorx_cksum.bin").read_bytes())
    b[1000] ^= 0x10
    (tmp_path / "taylorx_cksum.bin").write_bytes(bytes(b))
    rc, _ = table_info(str(tmp_path))
    assert rc == 4
