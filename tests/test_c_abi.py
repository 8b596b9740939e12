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


def test_c_caller_compiles_and_links(tmp_path):
    """A C99 caller written against the reference's KQED.h compiles against include/ and links."""
    import subprocess
    src = tmp_path / "caller.c"
    src.write_text(
        '#include "KQED.h"\n'
        "int main(void) {\n"
        "  struct QED_kernel_temps t; double kerv[6][4][4][4] KQED_ALIGN;\n"
        "  const double x[4] = {0, 0, 0.1, 0.2}, y[4] = {0, 0, 0, 0.3};\n"
        "  if (initialise(&t)) return 1;\n"
        "  QED_kernel_L0(x, y, t, kerv); QED_kernel_L2(x, y, t, kerv); QED_Mkernel_L2(0.5, x, y, t, kerv);\n"
        "  { struct QED_Kernels K; struct Kernels C[3]; const double M[4] = {0, .4, .8, 1.2};\n"
        "    compute_all_kernels(x, y, t, &K); compute_con_kernelsM_L2(M, x, y, t, C); swap_munu_Lyx(&K);\n"
        "    average_con_kernels(C, C); printf(\"%g\\n\", K.L2.xy[0][0][0][0] + C[2].L1[0][1][2][3]); }\n"
        "  printf(\"%g %g\\n\", kerv[0][0][0][0], AlfQED);\n"
        "  free_QED_temps(&t); return 0;\n}\n")
    inc = os.path.join(kqed_amd.ROOT, "include")
    libdir = os.path.dirname(kqed_amd.LIB_PATH)
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, str(src), "-L", libdir, "-lkqed_hip",
                    "-Wl,-rpath," + libdir, "-lm", "-o", str(exe)], check=True)
    assert exe.exists()
