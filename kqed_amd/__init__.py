"""kqed_amd -- MI355X (gfx950) implementation of KQED's position-space QED kernel.

Host-side mirror of the reference C interface (reference src/HEADERS/init.h, kernels.h) over the
C ABI of ``libkqed_hip.so`` (declared in ``include/kqed.h``):

    t = QEDKernelTemps.initialise(table_dir)      # initialise()
    k = t.QED_kernel_L0(x, y)                     # double kerv[6][4][4][4] as a (6,4,4,4) array
    t.QED_kernel_L1 / L2 / L3, t.QED_Mkernel_L2(M, x, y)
    t.free()                                      # free_QED_temps()

plus the batched entry points (host arrays or device pointers).  Every evaluation runs in the HIP
library; there is no CPU fallback, and loading fails loudly when the library is missing.
"""
import ctypes
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libkqed_hip.so")
# the reference's shipped Taylor-x and Taylor-y tables (reference PRECOMP/taylor{x,y}_cksum.bin),
# package data: every table directory this package builds pairs them with a form-factor table
TAYLOR_TABLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")

KQED_L0, KQED_L1, KQED_L2, KQED_L3, KQED_ML2 = range(5)
KQED_AOS, KQED_SOA, KQED_AOS_F32, KQED_SOA_F32 = 0, 1, 2, 3
KQED_SORT_CELL, KQED_SORT_XNODE, KQED_SORT_NONE = 0, 1, 2
# kernel combinations (reference all_kernels.c, con_kernel.c, sub_kernel.c, SYMXY.c, SYMXY0.c)
(KQED_COMB_ALL, KQED_COMB_ALL_M, KQED_COMB_ALL_M_V2, KQED_COMB_CON, KQED_COMB_CON_V2, KQED_COMB_CON_M_L2,
 KQED_COMB_SUB_M_L2, KQED_COMB_SYMXY, KQED_COMB_SYMXY_V2, KQED_COMB_SYMXY0, KQED_COMB_SYMXY0_V2) = range(11)

_lib = None


class _Grid(ctypes.Structure):
    _fields_ = [
        ("Nffa", ctypes.c_int), ("XX", ctypes.POINTER(ctypes.c_double)), ("YY", ctypes.POINTER(ctypes.c_double)),
        ("xstp", ctypes.c_double), ("ystp", ctypes.c_double), ("nstpx", ctypes.c_int), ("nstpy", ctypes.c_int),
        ("NtayY", ctypes.c_int), ("NY_tay", ctypes.c_int), ("NtayX", ctypes.c_int), ("NX_tay", ctypes.c_int),
        ("nfx", ctypes.POINTER(ctypes.c_int)),
    ]


class _Temps(ctypes.Structure):
    _fields_ = [("Grid", _Grid), ("G8", ctypes.POINTER(ctypes.c_int)), ("dev", ctypes.c_void_p)]


class Tune(ctypes.Structure):
    """kqed_tune: pass length, assembly grid, sort granularity, sort threshold."""
    _fields_ = [("chunk", ctypes.c_size_t), ("asm_grid", ctypes.c_int), ("sort_mode", ctypes.c_int),
                ("sort_min", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def lib():
    """Load libkqed_hip.so (built by kqed_amd.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m kqed_amd.build`")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 (same soname, another path).
    # Loading torch first lets this library bind to that runtime instead of starting a second one.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(_Temps)
    arr = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.initialise.argtypes = [P]
    L.kqed_initialise_dir.argtypes = [P, ctypes.c_char_p]
    L.free_QED_temps.argtypes = [P]
    for name in ("QED_kernel_L0", "QED_kernel_L1", "QED_kernel_L2", "QED_kernel_L3"):
        getattr(L, name).argtypes = [arr, arr, _Temps, arr]
        getattr(L, name).restype = None
    L.QED_Mkernel_L2.argtypes = [ctypes.c_double, arr, arr, _Temps, arr]
    L.QED_Mkernel_L2.restype = None
    L.kqed_batch.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.kqed_batch_host.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, arr, arr, ctypes.c_void_p,
                                  ctypes.c_int]
    szp = np.ctypeslib.ndpointer(np.uintp, flags="C_CONTIGUOUS")
    L.kqed_batch_ragged.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, szp, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.kqed_graph_create_ragged.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, szp, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)]
    L.kqed_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.kqed_graph_destroy.argtypes = [ctypes.c_void_p]
    L.kqed_graph_destroy.restype = None
    vp, f32p = ctypes.c_void_p, np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.kqed_surrogate_create.argtypes = [P, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.kqed_surrogate_destroy.argtypes = [vp]
    L.kqed_surrogate_destroy.restype = None
    L.kqed_surrogate_train.argtypes = [vp, ctypes.c_size_t, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]
    L.kqed_surrogate_gradients.argtypes = [vp, ctypes.c_int, f32p, f32p, f32p, f32p, ctypes.POINTER(ctypes.c_double)]
    L.kqed_surrogate_params.argtypes = [vp, vp, vp, vp, vp]
    L.kqed_surrogate_nparam.restype = ctypes.c_size_t
    L.kqed_surrogate_validate.argtypes = [vp, ctypes.c_size_t, vp, vp, ctypes.c_double, ctypes.c_int, arr]
    L.kqed_batch_surrogate.argtypes = [P, vp, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, vp, vp, vp, ctypes.c_int,
                                       vp]
    L.kqed_batch_inject.argtypes = [P, ctypes.c_size_t, vp, vp, vp, vp, ctypes.c_int, vp]
    L.kqed_surrogate_save.argtypes = [vp, ctypes.c_char_p]
    L.kqed_surrogate_load.argtypes = [vp, ctypes.c_char_p]
    L.kqed_comb_kernels.argtypes = [ctypes.c_int]
    L.kqed_combine.argtypes = [P, ctypes.c_int, vp, ctypes.c_size_t, vp, vp, vp, vp]
    L.kqed_combine_host.argtypes = [P, ctypes.c_int, vp, ctypes.c_size_t, arr, arr, vp]
    L.kqed_contract_comb.argtypes = [P, ctypes.c_int, vp, ctypes.c_size_t, vp, vp, ctypes.c_size_t, arr, arr, vp]
    L.kqed_pihat.argtypes = [P, ctypes.c_size_t, vp, vp, vp, vp]
    L.kqed_pihat_host.argtypes = [P, ctypes.c_size_t, arr, arr, arr]
    L.kqed_lepton_loop.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_size_t, vp, vp, vp,
                                   vp]
    L.pihat1.argtypes = [arr, arr, _Temps, arr, arr]
    L.pihat1.restype = None
    for name in ("amu_heavlept", "amu_lightlept", "vac_pol"):
        getattr(L, name).argtypes = [ctypes.c_double]
        getattr(L, name).restype = ctypes.c_double
    L.bessk.argtypes = [ctypes.c_int, ctypes.c_double, arr]
    L.bessk.restype = None
    L.diractrace.argtypes = [ctypes.c_int, np.ctypeslib.ndpointer(np.intc, flags="C_CONTIGUOUS")]
    L.kqed_get_tune.argtypes = [P, ctypes.POINTER(Tune)]
    L.kqed_set_tune.argtypes = [P, ctypes.POINTER(Tune)]
    L.kqed_autotune.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(Tune),
                                ctypes.POINTER(ctypes.c_double)]
    L.kqed_contract.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_size_t, arr, arr, ctypes.c_void_p]
    L.kqed_batch_workspace.argtypes = [ctypes.c_size_t]
    L.kqed_batch_workspace.restype = ctypes.c_size_t
    L.kqed_last_error.restype = ctypes.c_char_p
    L.kqed_profile.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    L.kqed_write_synthetic_ff.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong]
    _lib = L
    return L


# ----------------------------------------------------------------------------- table directories
# Synthetic form-factor tables (the reference's real FFxy_single_cksum.bin is not available):
#   "full":  100 x-nodes (the reference's x-grid layout), 810 y-nodes on the Taylor-x y step
#            (6.54/810), 32..64 Chebyshev terms per x-node (capped by float range at small x)
#   "small": the same x-grid, 218 y-nodes of step 0.03, 8..16 terms (fast CPU tests)
TABLES = {
    "full": dict(nstpx=100, nstpy=810, nf0=32, nf1=64, xh=6.54 / 810, yh=6.54 / 810, seed=20240611),
    "small": dict(nstpx=100, nstpy=218, nf0=8, nf1=16, xh=6.54 / 810, yh=0.03, seed=20240611),
    # the longest Chebyshev rows the reference's GPU port admits (nk = nfx + 2 <= 128)
    "wide": dict(nstpx=100, nstpy=64, nf0=96, nf1=127, xh=6.54 / 810, yh=0.1, seed=20240611),
}


def table_dir(name="small", root=None):
    """Directory holding the three table files for the named synthetic table (created on demand).

    Uses $KQED_TABLE_DIR unchanged when it is set (e.g. the real form-factor table)."""
    env = os.environ.get("KQED_TABLE_DIR")
    if env and name == "env":
        return env
    root = root or os.environ.get("KQED_TABLE_CACHE", os.path.join(ROOT, "_tables"))
    d = os.path.join(root, name)
    ff = os.path.join(d, "FFxy_single_cksum.bin")
    if not os.path.exists(ff):
        os.makedirs(d, exist_ok=True)
        p = TABLES[name]
        tmp = ff + f".tmp{os.getpid()}"
        rc = lib().kqed_write_synthetic_ff(tmp.encode(), p["nstpx"], p["nstpy"], p["nf0"], p["nf1"], p["xh"], p["yh"],
                                           p["seed"])
        if rc != 0:
            raise RuntimeError("synthetic table generation failed")
        os.replace(tmp, ff)
    for f in ("taylorx_cksum.bin", "taylory_cksum.bin"):
        dst = os.path.join(d, f)
        if not os.path.exists(dst):
            shutil.copyfile(os.path.join(TAYLOR_TABLES, f), dst)
    return d


# ----------------------------------------------------------------------------- kernels
class QEDKernelTemps:
    """struct QED_kernel_temps of the reference, owning device-resident tables."""

    def __init__(self, t):
        self._t = t

    @classmethod
    def initialise(cls, table_dir=None):
        L = lib()
        t = _Temps()
        rc = L.kqed_initialise_dir(ctypes.byref(t), table_dir.encode()) if table_dir else L.initialise(ctypes.byref(t))
        if rc != 0:
            raise RuntimeError("initialise failed: " + L.kqed_last_error().decode())
        return cls(t)

    def free(self):
        if self._t is not None and self._t.dev:
            lib().free_QED_temps(ctypes.byref(self._t))
        self._t = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def XX(self):
        g = self._t.Grid
        return np.ctypeslib.as_array(g.XX, (g.nstpx,)).copy()

    @property
    def YY(self):
        g = self._t.Grid
        return np.ctypeslib.as_array(g.YY, (g.nstpy,)).copy()

    @property
    def nfx(self):
        g = self._t.Grid
        return np.ctypeslib.as_array(g.nfx, (g.nstpx,)).copy()

    @property
    def G8(self):
        return np.ctypeslib.as_array(self._t.G8, (65536,)).copy()

    def _one(self, name, x, y, M=None):
        x = np.ascontiguousarray(x, np.float64).reshape(4)
        y = np.ascontiguousarray(y, np.float64).reshape(4)
        k = np.zeros(384)
        if M is None:
            getattr(lib(), name)(x, y, self._t, k)
        else:
            lib().QED_Mkernel_L2(float(M), x, y, self._t, k)
        return k.reshape(6, 4, 4, 4)

    def QED_kernel_L0(self, x, y):
        return self._one("QED_kernel_L0", x, y)

    def QED_kernel_L1(self, x, y):
        return self._one("QED_kernel_L1", x, y)

    def QED_kernel_L2(self, x, y):
        return self._one("QED_kernel_L2", x, y)

    def QED_kernel_L3(self, x, y):
        return self._one("QED_kernel_L3", x, y)

    def QED_Mkernel_L2(self, M, x, y):
        return self._one(None, x, y, M)

    def batch_host(self, which, x, y, M=0.0, layout=KQED_AOS):
        """n pairs from host arrays x[n,4], y[n,4] -> (n,384) (AoS) or (384,n) (SoA), float64 or float32."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        n = x.shape[0]
        out = np.zeros((n, 384) if layout in (KQED_AOS, KQED_AOS_F32) else (384, n),
                       np.float32 if layout in (KQED_AOS_F32, KQED_SOA_F32) else np.float64)
        rc = lib().kqed_batch_host(ctypes.byref(self._t), int(which), float(M), n, x.reshape(-1), y.reshape(-1),
                                   out.ctypes.data, int(layout))
        if rc != 0:
            raise RuntimeError("kqed_batch_host failed: " + lib().kqed_last_error().decode())
        return out

    def combine(self, comb, x, y, M=None):
        """A kernel combination (KQED_COMB_*) for n pairs of host arrays x[n,4], y[n,4] ->
        (n, nk, 384) float64, nk kernels per pair in the reference's struct order; M: 4 masses for
        the M variants."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 4)
        nk = lib().kqed_comb_kernels(int(comb))
        if nk == 0:
            raise ValueError(f"unknown combination {comb}")
        n = x.shape[0]
        out = np.zeros((n, nk, 384))
        M4 = None if M is None else np.ascontiguousarray(M, np.float64).reshape(4)
        rc = lib().kqed_combine_host(ctypes.byref(self._t), int(comb), None if M4 is None else M4.ctypes.data, n,
                                     x.reshape(-1), y.reshape(-1), out.ctypes.data)
        if rc != 0:
            raise RuntimeError("kqed_combine_host failed: " + lib().kqed_last_error().decode())
        return out

    def combine_device(self, comb, n, x_ptr, y_ptr, out_ptr, M=None, stream=None):
        """A kernel combination on device pointers (ints); out holds n * 384 * kqed_comb_kernels(comb) doubles."""
        M4 = None if M is None else np.ascontiguousarray(M, np.float64).reshape(4)
        rc = lib().kqed_combine(ctypes.byref(self._t), int(comb), None if M4 is None else M4.ctypes.data, int(n),
                                ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), ctypes.c_void_p(out_ptr),
                                ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_combine failed: " + lib().kqed_last_error().decode())

    def contract_comb_device(self, comb, nx, x_ptr, T_ptr, y, M=None, stream=None):
        """s[j, q] = sum_i K_q(x_i, y_j) . T_i for the output kernels of a combination; x[nx][4] FP64 and
        T[nx][384] FP32 on the device (pointers), y on the host."""
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 4)
        nk = lib().kqed_comb_kernels(int(comb))
        s = np.zeros((y.shape[0], nk))
        M4 = None if M is None else np.ascontiguousarray(M, np.float64).reshape(4)
        rc = lib().kqed_contract_comb(ctypes.byref(self._t), int(comb), None if M4 is None else M4.ctypes.data, int(nx),
                                      ctypes.c_void_p(x_ptr), ctypes.c_void_p(T_ptr), y.shape[0], y.reshape(-1),
                                      s.reshape(-1), ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_contract_comb failed: " + lib().kqed_last_error().decode())
        return s

    def pihat(self, x, y):
        """ipihatFermLoop_antisym for n pairs of host arrays -> (n, 384) float64 (computed on the GPU)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 4)
        out = np.zeros((x.shape[0], 384))
        rc = lib().kqed_pihat_host(ctypes.byref(self._t), x.shape[0], x.reshape(-1), y.reshape(-1), out.reshape(-1))
        if rc != 0:
            raise RuntimeError("kqed_pihat_host failed: " + lib().kqed_last_error().decode())
        return out

    def pihat1(self, x, y):
        """pihat1 of one pair -> (v1[4,4,4,4,4], r1[4,4,4,4])."""
        v = np.zeros(1024)
        r = np.zeros(256)
        lib().pihat1(np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64), self._t, v, r)
        return v.reshape(4, 4, 4, 4, 4), r.reshape(4, 4, 4, 4)

    def pihat_device(self, n, x_ptr, y_ptr, out_ptr, stream=None):
        rc = lib().kqed_pihat(ctypes.byref(self._t), int(n), ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr),
                              ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_pihat failed: " + lib().kqed_last_error().decode())

    def lepton_loop_device(self, which, n, x_ptr, y_ptr, out_ptr, M=0.0, Mv=1.0, stream=None):
        """out[i] = sum_c L(x_i, y_i)_c iPihat(Mv x_i, Mv y_i)_c on device pointers."""
        rc = lib().kqed_lepton_loop(ctypes.byref(self._t), int(which), float(M), float(Mv), int(n),
                                    ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), ctypes.c_void_p(out_ptr),
                                    ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_lepton_loop failed: " + lib().kqed_last_error().decode())

    def get_tune(self):
        tu = Tune()
        lib().kqed_get_tune(ctypes.byref(self._t), ctypes.byref(tu))
        return tu

    def set_tune(self, tune=None, **kw):
        tu = self.get_tune() if tune is None else tune
        for k, v in kw.items():
            setattr(tu, k, v)
        if lib().kqed_set_tune(ctypes.byref(self._t), ctypes.byref(tu)) != 0:
            raise ValueError("kqed_set_tune: " + lib().kqed_last_error().decode())
        return tu

    def autotune(self, which, n, x_ptr, y_ptr, out_ptr, M=0.0, layout=KQED_AOS, reps=2, stream=None):
        """Run the library's autotuner on a device batch; installs and returns (Tune, ms per run)."""
        tu = Tune()
        ms = ctypes.c_double()
        rc = lib().kqed_autotune(ctypes.byref(self._t), int(which), float(M), int(n), ctypes.c_void_p(x_ptr),
                                 ctypes.c_void_p(y_ptr), ctypes.c_void_p(out_ptr), int(layout), int(reps),
                                 ctypes.c_void_p(stream or 0), ctypes.byref(tu), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError("kqed_autotune failed: " + lib().kqed_last_error().decode())
        return tu, ms.value

    def batch_ragged(self, which, offsets, y_ptr, x_ptr, out_ptr, M=0.0, layout=KQED_AOS, stream=None):
        off = np.ascontiguousarray(offsets, np.uintp)
        rc = lib().kqed_batch_ragged(ctypes.byref(self._t), int(which), float(M), len(off) - 1, off,
                                     ctypes.c_void_p(y_ptr), ctypes.c_void_p(x_ptr), ctypes.c_void_p(out_ptr),
                                     int(layout), ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_batch_ragged failed: " + lib().kqed_last_error().decode())

    def graph_ragged(self, which, offsets, y_ptr, x_ptr, out_ptr, M=0.0, layout=KQED_AOS):
        off = np.ascontiguousarray(offsets, np.uintp)
        g = ctypes.c_void_p()
        rc = lib().kqed_graph_create_ragged(ctypes.byref(self._t), int(which), float(M), len(off) - 1, off,
                                            ctypes.c_void_p(y_ptr), ctypes.c_void_p(x_ptr), ctypes.c_void_p(out_ptr),
                                            int(layout), ctypes.byref(g))
        if rc != 0:
            raise RuntimeError("kqed_graph_create_ragged failed: " + lib().kqed_last_error().decode())
        return RaggedGraph(g)

    def profile(self, enable):
        """Read-and-reset the per-stage kernel times (ms, launches) and switch timing on/off."""
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_long * 3)()
        lib().kqed_profile(ctypes.byref(self._t), int(bool(enable)), ms, cnt)
        names = ("sort", "interp", "assemble")
        return {n: (ms[i], cnt[i]) for i, n in enumerate(names)}

    def contract_device(self, which, nx, x_ptr, T_ptr, y, M=0.0, stream=None):
        """s[j] = sum_i L(x_i, y_j) . T_i for device x[nx][4] (FP64) and T[nx][384] (FP32); y on the host."""
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 4)
        s = np.zeros(y.shape[0])
        rc = lib().kqed_contract(ctypes.byref(self._t), int(which), float(M), int(nx), ctypes.c_void_p(x_ptr),
                                 ctypes.c_void_p(T_ptr), y.shape[0], y.reshape(-1), s, ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_contract failed: " + lib().kqed_last_error().decode())
        return s

    def contract(self, which, x, T, y, M=0.0):
        """Host arrays in, host result out (the device copies are made with torch)."""
        import torch
        xd = torch.from_numpy(np.ascontiguousarray(x, np.float64)).to("cuda")
        Td = torch.from_numpy(np.ascontiguousarray(T, np.float32)).to("cuda")
        s = self.contract_device(which, xd.shape[0], xd.data_ptr(), Td.data_ptr(), y, M=M,
                                 stream=torch.cuda.current_stream().cuda_stream)
        return s

    def batch_device(self, which, n, x_ptr, y_ptr, out_ptr, M=0.0, layout=KQED_AOS, stream=None):
        """n pairs on device pointers (ints); enqueued on `stream` (an int hipStream_t or None)."""
        rc = lib().kqed_batch(ctypes.byref(self._t), int(which), float(M), int(n), ctypes.c_void_p(x_ptr),
                              ctypes.c_void_p(y_ptr), ctypes.c_void_p(out_ptr), int(layout),
                              ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_batch failed: " + lib().kqed_last_error().decode())


    def batch_inject_device(self, n, x_ptr, y_ptr, w_ptr, out_ptr, layout=KQED_AOS, stream=None):
        """L0 of n device pairs with the packed-path weight functions taken from w_ptr ([n][34] FP64,
        order of kqed_batch_inject) instead of the tables."""
        rc = lib().kqed_batch_inject(ctypes.byref(self._t), int(n), ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr),
                                     ctypes.c_void_p(w_ptr), ctypes.c_void_p(out_ptr), int(layout),
                                     ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_batch_inject failed: " + lib().kqed_last_error().decode())


class RaggedGraph:
    """A captured ragged batch (kqed_graph); launch() replays it on a stream."""

    def __init__(self, g):
        self._g = g

    def launch(self, stream=None):
        rc = lib().kqed_graph_launch(self._g, ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError("kqed_graph_launch failed: " + lib().kqed_last_error().decode())

    def destroy(self):
        if self._g:
            lib().kqed_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ----------------------------------------------------------------------------- surrogate
SURR_NB = 24
SURR_K = (16, 128, 128, 128, 128)
SURR_N = (128, 128, 128, 128, 48)
SURR_NOUT = 34


def _err(name):
    return RuntimeError(name + " failed: " + lib().kqed_last_error().decode())


class Surrogate:
    """The learned surrogate of the packed-path interpolation (kqed_surrogate_*)."""

    def __init__(self, temps, seed=1):
        self.temps = temps
        self._g = ctypes.c_void_p()
        if lib().kqed_surrogate_create(ctypes.byref(temps._t), int(seed), ctypes.byref(self._g)) != 0:
            raise _err("kqed_surrogate_create")

    def destroy(self):
        if self._g:
            lib().kqed_surrogate_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def train(self, n, x_ptr, y_ptr, epochs=4, batch=8192, lr=2e-3, seed=5):
        loss = ctypes.c_double()
        if lib().kqed_surrogate_train(self._g, int(n), x_ptr, y_ptr, int(epochs), int(batch), float(lr), int(seed),
                                      ctypes.byref(loss)) != 0:
            raise _err("kqed_surrogate_train")
        return loss.value

    def gradients(self, X3, tgt, wgt):
        X3 = np.ascontiguousarray(X3, np.float32)
        tgt = np.ascontiguousarray(tgt, np.float32)
        wgt = np.ascontiguousarray(wgt, np.float32)
        g = np.zeros(lib().kqed_surrogate_nparam(), np.float32)
        loss = ctypes.c_double()
        if lib().kqed_surrogate_gradients(self._g, X3.shape[0], X3.reshape(-1), tgt.reshape(-1), wgt.reshape(-1), g,
                                          ctypes.byref(loss)) != 0:
            raise _err("kqed_surrogate_gradients")
        return loss.value, g

    def params(self):
        P = np.zeros(lib().kqed_surrogate_nparam(), np.float32)
        lsc = np.zeros(SURR_NOUT * SURR_NB * SURR_NB, np.float32)
        mask = np.zeros(SURR_NB * SURR_NB, np.uint8)
        rng = np.zeros(4, np.float32)
        if lib().kqed_surrogate_params(self._g, P.ctypes.data, lsc.ctypes.data, mask.ctypes.data, rng.ctypes.data) != 0:
            raise _err("kqed_surrogate_params")
        return P, lsc.reshape(SURR_NOUT, SURR_NB, SURR_NB), mask.reshape(SURR_NB, SURR_NB), rng

    def validate(self, n, x_ptr, y_ptr, tol=1e-5, min_count=8):
        st = np.zeros(5)
        if lib().kqed_surrogate_validate(self._g, int(n), x_ptr, y_ptr, float(tol), int(min_count), st) != 0:
            raise _err("kqed_surrogate_validate")
        return {"pairs": int(st[0]), "packed": int(st[1]), "packed_in_passing_bins": int(st[2]),
                "max_err": st[3], "median_err": st[4]}

    def batch_device(self, which, n, x_ptr, y_ptr, out_ptr, M=0.0, layout=KQED_AOS, stream=None):
        if lib().kqed_batch_surrogate(ctypes.byref(self.temps._t), self._g, int(which), float(M), int(n), x_ptr, y_ptr,
                                      out_ptr, int(layout), stream or 0) != 0:
            raise _err("kqed_batch_surrogate")

    def save(self, path):
        if lib().kqed_surrogate_save(self._g, path.encode()) != 0:
            raise _err("kqed_surrogate_save")

    def load(self, path):
        if lib().kqed_surrogate_load(self._g, path.encode()) != 0:
            raise _err("kqed_surrogate_load")


def surrogate_layers(P):
    """Split a flat parameter vector into [(W_l [K][N], b_l [N])] (FP32)."""
    out, o = [], 0
    for K, N in zip(SURR_K, SURR_N):
        W = P[o:o + K * N].reshape(K, N)
        o += K * N
        b = P[o:o + N]
        o += N
        out.append((W, b))
    return out
