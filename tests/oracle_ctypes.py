"""ctypes binding of the CPU oracle (oracle/libkqed_oracle.so) for the test-suite.

Test infrastructure: the oracle is the checker, never the thing measured or shipped.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "libkqed_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class _Tables(ctypes.Structure):
    _fields_ = [
        ("nstpx", ctypes.c_int), ("nstpy", ctypes.c_int), ("Nffa", ctypes.c_int), ("nfmax", ctypes.c_int),
        ("XX", _dp), ("YY", _dp), ("xstp", ctypes.c_double), ("ystp", ctypes.c_double),
        ("nfx", ctypes.POINTER(ctypes.c_int)),
        ("Ffm", ctypes.POINTER(ctypes.c_float)), ("Ffp", ctypes.POINTER(ctypes.c_float)),
        ("TX", (ctypes.c_double * 810) * 23), ("TY", (ctypes.c_double * 100) * 14),
        ("NY_tay", ctypes.c_int), ("NX_tay", ctypes.c_int),
        ("LUT", _dp * 7), ("nLUT", ctypes.c_int),
    ]


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    T = ctypes.POINTER(_Tables)
    v4 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.oq_load.argtypes = [ctypes.c_char_p, T]
    L.oq_free.argtypes = [T]
    for n in ("oq_L0", "oq_L1", "oq_L2", "oq_L3"):
        getattr(L, n).argtypes = [T, v4, v4, v4]
    L.oq_ML2.argtypes = [T, ctypes.c_double, v4, v4, v4]
    L.oq_batch.argtypes = [T, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, v4, v4, v4, ctypes.c_int]
    f4 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.oq_contract.argtypes = [T, ctypes.c_int, ctypes.c_double, ctypes.c_size_t, v4, f4, ctypes.c_size_t, v4, v4, v4,
                              ctypes.c_int]
    L.oq_comb_len.argtypes = [ctypes.c_int]
    L.oq_combine.argtypes = [T, ctypes.c_int, v4, v4, v4, v4]
    L.oq_combine_batch.argtypes = [T, ctypes.c_int, v4, ctypes.c_size_t, v4, v4, v4, ctypes.c_int]
    L.oq_contract_comb.argtypes = [T, ctypes.c_int, v4, ctypes.c_size_t, v4, f4, ctypes.c_size_t, v4, v4, v4,
                                   ctypes.c_int]
    L.oq_chebU.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, v4]
    L.oq_chebU.restype = ctypes.c_double
    L.oq_chebU_avx.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, v4, T]
    L.oq_chebU_avx.restype = ctypes.c_double
    L.oq_stv.argtypes = [T, v4, v4, v4]
    L.oq_tabd_xeq0.argtypes = [T, v4, v4]
    L.oq_tabd_yeq0.argtypes = [T, v4, v4]
    L.oq_assembly_maps.argtypes = [v4, v4]
    L.oq_interp_scalars.argtypes = [T, v4, v4, v4]
    L.oq_regime.argtypes = [T, v4, v4]
    L.oq_stv_inject.argtypes = [T, v4, v4, v4, v4]
    L.oq_L0_inject.argtypes = [T, v4, v4, v4, v4]
    L.oq_plant.argtypes = [ctypes.c_int]
    L.oq_vterms.argtypes = [T, v4, v4, v4]
    L.oq_flops_get.restype = ctypes.c_double
    L.oq_flops_build_get.restype = ctypes.c_double
    L.oq_crc32c.argtypes = [ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
    L.oq_crc32c.restype = ctypes.c_uint32
    return L


class Oracle:
    """The CPU oracle over one table directory."""

    def __init__(self, table_dir):
        self.L = _lib()
        self.t = _Tables()
        rc = self.L.oq_load(table_dir.encode(), ctypes.byref(self.t))
        if rc != 0:
            raise RuntimeError(f"oracle could not load tables from {table_dir}")
        self.XX = np.ctypeslib.as_array(self.t.XX, (self.t.nstpx,)).copy()
        self.YY = np.ctypeslib.as_array(self.t.YY, (self.t.nstpy,)).copy()
        self.nfx = np.ctypeslib.as_array(self.t.nfx, (self.t.nstpx,)).copy()
        self.TX = np.array(self.t.TX)
        self.TY = np.array(self.t.TY)

    def close(self):
        if self.t is not None:
            self.L.oq_free(ctypes.byref(self.t))
            self.t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _k(self, name, x, y, M=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        k = np.zeros(384)
        if M is None:
            getattr(self.L, name)(ctypes.byref(self.t), x, y, k)
        else:
            self.L.oq_ML2(ctypes.byref(self.t), float(M), x, y, k)
        return k

    def L0(self, x, y):
        return self._k("oq_L0", x, y)

    def L1(self, x, y):
        return self._k("oq_L1", x, y)

    def L2(self, x, y):
        return self._k("oq_L2", x, y)

    def L3(self, x, y):
        return self._k("oq_L3", x, y)

    def ML2(self, M, x, y):
        return self._k(None, x, y, M)

    def batch(self, which, x, y, M=0.0, nthreads=0):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        n = x.shape[0]
        out = np.zeros((n, 384))
        self.L.oq_batch(ctypes.byref(self.t), int(which), float(M), n, x.ravel(), y.ravel(), out.ravel(), int(nthreads))
        return out

    def contract(self, which, x, Tc, y, M=0.0, nthreads=0):
        """s[j] = sum_i L(x_i, y_j) . Tc[i] and the error scale a[j] = sum_i sum_c |L_c Tc[i][c]|."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        Tc = np.ascontiguousarray(Tc, np.float32)
        ny = y.shape[0]
        s = np.zeros(ny)
        a = np.zeros(ny)
        self.L.oq_contract(ctypes.byref(self.t), int(which), float(M), x.shape[0], x.ravel(), Tc.ravel(), ny, y.ravel(),
                           s, a, int(nthreads))
        return s, a

    def combine(self, comb, x, y, M=None):
        """One pair through a kernel combination: array (nk, 384) in the reference's struct order."""
        nk = self.L.oq_comb_len(int(comb)) // 384
        out = np.zeros(nk * 384)
        M4 = np.zeros(4) if M is None else np.ascontiguousarray(M, np.float64)
        rc = self.L.oq_combine(ctypes.byref(self.t), int(comb), M4, np.ascontiguousarray(x, np.float64),
                               np.ascontiguousarray(y, np.float64), out)
        if rc:
            raise ValueError(f"bad combination {comb}")
        return out.reshape(nk, 384)

    def combine_batch(self, comb, x, y, M=None, nthreads=0):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        nk = self.L.oq_comb_len(int(comb)) // 384
        out = np.zeros((x.shape[0], nk, 384))
        M4 = np.zeros(4) if M is None else np.ascontiguousarray(M, np.float64)
        self.L.oq_combine_batch(ctypes.byref(self.t), int(comb), M4, x.shape[0], x.ravel(), y.ravel(), out.ravel(),
                                int(nthreads))
        return out

    def contract_comb(self, comb, x, Tc, y, M=None, nthreads=0):
        """s[j, q] = sum_i K_q(x_i, y_j) . Tc[i] for the output kernels of a combination, and the error
        scale a[j, q] = sum of |terms|."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 4)
        Tc = np.ascontiguousarray(Tc, np.float32)
        nk = self.L.oq_comb_len(int(comb)) // 384
        s = np.zeros((y.shape[0], nk))
        a = np.zeros((y.shape[0], nk))
        M4 = np.zeros(4) if M is None else np.ascontiguousarray(M, np.float64)
        self.L.oq_contract_comb(ctypes.byref(self.t), int(comb), M4, x.shape[0], x.ravel(), Tc.ravel(), y.shape[0],
                                y.ravel(), s.ravel(), a.ravel(), int(nthreads))
        return s, a

    def stv(self, x, y):
        s = np.zeros(200)
        rc = self.L.oq_stv(ctypes.byref(self.t), np.asarray(x, np.float64), np.asarray(y, np.float64), s)
        return rc, s

    def tabd_xeq0(self, y):
        t = np.zeros(192)
        rc = self.L.oq_tabd_xeq0(ctypes.byref(self.t), np.asarray(y, np.float64), t)
        return rc, t

    def tabd_yeq0(self, x):
        t = np.zeros(192)
        rc = self.L.oq_tabd_yeq0(ctypes.byref(self.t), np.asarray(x, np.float64), t)
        return rc, t

    def scalars(self, x, y):
        s = np.zeros(45)
        reg = self.L.oq_interp_scalars(ctypes.byref(self.t), np.asarray(x, np.float64), np.asarray(y, np.float64), s)
        return reg, s

    def stv_inject(self, x, y, w):
        """S/T/V (oq_stv layout) of a standard-regime pair with the 34 weight functions w injected."""
        s = np.zeros(200)
        rc = self.L.oq_stv_inject(ctypes.byref(self.t), np.asarray(x, np.float64), np.asarray(y, np.float64),
                                  np.ascontiguousarray(w, np.float64), s)
        return rc, s

    def L0_inject(self, x, y, w):
        k = np.zeros(384)
        self.L.oq_L0_inject(ctypes.byref(self.t), np.asarray(x, np.float64), np.asarray(y, np.float64),
                            np.ascontiguousarray(w, np.float64), k)
        return k

    def vterms(self, x, y):
        """The 27 g1/g2/g3 derivatives chnr_dV passes to its tensor step (oq_vterms)."""
        v = np.zeros(27)
        rc = self.L.oq_vterms(ctypes.byref(self.t), np.asarray(x, np.float64), np.asarray(y, np.float64), v)
        return rc, v

    def plant(self, k):
        """Switch on deliberate mistake k of the oracle (0 = none); see oq_plant."""
        self.L.oq_plant(int(k))

    def regime(self, x, y):
        return self.L.oq_regime(ctypes.byref(self.t), np.asarray(x, np.float64), np.asarray(y, np.float64))

    def chebU(self, order, co, x):
        co = np.ascontiguousarray(co, np.float64)
        return self.L.oq_chebU(int(order), len(co), float(x), co)

    def chebU_avx(self, order, co, x):
        co = np.ascontiguousarray(co, np.float64)
        return self.L.oq_chebU_avx(int(order), len(co), float(x), co, ctypes.byref(self.t))

    def flops(self):
        return self.L.oq_flops_get()

    def flops_build(self):
        return self.L.oq_flops_build_get()

    def flops_reset(self):
        self.L.oq_flops_reset()


def assembly_maps():
    L = _lib()
    a = np.zeros(384 * 200)
    b = np.zeros(384 * 192)
    L.oq_assembly_maps(a, b)
    return a.reshape(384, 200), b.reshape(384, 192)
