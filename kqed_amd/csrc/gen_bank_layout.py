"""Generate kqed_bank_layout.h: LDS slots of the standard assembly map's tensor entries and the
lane order of its rows, chosen so that the map's tensor gathers and output-row writes hit
distinct LDS banks.

    python kqed_amd/csrc/gen_bank_layout.py     # needs kqed_amd/libkqed_hip.so (for the map)

The assembly kernels read, at unrolled step k of row slot m, the k-th column of the row held by
each lane, as 16-byte (two-pair) or 8-byte values.  Neither the slot an entry is stored in nor the
lane a row of a slot group is dealt to changes any arithmetic (each row keeps its own term order),
only the LDS addresses, so both are free.  They are chosen by simulated annealing (fixed seed)
against the number of LDS cycles of those accesses under two banking models of a 128-bit access
(8 lanes over 8 groups of 16 bytes, 16 lanes over 16 groups), the first of which reproduces the
bank-conflict counter of the kernel with the natural layout.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KB = [13, 9, 6, 6, 4, 3]
STARTS = np.cumsum([0] + KB)


class _Map(ctypes.Structure):
    _fields_ = [("rowptr", ctypes.c_int * 385), ("col", ctypes.c_short * 8192), ("val", ctypes.c_double * 8192),
                ("nnz", ctypes.c_int), ("kmax", ctypes.c_int)]


def standard_map():
    lib = ctypes.CDLL(os.path.join(os.path.dirname(HERE), "libkqed_hip.so"))
    maps = [_Map(), _Map(), _Map()]
    if lib.kqed_build_maps(*[ctypes.byref(m) for m in maps]) != 0:
        raise RuntimeError("kqed_build_maps failed")
    m = maps[0]
    rp = np.array(m.rowptr)
    return rp, np.array(m.col[:m.nnz])


def cycles(addr, G, NB):
    """LDS cycles of [steps][64] 16-byte slot addresses: per group of G lanes, the largest number of
    distinct addresses in one of NB bank groups (equal addresses are one broadcast)."""
    ns = addr.shape[0]
    s = np.sort(addr.reshape(ns, 64 // G, G), axis=2)
    uniq = np.ones_like(s, dtype=bool)
    uniq[:, :, 1:] = s[:, :, 1:] != s[:, :, :-1]
    gid = np.arange(ns * (64 // G)).reshape(ns, 64 // G, 1)
    cnt = np.bincount((gid * NB + s % NB)[uniq], minlength=ns * (64 // G) * NB).reshape(-1, NB)
    return int(cnt.max(axis=1).sum())


def main():
    rp, col = standard_map()
    lens = np.diff(rp)
    order = sorted(range(384), key=lambda r: -lens[r])  # as build_wave_maps: stable, longest first
    rows = np.array(order).reshape(6, 64)               # rows[m][lane]
    steps = []
    for m in range(6):
        for k in range(KB[m]):
            steps.append([col[rp[r] + k] if k < lens[r] else 200 for r in rows[m]])
    S = np.array(steps)                                   # [41][64] columns read

    def cost(pi, S, rows):
        a = pi[S]
        return cycles(a, 8, 8) + cycles(a, 16, 16) + cycles(rows, 8, 8) + cycles(rows, 16, 16)

    pi = np.arange(201)
    rng = np.random.default_rng(20240611)
    cur = start = cost(pi, S, rows)
    T = 2.0
    for _ in range(120000):
        if rng.random() < 0.5:
            a, b = rng.integers(0, 200, 2)
            pi[a], pi[b] = pi[b], pi[a]
            new = cost(pi, S, rows)
            if new <= cur or rng.random() < np.exp((cur - new) / T):
                cur = new
            else:
                pi[a], pi[b] = pi[b], pi[a]
        else:
            m = rng.integers(0, 6)
            a, b = rng.integers(0, 64, 2)
            r0, r1 = STARTS[m], STARTS[m + 1]
            S[r0:r1, [a, b]] = S[r0:r1, [b, a]]
            rows[m, [a, b]] = rows[m, [b, a]]
            new = cost(pi, S, rows)
            if new <= cur or rng.random() < np.exp((cur - new) / T):
                cur = new
            else:
                S[r0:r1, [a, b]] = S[r0:r1, [b, a]]
                rows[m, [a, b]] = rows[m, [b, a]]
        T = max(0.05, T * 0.99993)
    out = os.path.join(HERE, "kqed_bank_layout.h")
    with open(out, "w") as f:
        f.write("/* Generated by gen_bank_layout.py: do not edit.  LDS layout of the standard assembly map\n"
                " * (see the generator for the method).  Modelled LDS cycles of the map's gathers and output\n"
                f" * writes per two-pair evaluation: {start} with the natural layout, {cur} with this one. */\n")
        f.write("#ifndef KQED_BANK_LAYOUT_H\n#define KQED_BANK_LAYOUT_H\n")
        f.write("/* LDS slot of tensor entry c (c < 200; the zero entry keeps slot 200) */\n")
        f.write("#define KQED_STD_SLOT { " + ", ".join(str(int(v)) for v in pi[:200]) + " }\n")
        f.write("/* output row held by lane l of row slot m of the standard map: [6][64] */\n")
        f.write("#define KQED_STD_ROWS { " + ", ".join(str(int(v)) for v in rows.reshape(-1)) + " }\n")
        f.write("#endif\n")
    print(f"{out}: modelled cycles {start} -> {cur}")


if __name__ == "__main__":
    sys.exit(main())
