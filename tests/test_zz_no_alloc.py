"""kqed_batch_noalloc: kqed_batch with its workspace in a fixed device array, no allocation.

Called through the C ABI (ctypes) and checked against kqed_batch on the same inputs: outputs do
not depend on the pass length, so they must be bit-identical.  Two calls in a row, the second
smaller, share the scratch array; device free memory must not drop by a workspace's worth; batches
above the maximum are rejected with an error.
"""
import ctypes

import numpy as np
import pytest

import kqed_amd
from kqed_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_PAIRS = 262144


def _noalloc(t, which, M, n, xd, yd, out, layout, stream):
    L = kqed_amd.lib()
    f = L.kqed_batch_noalloc
    f.argtypes = [ctypes.POINTER(kqed_amd._Temps), ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return f(ctypes.byref(t._t), int(which), float(M), int(n), ctypes.c_void_p(xd), ctypes.c_void_p(yd),
             ctypes.c_void_p(out), int(layout), ctypes.c_void_p(stream))


@pytest.mark.parametrize("which,layout", [(0, kqed_amd.KQED_AOS), (2, kqed_amd.KQED_AOS_F32), (4, kqed_amd.KQED_SOA)])
def test_noalloc_equals_batch_two_calls(gpu_full, which, layout):
    import torch
    f32 = layout == kqed_amd.KQED_AOS_F32
    dt = torch.float32 if f32 else torch.float64
    M = 0.5 if which == 4 else 0.0
    cases = [W.pairs_log_uniform(61 + which, 150000), W.pairs_log_uniform(71 + which, 9000)]
    bufs = []
    for x, y in cases:
        xd = torch.from_numpy(x).cuda()
        yd = torch.from_numpy(y).cuda()
        ref = torch.empty((len(x), 384), dtype=dt, device="cuda")
        got = torch.full((len(x), 384), float("nan"), dtype=dt, device="cuda")
        bufs.append((xd, yd, ref, got))
    # the reference results first, on the same stream: the first dispatch of a spilling kernel to a
    # hardware queue makes the HIP runtime reserve private (register-spill) memory for that queue,
    # which is not an allocation of the library and must not count against the entry below
    s_na = torch.cuda.Stream()
    for xd, yd, ref, _ in bufs:
        gpu_full.batch_device(which, len(xd), xd.data_ptr(), yd.data_ptr(), ref.data_ptr(), M=M, layout=layout,
                              stream=s_na.cuda_stream)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for xd, yd, _, got in bufs:
        rc = _noalloc(gpu_full, which, M, len(xd), xd.data_ptr(), yd.data_ptr(), got.data_ptr(), layout,
                      s_na.cuda_stream)
        assert rc == 0, kqed_amd.lib().kqed_last_error()
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    # a workspace for 150000 pairs would take 150000 * 1152 B = 173 MB of records alone
    assert free0 - free1 < 16 << 20, (free0, free1)
    for _, _, ref, got in bufs:
        assert torch.equal(got, ref)


def test_noalloc_rejects_too_many_pairs(gpu_full):
    import torch
    x = torch.zeros((4, 4), dtype=torch.float64, device="cuda")
    rc = _noalloc(gpu_full, 0, 0.0, MAX_PAIRS + 1, x.data_ptr(), x.data_ptr(), x.data_ptr(), kqed_amd.KQED_AOS,
                  torch.cuda.current_stream().cuda_stream)
    assert rc < 0
    assert b"maximum" in kqed_amd.lib().kqed_last_error()
