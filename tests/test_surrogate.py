"""The learned surrogate: MFMA forward/backward kernels against the CPU reference of the network,
training, the per-bin error check and the table fallback.

Bars: gradients of every layer within 2e-2 normwise of the CPU reference (BF16 operands on both
sides; the difference is FP32 vs FP64 accumulation and rare BF16 rounding flips), loss within 1e-3;
with a tolerance at which some bins pass, the pairs of those bins take the network (within the
tolerance) and every other pair is bit-identical to the table path.
"""
import numpy as np
import pytest

import kqed_amd
from kqed_amd import workloads as W
from pairs import normwise
from surrogate_ref import forward_backward

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_gradients_match_cpu_reference(gpu_small):
    s = kqed_amd.Surrogate(gpu_small, seed=3)
    rng = np.random.default_rng(0)
    B = 512
    X3 = rng.uniform(-1, 1, (B, 3)).astype(np.float32)
    tgt = rng.normal(0, 0.5, (B, 34)).astype(np.float32)
    wgt = (rng.uniform(size=(B, 34)) < 0.9).astype(np.float32)
    loss, g = s.gradients(X3, tgt, wgt)
    P, _, _, _ = s.params()
    layers = kqed_amd.surrogate_layers(P)
    ref_loss, ref = forward_backward(layers, X3.astype(np.float64), tgt.astype(np.float64), wgt.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    gl = kqed_amd.surrogate_layers(g)
    for l in range(5):
        dW, db = ref[l]
        for got, want, name in ((gl[l][0], dW, "W"), (gl[l][1], db, "b")):
            e = np.abs(got - want).max() / np.abs(want).max()
            assert e <= 2e-2, f"layer {l} d{name}: {e:.3e}"
    s.destroy()


@pytest.fixture(scope="module")
def trained(gpu_full):
    x, y = W.pairs_log_uniform(21, 1 << 17)
    xd, yd = _dev(x), _dev(y)
    s = kqed_amd.Surrogate(gpu_full, seed=7)
    s.train(len(x), xd.data_ptr(), yd.data_ptr(), epochs=3, batch=4096, lr=2e-3)
    yield s
    s.destroy()


def test_training_gain(gpu_full, trained):
    """The trained network's L0 error is at least ten times below the untrained network's."""
    xv, yv = W.pairs_log_uniform(22, 1 << 15)
    xvd, yvd = _dev(xv), _dev(yv)
    untrained = kqed_amd.Surrogate(gpu_full, seed=7)
    v0 = untrained.validate(len(xv), xvd.data_ptr(), yvd.data_ptr(), tol=1e-3)
    untrained.destroy()
    v1 = trained.validate(len(xv), xvd.data_ptr(), yvd.data_ptr(), tol=1e-3, min_count=4)
    assert v1["median_err"] <= 0.1 * v0["median_err"], (v0, v1)


def test_substitution_path_with_passing_bins(gpu_full, trained):
    """With a tolerance at which some bins pass (three times the median error), a surrogate batch on
    the validation pairs themselves puts the network on the packed pairs of the passing bins, each
    within the tolerance of the table path, and leaves every other pair bit-identical to it."""
    import torch
    xv, yv = W.pairs_log_uniform(23, 1 << 15)
    xvd, yvd = _dev(xv), _dev(yv)
    v_all = trained.validate(len(xv), xvd.data_ptr(), yvd.data_ptr(), tol=1e30, min_count=4)
    tol = 3.0 * v_all["median_err"]
    v = trained.validate(len(xv), xvd.data_ptr(), yvd.data_ptr(), tol=tol, min_count=4)
    assert 0 < v["packed_in_passing_bins"] < v["packed"], v
    out = torch.empty((len(xv), 384), dtype=torch.float64, device="cuda")
    trained.batch_device(0, len(xv), xvd.data_ptr(), yvd.data_ptr(), out.data_ptr(),
                         stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = gpu_full.batch_host(0, xv, yv)
    same = np.all(got == ref, axis=1)
    err = normwise(got, ref)
    # the bins' errors are held in FP32: allow that rounding
    assert np.all(same | (err <= tol * (1 + 1e-6))), (err[~same].max(), tol)
    assert (~same).sum() == v["packed_in_passing_bins"], ((~same).sum(), v)


def test_save_load_roundtrip(gpu_small, tmp_path):
    a = kqed_amd.Surrogate(gpu_small, seed=11)
    b = kqed_amd.Surrogate(gpu_small, seed=12)
    f = str(tmp_path / "surr.bin")
    a.save(f)
    b.load(f)
    pa, la, ma, ra = a.params()
    pb, lb, mb, rb = b.params()
    assert np.array_equal(pa, pb) and np.array_equal(la, lb) and np.array_equal(ma, mb) and np.array_equal(ra, rb)
