#!/usr/bin/env python3
"""Throughput of the FP64 QED kernel L(x,y) on one MI355X, in evaluations per second.

Workload (the plain-FP64 GPU configuration): 2^24 (x,y) pairs, |x| and |y| log-uniform in
[1e-3, 4.0], isotropic directions (seed 2), FP64 in and out, synthetic form-factor table "full"
(reference file format; the real table is not available).  One step = one batched evaluation of
QED_kernel_L0 over all pairs (inputs resident in HBM, 384 FP64 outputs per pair written to HBM).
L1, L2, L3 and M-L2 (M = 0.5) rates are measured on the same inputs and reported beside it.

Prints one JSON line.  --gpus N > 1 is refused: the library runs on a single GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PEAK_FP64 = 78.6e12  # MI355X FP64 vector / matrix, FLOP/s
PEAK_HBM = 8.0e12


def workload(n, seed=2):
    from pairs import log_uniform
    return log_uniform(seed, n)


def cpu_baseline(table, x2, y2, reps=9):
    """The CPU oracle (restatement of the reference's AVX+FMA FP64 path, built -O3 -mavx2 -mfma
    -ffp-contract=off) on the CPU L0 benchmark law: 4096 pairs, |x|, |y| uniform in [0.01, 3.0],
    isotropic, seed 1 (the reference's timing loop is amu_forlattice.c:1402-1433).  Median of
    `reps` runs after one warm-up, on one thread and on the host cores this process may use (at most
    16, the box's CPU share per GPU).  Also counts the algorithmic flops of one L0 on 2000 pairs of
    the GPU workload (x2, y2) for the roofline."""
    from oracle_ctypes import Oracle
    from kqed_amd import workloads as W
    o = Oracle(table)
    x, y = W.pairs_uniform(1, 4096)
    cores = min(16, len(os.sched_getaffinity(0)))

    def rate(nt):
        o.batch(0, x, y, nthreads=nt)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            o.batch(0, x, y, nthreads=nt)
            ts.append(time.perf_counter() - t0)
        return len(x) / float(np.median(ts)), float(np.sum(ts))

    single, t_single = rate(1)
    multi, t_multi = rate(cores)
    o.flops_reset()
    o.batch(0, x2[:2000], y2[:2000], nthreads=1)
    flops = {"total": o.flops() / 2000, "node_only": o.flops_build() / 2000}
    o.close()
    return {"value": multi, "unit": "evals/s", "cores": cores, "kind": "port",
            "single_thread": single, "host_nproc": os.cpu_count(),
            "sample": f"4096 pairs, |x|,|y| U[0.01,3.0], seed 1 (the CPU L0 law), L0, synthetic table "
                      f"'{os.path.basename(table)}'; median of {reps} runs after a warm-up; "
                      f"{t_single + t_multi * cores:.1f} CPU-seconds"}, flops


def parity_sample(table, which, M, x, y, got, nthreads):
    """max and median per-pair normwise error of GPU rows `got` against the oracle on the same pairs."""
    from oracle_ctypes import Oracle
    from pairs import normwise
    o = Oracle(table)
    ref = o.batch(which, x, y, M=M, nthreads=nthreads)
    o.close()
    err = normwise(got, ref)
    return {"pairs": len(x), "max_rel_err": float(err.max()), "median_rel_err": float(np.median(err))}


def pass_traffic(pairs_per_pass):
    """HBM bytes of one pipeline pass from the committed counter profile (FETCH_SIZE x its copy-kernel
    calibration factor + WRITE_SIZE, summed over the pass's kernels), scaled to this pass length."""
    f = os.path.join(ROOT, "profiles", "traffic_l0_pass.txt")
    try:
        last = [ln for ln in open(f).read().splitlines() if ln.strip()][-1]  # the summary's final JSON line
        d = json.loads(last)
        return d["bytes_per_pass"] * pairs_per_pass / d["pairs_per_pass"]
    except (OSError, KeyError, ValueError):
        return None


def contraction_leg(t, dev, stream, L):
    """sum_x L2(x,y) . T(x) over an L^4 lattice for the 16 fixed y sites (FP32 T, FP64 accumulation)."""
    import torch
    from kqed_amd import workloads as W
    x = W.lattice_sites(L)
    T = W.site_tensors(x, seed=3)
    y = W.y_sites()
    xd = torch.from_numpy(x).to(dev)
    Td = torch.from_numpy(T).to(dev)
    del T
    t.contract_device(2, len(x), xd.data_ptr(), Td.data_ptr(), y, stream=stream)  # warm-up: the timed call's workspace
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s = t.contract_device(2, len(x), xd.data_ptr(), Td.data_ptr(), y, stream=stream)
    dt = time.perf_counter() - t0
    del xd, Td
    return {"workload": f"sum_x L2(x,y).T(x), {L}^4 lattice, a=0.0536, 16 y sites, FP32 T, FP64 accumulation",
            "l2_evals": len(x) * len(y), "seconds": dt, "evals_per_s": len(x) * len(y) / dt,
            "s_first": float(s[0])}


def lattice_comb_leg(t, dev, stream, L, ny=4):
    """Per-y contraction of a lattice through the x<->y symmetrised kernels (SYMXY.c, all four L^k):
    s_k(y) = sum_x L^k_sym(x,y) . T(x), L^4 lattice, the first ny of the 16 fixed y sites."""
    import torch
    import kqed_amd
    from kqed_amd import workloads as W
    x = W.lattice_sites(L)
    T = W.site_tensors(x, seed=3)
    y = W.y_sites()[:ny]
    xd, Td = torch.from_numpy(x).to(dev), torch.from_numpy(T).to(dev)
    del T
    comb = kqed_amd.KQED_COMB_SYMXY_V2
    t.contract_comb_device(comb, len(x), xd.data_ptr(), Td.data_ptr(), y, stream=stream)  # warm-up: the timed call's workspace
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s = t.contract_comb_device(comb, len(x), xd.data_ptr(), Td.data_ptr(), y, stream=stream)
    dt = time.perf_counter() - t0
    del xd, Td
    return {"workload": f"sum_x L^k_SYMXY(x,y).T(x), k = 0..3, {L}^4 lattice, {ny} y sites, FP32 T, FP64 sums",
            "pairs": len(x) * ny, "seconds": dt, "pairs_per_s": len(x) * ny / dt,
            "l0_evals_per_s": 6 * len(x) * ny / dt, "s_first": [float(v) for v in s[0]]}


def combination_leg(t, dev, stream, xd, yd, n):
    """The reference's kernel combinations (all_kernels.c, con_kernel.c, SYMXY0.c) on the first n
    pairs of the workload: combinations/s and the L0 evaluations/s inside them."""
    import torch
    import kqed_amd
    M = np.array([0.0, 0.4, 0.8, 1.2])
    legs = (("compute_all_kernels", kqed_amd.KQED_COMB_ALL, 6, None),
            ("compute_con_kernelsM_L2", kqed_amd.KQED_COMB_CON_M_L2, 8, M),
            ("compute_all_kernels_SYMXY0_v2", kqed_amd.KQED_COMB_SYMXY0_V2, 12, None))
    out = torch.empty((n, 12 * 384), dtype=torch.float64, device=dev)
    res = {"workload": f"first {n} pairs of the L0 workload, FP64, M = (0, 0.4, 0.8, 1.2) for the M variant"}
    for name, comb, evals, m in legs:
        t.combine_device(comb, n, xd.data_ptr(), yd.data_ptr(), out.data_ptr(), M=m, stream=stream)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t.combine_device(comb, n, xd.data_ptr(), yd.data_ptr(), out.data_ptr(), M=m, stream=stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = {"pairs_per_s": n / dt, "l0_evals_per_s": evals * n / dt, "ms": dt * 1e3}
    del out
    return res


# FP64 flops of iPihat per pair as the GPU executes them: per pihat1 evaluation and output thread
# (256), 76 additions forming the 25 partial sums of the nonzero trace terms, 121 fused multiply-adds
# and 4 additions contracting them with the propagator tensors, 5 scalings; three evaluations, and
# 14 per output of the antisymmetrised sum.  The reference's loop order (pi_pert.c:320-377) takes
# 3 * 256 * (16 * (16 * 11 + 47) + 12) + 384 * 14 = 2.75 M per pair for the same result.
PIHAT_FLOPS = 3 * 256 * (76 + 2 * 121 + 4 + 5) + 384 * 14


def lepton_loop_leg(t, dev, stream, n):
    """iPihat of the free-lepton loop (pi_pert.c) and the example1 integrand sum_c L0 . iPihat on n
    pairs with |x|, |y| uniform in [0.05, 1], isotropic, Mv = 2."""
    import torch
    from pairs import iso
    rng = np.random.default_rng(5)
    x = iso(rng, n) * rng.uniform(0.05, 1.0, (n, 1))
    y = iso(rng, n) * rng.uniform(0.05, 1.0, (n, 1))
    xd, yd = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    P = torch.empty((n, 384), dtype=torch.float64, device=dev)
    o = torch.empty(n, dtype=torch.float64, device=dev)
    t.pihat_device(1024, xd.data_ptr(), yd.data_ptr(), P.data_ptr(), stream=stream)  # warm-up (tables, workspace)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t.pihat_device(n, xd.data_ptr(), yd.data_ptr(), P.data_ptr(), stream=stream)
    torch.cuda.synchronize()
    dt_p = time.perf_counter() - t0
    t0 = time.perf_counter()
    t.lepton_loop_device(0, n, xd.data_ptr(), yd.data_ptr(), o.data_ptr(), Mv=2.0, stream=stream)
    torch.cuda.synchronize()
    dt_l = time.perf_counter() - t0
    del xd, yd, P, o
    return {"workload": f"{n} pairs, |x|,|y| U[0.05,1], isotropic, FP64; integrand with L0 and Mv = 2",
            "pihat_pairs_per_s": n / dt_p, "pihat_executed_tflops": PIHAT_FLOPS * n / dt_p / 1e12,
            "pihat_executed_flops_per_pair": PIHAT_FLOPS, "integrand_pairs_per_s": n / dt_l}


def ragged_leg(t, dev, stream, reps=3):
    """256 y points with lognormal x-sets (median 16384, sigma 1, clipped [1024, 262144]), L0, FP32
    output (FP32 assembly maps, FP64 interpolation and Taylor branch), autotuned, one graph launch."""
    import torch
    import kqed_amd
    from kqed_amd import workloads as W
    y, sizes, x = W.ragged(seed=4)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uintp)
    n = int(off[-1])
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y).to(dev)
    out = torch.empty((n, 384), dtype=torch.float32, device=dev)
    # autotune on the first 2^20 pairs of the concatenation (their y vectors expanded)
    m = min(n, 1 << 20)
    ys = torch.from_numpy(np.repeat(y, sizes, axis=0)[:m]).to(dev)
    t0 = t.get_tune()
    tune, tune_ms = t.autotune(0, m, xd.data_ptr(), ys.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32,
                               reps=1, stream=stream)
    del ys
    g = t.graph_ragged(0, off, yd.data_ptr(), xd.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32)
    g.launch(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        g.launch(stream)
    torch.cuda.synchronize()
    dt_graph = (time.perf_counter() - t1) / reps
    t1 = time.perf_counter()
    t.batch_ragged(0, off, yd.data_ptr(), xd.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32, stream=stream)
    torch.cuda.synchronize()
    dt_eager = time.perf_counter() - t1
    g.destroy()
    t.set_tune(t0)
    del out, xd, yd
    return {"workload": "256 y (|y| U[0.05,3]), lognormal x-sets (median 16384, sigma 1.0, clipped [1024,262144]), "
                        "|x| log-uniform [1e-3,4], L0, FP32 assembly/output, one graph launch",
            "groups": len(sizes), "pairs": n, "seconds_per_launch": dt_graph, "evals_per_s": n / dt_graph,
            "evals_per_s_eager": n / dt_eager, "tune": tune.as_dict(), "tune_ms_per_2^20": tune_ms}


def surrogate_leg(t, dev, stream, n_train, epochs, n_infer, n_plain_rate):
    """Surrogate + FP32 form: train the network on interpolator samples of n_train pairs, validate at
    1e-5 (per-bin error check), then evaluate n_infer pairs of the benchmark law with FP32 assembly
    (network in passing bins, tables elsewhere); accuracy against the plain FP64 path on a sample."""
    import torch
    import kqed_amd
    from kqed_amd import workloads as W
    from pairs import normwise
    x, y = W.pairs_log_uniform(31, n_train)
    xd, yd = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    s = kqed_amd.Surrogate(t, seed=7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = s.train(n_train, xd.data_ptr(), yd.data_ptr(), epochs=epochs, batch=8192, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del xd, yd
    xv, yv = W.pairs_log_uniform(32, 1 << 20)
    xvd, yvd = torch.from_numpy(xv).to(dev), torch.from_numpy(yv).to(dev)
    chunk = min(n_infer, 1 << 24)
    xi, yi = W.pairs_log_uniform(2, chunk)
    xid, yid = torch.from_numpy(xi).to(dev), torch.from_numpy(yi).to(dev)
    out = torch.empty((chunk, 384), dtype=torch.float32, device=dev)
    m = 1 << 16
    ref = np.empty((m, 384))
    ref[:] = t.batch_host(0, xi[:m], yi[:m])

    def timed(reps):
        s.batch_device(0, chunk, xid.data_ptr(), yid.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32,
                       stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            s.batch_device(0, chunk, xid.data_ptr(), yid.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32,
                           stream=stream)
        torch.cuda.synchronize()
        return reps * chunk / (time.perf_counter() - t0)

    # the network on every packed pair (every bin passes an unbounded tolerance): what it costs and
    # how accurate it is when nothing falls back to the tables
    v_all = s.validate(len(xv), xvd.data_ptr(), yvd.data_ptr(), tol=1e30, min_count=1)
    rate_all = timed(2)
    err_all = normwise(out[:m].double().cpu().numpy(), ref)
    everywhere = {"packed_fraction": v_all["packed"] / v_all["pairs"], "evals_per_s": rate_all,
                  "speedup_vs_plain_fp64": rate_all / n_plain_rate, "max_rel_err_vs_fp64": float(err_all.max()),
                  "median_rel_err_vs_fp64": float(np.median(err_all))}
    val = {f"{tol:g}": s.validate(len(xv), xvd.data_ptr(), yvd.data_ptr(), tol=tol, min_count=8)
           for tol in (1e-2, 1e-3, 1e-5)}  # the last one installs the 1e-5 mask
    del xvd, yvd
    s.batch_device(0, chunk, xid.data_ptr(), yid.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32,
                   stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < n_infer:
        s.batch_device(0, chunk, xid.data_ptr(), yid.data_ptr(), out.data_ptr(), layout=kqed_amd.KQED_AOS_F32,
                       stream=stream)
        done += chunk
    torch.cuda.synchronize()
    rate = done / (time.perf_counter() - t0)
    err = normwise(out[:m].double().cpu().numpy(), ref)
    s.destroy()
    del out, xid, yid
    return {"workload": f"train on {n_train} interpolator samples x {epochs} epochs; {n_infer} pairs of the "
                        "benchmark law, L0, FP32 assembly/output, network in bins passing a 1e-5 check",
            "train_seconds": t_train, "final_loss": loss, "validation": val, "network_everywhere": everywhere,
            "evals_per_s": rate,
            "speedup_vs_plain_fp64": rate / n_plain_rate,
            "max_rel_err_vs_fp64": float(err.max()), "median_rel_err_vs_fp64": float(np.median(err))}


KNAMES = ("L0", "L1", "L2", "L3", "ML2_M0.5")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--table", default="full")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the L1/L2/L3/M-L2 rates and the contraction")
    ap.add_argument("--kernel", type=int, default=0, choices=range(5),
                    help="kernel of the timed workload: 0 L0 (the headline), 1 L1, 2 L2, 3 L3, 4 M-L2 (M = 0.5); "
                         "kernels other than L0 time the GPU only (no CPU baseline or roofline)")
    ap.add_argument("--legs", default="all",
                    help="comma list of extra legs to run (other_kernels, contraction, ragged, combinations, "
                         "lepton_loop, lattice_symxy, surrogate) or 'all'")
    ap.add_argument("--lattice", type=int, default=32, help="lattice extent of the contraction leg")
    ap.add_argument("--surrogate", type=int, default=1, help="run the surrogate + FP32 leg (0 = skip)")
    ap.add_argument("--surrogate-train", type=int, default=1 << 24, help="training pairs of the surrogate leg")
    ap.add_argument("--surrogate-epochs", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (reference kernel for HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: the library runs on one GPU only (asked for --gpus {a.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    import kqed_amd
    table = kqed_amd.table_dir(a.table)
    t = kqed_amd.QEDKernelTemps.initialise(table)
    x, y = workload(a.n)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y).to(dev)
    out = torch.empty((a.n, 384), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def step(which=a.kernel, M=0.5 if a.kernel == 4 else 0.0, n=a.n):
        t.batch_device(which, n, xd.data_ptr(), yd.data_ptr(), out.data_ptr(), M=M, stream=stream)

    if a.kernel != 0:
        a.no_cpu = True

    if a.copy_calibration:
        src = torch.empty(a.copy_calibration << 18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t.profile(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stages = t.profile(False)
    evals = a.n * a.steps / dt


    # uniform samples of the measured outputs (copied out before the buffer is reused), checked
    # against the oracle after the GPU legs: 2^18 pairs of the timed kernel, 2^16 of the others
    rng = np.random.default_rng(17)
    samples = {}

    def take_sample(which, M, m):
        idx = np.sort(rng.choice(a.n, min(m, a.n), replace=False))
        rows = out[torch.from_numpy(idx).to(dev)].cpu().numpy()
        samples[KNAMES[which]] = (which, M, idx, rows)

    extra = {}
    legs = set(a.legs.split(",")) if a.legs != "all" else {"other_kernels", "contraction", "ragged", "combinations",
                                                            "lepton_loop", "lattice_symxy", "surrogate"}
    if a.no_extra:
        legs = set()
    take_sample(a.kernel, 0.5 if a.kernel == 4 else 0.0, 1 << 18)
    # achievable HBM write rate on this GPU, measured live: torch's fill of 2 GiB of the output buffer
    # (after the parity sample of the timed outputs is taken: the fill overwrites them)
    fill_gbps = None
    if a.n * 384 >= 1 << 28:
        buf = out.view(-1)[: 1 << 28]
        buf.fill_(0.0)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(5):
            buf.fill_(1.0)
        ev1.record()
        torch.cuda.synchronize()
        fill_gbps = 5 * 8 * (1 << 28) / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    if "other_kernels" in legs:
        rates = {}
        for name, which, M in (("L1", 1, 0.0), ("L2", 2, 0.0), ("L3", 3, 0.0), ("ML2_M0.5", 4, 0.5)):
            if which == a.kernel:
                continue
            step(which, M)  # warm-up
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(3):
                step(which, M)
            torch.cuda.synchronize()
            rates[name] = 3 * a.n / (time.perf_counter() - t1)
            take_sample(which, M, 1 << 16)
        extra["evals_per_s_other_kernels"] = rates

    if "contraction" in legs:
        extra["contraction"] = contraction_leg(t, dev, stream, a.lattice)
    if "ragged" in legs:
        extra["ragged"] = ragged_leg(t, dev, stream)
    if "combinations" in legs:
        extra["combinations"] = combination_leg(t, dev, stream, xd, yd, min(a.n, 1 << 19))
    if "lepton_loop" in legs:
        extra["lepton_loop"] = lepton_loop_leg(t, dev, stream, 1 << 18)
    if "lattice_symxy" in legs:
        extra["lattice_symxy"] = lattice_comb_leg(t, dev, stream, a.lattice)
    if "surrogate" in legs and a.surrogate:
        del out
        extra["surrogate"] = surrogate_leg(t, dev, stream, a.surrogate_train, a.surrogate_epochs, 1 << 26, evals)

    cpu = None
    flops_per_eval = None
    if not a.no_cpu:
        cpu, flops_per_eval = cpu_baseline(table, x, y)
    nthr = min(16, len(os.sched_getaffinity(0)))
    extra["parity_sample"] = {
        name: dict(parity_sample(table, which, M, x[idx], y[idx], rows, nthr), law="config-2 workload, uniform sample")
        for name, (which, M, idx, rows) in samples.items()}

    ms_i, cnt_i = stages["interp"]
    ms_a, cnt_a = stages["assemble"]
    ms_s, cnt_s = stages["sort"]
    dom = "interp" if ms_i >= ms_a else "assemble"
    roof = None
    if flops_per_eval:
        # Algorithmic FP64 flops of one L0 evaluation (the oracle's count of the reference algorithm
        # without the node-only coefficient-array construction, which the tables hold precomputed)
        # times the pairs of one pipeline pass, over the measured duration of that pass's kernels
        # (interpolation classes + chain-rule algebra + assembly; HIP events on the launch stream).
        per_eval = flops_per_eval["total"] - flops_per_eval["node_only"]
        pairs_per_pass = a.n * a.steps / max(cnt_i, 1)
        avg = (ms_i + ms_a) / max(cnt_i, 1) / 1e3
        achieved = per_eval * pairs_per_pass / avg
        roof = {"bound": "valu", "resource": "FP64 vector ALU: no matrix instructions on this path; the MI355X "
                "FP64 vector peak (78.6 TFLOP/s) equals its FP64 matrix peak",
                "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": pass_traffic(pairs_per_pass),
                "traffic_source": "profiles/traffic_l0_pass.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of the "
                                  "same pass; not measured in this run)",
                "flop_method": "the oracle's count of the reference algorithm's FP64 operations (add, mul, div, "
                               "sqrt 1 each, FMA 2), tallied per operation group on 2000 pairs of this workload",
                "kernel": "one pipeline pass: k_interp<0,1,2> (interpolation and chain-rule algebra) + k_assemble2 / k_assemble",
                "pairs_per_pass": pairs_per_pass, "pass_ms": avg * 1e3,
                "flops_per_eval": per_eval, "flops_per_eval_with_node_arrays": flops_per_eval["total"]}
    roof_asm = None
    if a.kernel == 0 and cnt_a:
        # the assembly stage (k_assemble2 + k_assemble of one pass) against HBM: it writes the 384
        # FP64 outputs of every pair (3072 B) and reads each pair's 640-byte record
        pairs_per_pass = a.n * a.steps / max(cnt_a, 1)
        t_pass = ms_a / cnt_a / 1e3
        alg = (64 + 3072) * pairs_per_pass / t_pass
        rec = (640 + 3072) * pairs_per_pass / t_pass
        roof_asm = {"bound": "hbm", "kernel": "assembly stage of one pass: k_assemble2 (two-pair) + k_assemble (special points)",
                    "achieved": alg / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": alg / PEAK_HBM,
                    "bytes_per_pair": 3136, "bytes_per_pair_note": "algorithmic: x, y in (64 B) + 384 FP64 outputs (3072 B)",
                    "with_record_reads_gbps": rec / 1e9, "measured_fill_gbps": fill_gbps,
                    "with_record_reads_frac_of_fill": (rec / 1e9 / fill_gbps) if fill_gbps else None,
                    "pass_ms": t_pass * 1e3, "pairs_per_pass": pairs_per_pass}
    res = {
        "metric": "L(x,y) evals/s, FP64 plain vs surrogate+FP32 at <=1e-5 max rel err",
        "value": evals, "unit": "evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"QED_kernel_{KNAMES[a.kernel]}, 2^24 pairs, |x|,|y| log-uniform [1e-3,4], isotropic, FP64",
                   "pairs": a.n, "table": a.table, "kernel": KNAMES[a.kernel], "path": "plain FP64",
                   "taylor_fraction": float(np.mean(np.linalg.norm(x, axis=1) < t.XX[0])),
                   "out_of_grid_fraction": float(np.mean(np.linalg.norm(x, axis=1) > t.XX[-1]))},
        "stages_ms_per_step": {"sort": ms_s / a.steps, "interp": ms_i / a.steps, "assemble": ms_a / a.steps},
        "dominant_kernel": dom,
        "roofline": roof,
        "roofline_assembly": roof_asm,
        "cpu_baseline": cpu,
    }
    res.update(extra)
    t.free()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
