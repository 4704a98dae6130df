"""Benchmark: A x + A^T y pairs per second on one MI355X.

Workload (default): the full-brain drop-in configuration — 145x174x145 grid at
1.25 mm, ~600k-voxel white-matter mask, 1M synthetic fibres of 20-200 mm
(~8.5e7 segments), nS=288 (18 b0 + 3 shells x 90), 500 directions,
Stick-Zeppelin-Ball (1 IC + 1 EC + 2 ISO response functions), FP64 x and y.

A step is one pair: y = A x followed by x' = A^T y, with x, y and the
dictionary already resident in HBM.  W untimed warm-up pairs, then K timed
pairs bracketed by torch.cuda.synchronize().  Each entry point is also timed
alone with HIP events on the stream it is launched on, to report achieved
algorithmic bandwidth against the 8 TB/s HBM peak.  The CPU baseline is the
oracle's C restatement of the reference operator (same algorithm, threads and
compiler flags as the reference's generated C) timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(D, K, bx=8, by=8):
    """Compulsory bytes of one product (A x or A^T y):
    10 B per IC segment (u32 id + u16 orientation + f32 length), 6 B per EC
    peak, 4 B per ISO voxel, x and y once, and the LUTs once."""
    nIC, nEC, nISO = K["wmr"].shape[0], K["wmh"].shape[0], K["iso"].shape[0]
    nS = K["wmr"].shape[2]
    ndirs = K["wmr"].shape[1]
    nnz, nE, nI = D["IC"]["n"], D["EC"]["n"], D["ISO"]["n"]
    ncols = D["IC"]["nSTR"] * nIC + nE * nEC + nI * nISO
    nrows = D["IC"]["nVOX"] * nS
    return (10 * nnz + 6 * nE + 4 * nI + ncols * bx + nrows * by
            + 4 * nS * (ndirs * (nIC + nEC) + nISO))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="brain")
    ap.add_argument("--fibres", type=int, default=None)
    ap.add_argument("--cpu-pairs", type=int, default=2, help="timed pairs of the CPU baseline (0: skip)")
    ap.add_argument("--rhs", type=int, default=16, help="right-hand sides of the batched leg (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 256 MiB device copies (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the operator runs on a single MI355X; see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    from commit_amd import synth
    from commit_amd.operator.operator import LinearOperator

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    seed = 1002
    t0 = time.time()
    print(f"[bench] generating workload '{args.workload}'", file=sys.stderr, flush=True)
    D, K, raw = synth.make_problem(args.workload, seed=seed, nF=args.fibres)
    print(f"[bench] dictionary ready in {time.time() - t0:.1f}s: nnz={D['IC']['n']} "
          f"nVOX={D['IC']['nVOX']} nF={D['IC']['nSTR']}", file=sys.stderr, flush=True)
    t0 = time.time()
    A = LinearOperator(D, K, None)
    At = A.T
    print(f"[bench] device plan built in {time.time() - t0:.1f}s "
          f"({A._plan.device_bytes / 2**30:.2f} GiB)", file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(seed)
    x = torch.rand(A.shape[1], dtype=torch.float64, device=dev, generator=g)   # dense: no zero-skip
    y = torch.empty(A.shape[0], dtype=torch.float64, device=dev)
    xo = torch.empty_like(x)

    def pair():
        A.matvec(x, out=y)
        At.matvec(y, out=xo)

    for _ in range(args.warmup):
        pair()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pair()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = args.steps / dt
    print(f"[bench] {args.steps} pairs in {dt:.3f}s", file=sys.stderr, flush=True)

    # per-entry-point timing with HIP events on the launch stream
    stream = torch.cuda.current_stream(dev)

    def time_op(fn, reps):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        fn()
        ev0.record(stream)
        for _ in range(reps):
            fn()
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_time(ev1) / reps

    reps = max(5, args.steps)
    import ctypes
    from commit_amd import _lib
    lib = _lib.load()
    h = A._plan.handle
    sp = ctypes.c_void_p(stream.cuda_stream)

    def at_pass(p):
        _lib.check(lib.commit_hip_At_pass(h, ctypes.c_void_p(y.data_ptr()), None, ctypes.c_void_p(xo.data_ptr()),
                                          p, sp), "commit_hip_At_pass")

    ms_A = time_op(lambda: A.matvec(x, out=y), reps)
    ms_At = time_op(lambda: At.matvec(y, out=xo), reps)
    ms_At1 = time_op(lambda: at_pass(1), reps)
    ms_At2 = time_op(lambda: at_pass(2), reps)
    bytes_pass = algorithmic_bytes(D, K)
    kernels = {"k_A_fp64": ms_A, "k_At_fp64": ms_At1}
    dom_name = max(kernels, key=kernels.get)
    dom_ms = kernels[dom_name]
    dom_gbs = bytes_pass / (dom_ms * 1e-3) / 1e9
    print(f"[bench] A x {ms_A:.3f} ms; A^T y {ms_At:.3f} ms (pass 1 {ms_At1:.3f}, pass 2 {ms_At2:.3f}); "
          f"dominant {dom_name}: {dom_gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
    traffic = measured_traffic(dom_name)

    # sparse x (20% nonzero, the non-negative FISTA regime): extra figure
    xs = torch.from_numpy(synth.x_true(A.shape[1], seed + 1)).to(dev)
    ms_A_sparse = time_op(lambda: A.matvec(xs, out=y), reps)

    # FP32 storage mode (float32 x and y, FP64 accumulation) on the same plan
    from commit_amd import _lib as L
    A.set_precision(L.FP32)
    x32, y32, xo32 = x.float(), torch.empty(A.shape[0], dtype=torch.float32, device=dev), xo.float()
    ms_A32 = time_op(lambda: A.matvec(x32, out=y32), reps)
    ms_At32 = time_op(lambda: At.matvec(y32, out=xo32), reps)
    A.set_precision(L.FP64)
    print(f"[bench] fp32 storage: A x {ms_A32:.3f} ms; A^T y {ms_At32:.3f} ms", file=sys.stderr, flush=True)

    batched = None
    if args.rhs > 0:
        batched = batched_leg(A, At, args.rhs, reps, time_op, ms_A + ms_At, seed, dev)

    cpu = None
    if args.cpu_pairs > 0:
        cpu = cpu_baseline(D, K, x.cpu().numpy(), args.cpu_pairs)

    line = {
        "metric": "A·x+Aᵀ·y pairs/s @1M fibers nS=288",
        "value": round(value, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {D['IC']['nSTR']} fibres, {D['IC']['n']} segments, "
                        f"{D['IC']['nVOX']} voxels, nS={K['wmr'].shape[2]}, ndirs={K['wmr'].shape[1]}, "
                        "SZB 1 IC + 1 EC + 2 ISO, dense x~U(0,1)",
            "nnz": D["IC"]["n"], "nVOX": D["IC"]["nVOX"], "nF": D["IC"]["nSTR"],
            "nE": D["EC"]["n"], "nI": D["ISO"]["n"], "nS": int(K["wmr"].shape[2]),
            "ms_A": round(ms_A, 4), "ms_At": round(ms_At, 4), "ms_At_pass1": round(ms_At1, 4),
            "ms_At_pass2": round(ms_At2, 4), "ms_A_sparse_x": round(ms_A_sparse, 4),
            "algorithmic_bytes_per_product": int(bytes_pass),
            "fp32_storage": {"ms_A": round(ms_A32, 4), "ms_At": round(ms_At32, 4),
                             "pairs_per_s": round(1e3 / (ms_A32 + ms_At32), 2),
                             "algorithmic_bytes_per_product": int(algorithmic_bytes(D, K, 4, 4))},
            "batched": batched,
        },
        "roofline": {
            "bound": "hbm", "kernel": dom_name, "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def batched_leg(A, At, R, reps, time_op, ms_plain_pair, seed, dev):
    """R right-hand sides at once (float32 X [ncols, R], Y [nrows, R]) through
    the matrix-core products; per-RHS pair time against the plain pair."""
    import torch
    g = torch.Generator(device=dev).manual_seed(seed + 7)
    X = torch.rand((A.shape[1], R), dtype=torch.float32, device=dev, generator=g)
    Y = torch.empty((A.shape[0], R), dtype=torch.float32, device=dev)
    Xo = torch.empty_like(X)
    ms_Am = time_op(lambda: A.matmat(X, out=Y), reps)
    ms_Atm = time_op(lambda: At.matmat(Y, out=Xo), reps)
    per_rhs = (ms_Am + ms_Atm) / R
    out = {"nrhs": R, "dtype": "f32 storage, f32 matrix-core products, f64 per-fibre sums",
           "ms_A": round(ms_Am, 4), "ms_At": round(ms_Atm, 4), "ms_pair_per_rhs": round(per_rhs, 4),
           "pairs_per_s_per_rhs": round(1e3 / per_rhs, 2),
           "speedup_vs_plain_per_rhs": round(ms_plain_pair / per_rhs, 3)}
    print(f"[bench] batched R={R}: A {ms_Am:.3f} ms, A^T {ms_Atm:.3f} ms -> {per_rhs:.3f} ms per RHS pair "
          f"({out['speedup_vs_plain_per_rhs']:.2f}x the plain pair)", file=sys.stderr, flush=True)
    return out


def copy_calibration():
    """Three device-to-device copies of 256 MiB each: a kernel whose HBM
    traffic is known exactly (256 MiB read + 256 MiB written per launch)."""
    import torch
    n = (256 << 20) // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"[bench] copy calibration: 3 copies of {n * 8} bytes", flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel
    calibration recorded with it), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "traffic_brain1M.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(D, K, x, pairs):
    """The oracle's C operator with the reference's flags (-O3 -Ofast) and thread
    partition, on the host cores: median of `pairs` pairs after one warm-up."""
    from oracle import oracle

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    n = max(1, min(cores, 255))
    t0 = time.time()
    T = oracle.set_threads(D, n)
    print(f"[bench] cpu baseline: partition for {n} threads in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    R = oracle.OracleOperator(D, K, T, fast=True)
    Rt = R.T
    y = R.dot(x)
    Rt.dot(y)
    times = []
    for i in range(pairs):
        t0 = time.perf_counter()
        y = R.dot(x)
        Rt.dot(y)
        times.append(time.perf_counter() - t0)
        print(f"[bench] cpu pair {i}: {times[-1]:.2f}s", file=sys.stderr, flush=True)
    med = float(np.median(times))
    return {"value": round(1.0 / med, 5), "unit": "pairs/s", "cores": n, "kind": "port",
            "sample": f"full workload, 1 warm-up + median of {pairs} pairs (A x then A^T y, "
                      f"zero-filled outputs), oracle C operator built -O3 -Ofast, {n} threads"}


if __name__ == "__main__":
    main()
