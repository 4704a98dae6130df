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
FP64_PEAK_TFLOPS = 78.6
FP64_LANE_OPS = 256 * 64 * 2.4e9        # FP64 VALU lane-operations per second (39.3e12)


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


def distinct_pairs(D, K):
    """Distinct (voxel, LUT row) pairs of one product: IC segments of one voxel
    and orientation share their row product (per IC RF), plus every EC peak
    and RF and every ISO voxel and RF."""
    nIC, nEC, nISO = K["wmr"].shape[0], K["wmh"].shape[0], K["iso"].shape[0]
    ndirs = int(K["wmr"].shape[1])
    seen = np.zeros(int(D["IC"]["nVOX"]) * ndirs, dtype=bool)
    seen[D["IC"]["vox"].astype(np.int64) * ndirs + D["IC"]["dir"].astype(np.int64)] = True
    return int(seen.sum()) * nIC + D["EC"]["n"] * nEC + D["ISO"]["n"] * nISO


def algorithmic_flops(D, K, pairs):
    """Multiply-adds of one product (A x or A^T y), 2 flops each: nS per
    distinct (voxel, LUT row) pair."""
    return 2 * int(K["wmr"].shape[2]) * pairs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="brain")
    ap.add_argument("--fibres", type=int, default=None)
    ap.add_argument("--cpu-pairs", type=int, default=5, help="timed pairs of the CPU baseline (0: skip)")
    ap.add_argument("--rhs", type=int, default=16, help="right-hand sides of the batched leg (0: skip)")
    ap.add_argument("--fit-iters", type=int, default=500,
                    help="FISTA iterations of the mixed-precision fit leg (0: skip); batched-fit mode: max_iter")
    ap.add_argument("--mode", default="headline",
                    choices=("headline", "batched-fit", "autograd", "products", "voxelize"),
                    help="headline: the operator pairs/s line; batched-fit: 8 tractograms x 16 lambdas; "
                         "autograd: differentiable operator on 200k fibres; products: the single-vector "
                         "products alone with a fixed launch option (--slab), for counter and trace runs; "
                         "voxelize: the tractogram voxeliser on 1M streamlines")
    ap.add_argument("--slab", type=int, default=2, help="products mode: slab option of the plan (0, 1, 2)")
    ap.add_argument("--residual", default="combine", choices=("combine", "product"),
                    help="batched-fit mode: momentum residual from r(x) and r(x_prev) inside A^T, or a second A X")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 256 MiB device copies (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the operator runs on a single MI355X; see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.mode == "batched-fit":
        batched_fit_mode(args)
        return
    if args.mode == "autograd":
        autograd_mode(args)
        return
    if args.mode == "products":
        products_mode(args)
        return
    if args.mode == "voxelize":
        voxelize_mode(args)
        return

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
    # launch options of the plan chosen by timing the variants on this dictionary
    from commit_amd import autotune
    tuned = autotune.tune(A, nrhs=max(args.rhs, 16), reps=3)
    print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
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
    nj = (int(K["wmr"].shape[2]) + 63) // 64
    slab = tuned["slab"]["choice"]
    wpb = 8 if slab == 1 else 16
    name_A = ("k_A_dense<double>" if tuned["a_dense"]["choice"] else
              f"k_A_slab<{nj}, {_lib.SLAB_VA}, {wpb}, double>" if slab else f"k_A<{nj}, double>")
    name_At = f"k_At_slab<{nj}, {_lib.SLAB_VAT}, {wpb}, double>" if slab else f"k_At<{nj}, double>"
    kernels = {name_A: ms_A, name_At: ms_At1}
    dom_name = max(kernels, key=kernels.get)
    dom_ms = kernels[dom_name]
    dom_gbs = bytes_pass / (dom_ms * 1e-3) / 1e9
    pairs = distinct_pairs(D, K)
    flops_pass = algorithmic_flops(D, K, pairs)
    dom_tflops = flops_pass / (dom_ms * 1e-3) / 1e12
    # the compute roof of the FP64 kernels: a v_cvt_f64_f32 and a v_fma_f64 per
    # sample and distinct pair, at 256 CUs x 64 lanes x 2.4 GHz FP64 lane-ops/s
    lane_ops = 2 * int(K["wmr"].shape[2]) * pairs
    dom_lops = lane_ops / (dom_ms * 1e-3)
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
        batched["speedup_vs_plain_fp32_per_rhs"] = round((ms_A32 + ms_At32) / batched["ms_pair_per_rhs"], 3)

    fit = None
    if args.fit_iters > 0:
        fit = mixed_precision_leg(A, seed, args.fit_iters)

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
            "autotune": tuned,
            "mixed_precision_fit": fit,
        },
        "roofline": {
            "bound": "hbm", "kernel": dom_name, "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            # the same launch against the FP64 arithmetic peak (vector and matrix FP64 peaks
            # coincide on MI355X): multiply-adds of one product, 2 flops each
            "fp64": {"achieved": round(dom_tflops, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(dom_tflops / FP64_PEAK_TFLOPS, 4), "flops": int(flops_pass),
                     "distinct_pairs": pairs},
            "compute": {"achieved": round(dom_lops / 1e12, 3), "peak": FP64_LANE_OPS / 1e12,
                        "unit": "T FP64 lane-op/s", "frac": round(dom_lops / FP64_LANE_OPS, 4),
                        "lane_ops": int(lane_ops),
                        "model": "one v_cvt_f64_f32 and one v_fma_f64 per sample and distinct (voxel, LUT row)"},
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
           "speedup_vs_plain_per_rhs": round(ms_plain_pair / per_rhs, 3)}     # against the FP64 plain pair
    print(f"[bench] batched R={R}: A {ms_Am:.3f} ms, A^T {ms_Atm:.3f} ms -> {per_rhs:.3f} ms per RHS pair "
          f"({out['speedup_vs_plain_per_rhs']:.2f}x the plain pair)", file=sys.stderr, flush=True)
    return out


def mixed_precision_leg(A, seed, iters, ngroups=1000):
    """The same group-lasso + non-negativity fit (1,000 random bundles) run with
    float32 vectors and with float64 vectors on the same plan; reports the
    float32 fit's x and objective error against the float64 one."""
    import torch
    from commit_amd import _lib, solvers, synth
    dev = A.device
    n, m = A.shape[1], A.shape[0]
    P = A._plan
    nF, nIC, nEC, nISO = P.nF, P.nIC, P.nEC, P.nISO
    xt = torch.from_numpy(synth.x_true(n, seed + 11)).to(dev)
    s = A.matvec(xt)
    sigma = float(s.mean()) / 30.0
    g = torch.Generator(device=dev).manual_seed(seed + 12)
    y = torch.sqrt((s + sigma * torch.randn(m, generator=g, dtype=torch.float64, device=dev)) ** 2 +
                   (sigma * torch.randn(m, generator=g, dtype=torch.float64, device=dev)) ** 2)
    del s
    rng = np.random.default_rng(seed + 13)
    lab = rng.integers(0, ngroups, nF)
    order = np.argsort(lab, kind="stable")
    cuts = np.cumsum(np.bincount(lab, minlength=ngroups))[:-1]
    groups = np.empty(ngroups, dtype=object)
    for k, gidx in enumerate(np.split(order, cuts)):
        groups[k] = gidx.astype(np.int32)
    gw = np.sqrt([gi.size for gi in groups]).astype(np.float64)
    aty = A.T.matvec(y).cpu().numpy()
    lam_max = max(np.linalg.norm(aty[gi]) / w for gi, w in zip(groups, gw) if gi.size)   # (core.pyx:1196-1211)
    rp = {"startIC": 0, "sizeIC": nF * nIC, "startEC": nF * nIC, "sizeEC": P.nE * nEC,
          "startISO": nF * nIC + P.nE * nEC, "sizeISO": P.nV * nISO, "regIC": "group_lasso", "regEC": None,
          "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True, "lambdaIC": 0.05 * lam_max,
          "dictIC_params": {"group_idx_kept": groups, "group_weights": gw}}
    out = {}
    for name, prec in (("fp32", _lib.FP32), ("fp64", _lib.FP64)):
        A.set_precision(prec)
        reg = solvers.init_regularisation(dict(rp), n=n, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, d = solvers.solve(y, A, A.T, tol_fun=0.0, tol_x=0.0, max_iter=iters, verbose=0, regularisation=reg,
                             lipschitz_iters=30)
        torch.cuda.synchronize()
        out[name] = (x, d, time.perf_counter() - t0)
        print(f"[bench] {name} fit: {d['iterations']} iterations in {out[name][2]:.2f}s, cost {d['cost_function']:.9e}",
              file=sys.stderr, flush=True)
    A.set_precision(_lib.FP64)
    x32, d32, t32 = out["fp32"]
    x64, d64, t64 = out["fp64"]
    dxf = np.abs(x32[:nF * nIC] - x64[:nF * nIC]) / max(np.abs(x64[:nF * nIC]).max(), 1e-300)
    rel_x = float(np.linalg.norm(x32 - x64) / max(np.linalg.norm(x64), 1e-300))
    return {"workload": f"group lasso over {ngroups} random bundles + non-negativity, lambda = 0.05 lambda_max, "
                        f"y = A x_true + Rician noise (SNR 30); both: Lipschitz by 30 FP64 power iterations",
            "iterations": iters, "s_fp32": round(t32, 3), "s_fp64": round(t64, 3),
            "ms_per_iteration_fp32": round(1e3 * t32 / max(d32["iterations"], 1), 3),
            "ms_per_iteration_fp64": round(1e3 * t64 / max(d64["iterations"], 1), 3),
            "rel_err_x": rel_x, "per_fibre_err_max": float(dxf.max()), "per_fibre_err_p99": float(np.percentile(dxf, 99)),
            "rel_err_objective": float(abs(d32["cost_function"] - d64["cost_function"]) / abs(d64["cost_function"])),
            "fp64_cost": d64["cost_function"], "fp32_cost": d32["cost_function"]}


def batched_fit_mode(args):
    """Eight tractograms of 50k..400k fibres (2.5 mm head, nS=288), each with a
    16-value group-lasso lambda sweep over 64 random bundles, fitted together:
    one stacked plan, float32 multi-RHS products, device-side FISTA control,
    graph-replayed trial steps."""
    import torch
    from commit_amd import _lib, autotune, batched, solvers, synth
    from commit_amd.operator.operator import LinearOperator
    dev = torch.device("cuda", 0)
    sizes = [50_000 * (k + 1) for k in range(8)]
    t0 = time.time()
    probs = [synth.make_problem("brain25", seed=2000 + k, nF=nf) for k, nf in enumerate(sizes)]
    K = probs[0][1]
    C = batched.Combined([p[0] for p in probs], K)
    A = C.operator(precision=_lib.FP64)
    print(f"[bench] {len(sizes)} tractograms stacked in {time.time() - t0:.1f}s: nF={C.F} nVOX={C.V} "
          f"nnz={C.D['IC']['n']} rows={C.m}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(3000)
    xt = torch.from_numpy(synth.x_true(C.n, 3001)).to(dev)
    s = A.matvec(xt)
    g = torch.Generator(device=dev).manual_seed(3002)
    sig = float(s.mean()) / 30.0
    y = torch.sqrt((s + sig * torch.randn(C.m, generator=g, dtype=torch.float64, device=dev)) ** 2 +
                   (sig * torch.randn(C.m, generator=g, dtype=torch.float64, device=dev)) ** 2)
    aty = A.T.matvec(y).cpu().numpy()
    groups, gw, gs, lam_max = [], [], [], np.zeros(C.S)
    for sidx in range(C.S):
        lab = rng.integers(0, 64, C.nF[sidx])
        for b in range(64):
            gi = (np.nonzero(lab == b)[0] + C.fo[sidx]).astype(np.int32)
            if gi.size:
                groups.append(gi)
                gw.append(np.sqrt(gi.size))
                gs.append(sidx)
                lam_max[sidx] = max(lam_max[sidx], np.linalg.norm(aty[gi]) / gw[-1])
    gidx = np.empty(len(groups), dtype=object)
    for k, gi in enumerate(groups):
        gidx[k] = gi
    perc = np.logspace(-4, -1, 16)
    lam = lam_max[:, None] * perc[None, :]
    rp = {"startIC": 0, "sizeIC": C.nIC * C.F, "startEC": C.nIC * C.F, "sizeEC": C.nEC * C.E,
          "startISO": C.nIC * C.F + C.nEC * C.E, "sizeISO": C.nISO * C.I, "regIC": "group_lasso", "regEC": None,
          "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True, "lambdaIC": 0.0,
          "dictIC_params": {"group_idx_kept": gidx, "group_weights": np.asarray(gw)}}
    tuned = autotune.tune(A, nrhs=16, reps=2)
    y32 = y.float()
    bf = batched.BatchedFista(A, y32, rp, lam, 16, layout=C, tol_fun=1e-5, tol_x=1e-6, max_iter=args.fit_iters,
                              combine=args.residual == "combine")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bf.initialise()
    bf.capture()
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    steps = bf.run(use_graph=True)
    torch.cuda.synchronize()
    t_run = time.perf_counter() - t0
    det = bf.details()
    iters = [det[s_][r]["iterations"] for s_ in range(C.S) for r in range(16)]
    print(f"[bench] batched fits: {steps} trial steps in {t_run:.2f}s (init {t_init:.2f}s); iterations "
          f"min {min(iters)} max {max(iters)} total {sum(iters)}", file=sys.stderr, flush=True)
    # one fit re-run alone with the float64 single-vector solver on its own tractogram
    s0, r0 = 0, 8
    D0, K0 = probs[s0][0], probs[s0][1]
    A0 = LinearOperator(D0, K0, None)
    rows0, cols0 = C.rows(s0), C.columns(s0)
    gi0 = [gi - C.fo[s0] for gi in groups if gi[0] < C.fo[s0] + C.nF[s0] and gi[0] >= C.fo[s0]]
    ga = np.empty(len(gi0), dtype=object)
    for k, gi in enumerate(gi0):
        ga[k] = gi.astype(np.int32)
    rp0 = {"startIC": 0, "sizeIC": C.nF[s0] * C.nIC, "startEC": C.nF[s0] * C.nIC, "sizeEC": C.nE[s0] * C.nEC,
           "startISO": C.nF[s0] * C.nIC + C.nE[s0] * C.nEC, "sizeISO": C.nI[s0] * C.nISO, "regIC": "group_lasso",
           "regEC": None, "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True, "lambdaIC": float(lam[s0, r0]),
           "dictIC_params": {"group_idx_kept": ga, "group_weights": np.sqrt([gi.size for gi in gi0])}}
    y0 = y32[torch.from_numpy(rows0).to(dev)].double()
    x0, d0 = solvers.solve(y0, A0, A0.T, tol_fun=1e-5, tol_x=1e-6, max_iter=args.fit_iters, verbose=0,
                           regularisation=solvers.init_regularisation(rp0, device=dev))
    xb = bf.solution()[cols0, r0]
    err = float(np.linalg.norm(xb - x0) / max(np.linalg.norm(x0), 1e-300))
    # The plain GPU form per right-hand side: the single-vector float32 solver
    # on the same stacked plan, one lambda column (r0) for all 8 tractograms at
    # once (each tractogram's lambda carried in its groups' weights).
    del bf
    torch.cuda.empty_cache()
    A.set_precision(_lib.FP32)
    rps = dict(rp)
    rps["lambdaIC"] = 1.0
    rps["dictIC_params"] = {"group_idx_kept": gidx,
                            "group_weights": np.asarray(gw) * lam[np.asarray(gs), r0]}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, ds = solvers.solve(y32.double(), A, A.T, tol_fun=1e-5, tol_x=1e-6, max_iter=args.fit_iters, verbose=0,
                           regularisation=solvers.init_regularisation(rps, device=dev))
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    A.set_precision(_lib.FP64)
    ms_single = 1e3 * t_single / max(ds["iterations"], 1)
    ms_batched = 1e3 * t_run / max(sum(iters), 1)
    print(f"[bench] single-vector FP32 fit of lambda column {r0} on the stack: {ds['iterations']} iterations in "
          f"{t_single:.2f}s", file=sys.stderr, flush=True)
    line = {"metric": "batched group-lasso fits: fit iterations/s (8 tractograms x 16 lambdas, nS=288)",
            "value": round(sum(iters) / t_run, 2), "unit": "fit-iterations/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f32 storage, f64 control", "data": "synthetic",
            "config": {"workload": f"8 tractograms of {sizes[0]}..{sizes[-1]} fibres (2.5 mm head), "
                                   f"{C.F} fibres, {C.D['IC']['n']} segments, {C.V} voxels, 16 lambdas "
                                   "log-spaced 1e-4..1e-1 x lambda_max over 64 random bundles each",
                       "fits": C.S * 16, "residual": args.residual,
                       "trial_steps": steps, "s_run": round(t_run, 3), "s_init": round(t_init, 3),
                       "ms_per_trial_step": round(1e3 * t_run / max(steps, 1), 3),
                       "iterations_min": min(iters), "iterations_max": max(iters),
                       "check_fit": {"tractogram": s0, "lambda_index": r0, "rel_err_x_vs_fp64_single": err,
                                     "iterations_batched": det[s0][r0]["iterations"],
                                     "iterations_single": d0["iterations"]},
                       "plain_fp32_fit": {"what": f"single-vector FP32 solver, all {C.S} tractograms in one plan, "
                                                  f"lambda column {r0}", "iterations": ds["iterations"],
                                          "s": round(t_single, 3), "ms_per_iteration": round(ms_single, 3),
                                          "ms_per_tractogram_fit_iteration": round(ms_single / C.S, 4)},
                       "ms_per_fit_iteration": round(ms_batched, 4),
                       "speedup_vs_plain_fp32_per_fit_iteration": round(ms_single / C.S / ms_batched, 2),
                       "autotune": tuned}}
    print(json.dumps(line), flush=True)


def autograd_mode(args):
    """Differentiable operator on 200k fibres, nS=288, float32 vectors: time of
    one forward + backward (x and table gradients) and the diffusivity fit
    from (1.7e-3, 0.6e-3) to the (1.9e-3, 0.5e-3) that generated y."""
    import torch
    from commit_amd import _lib, autograd, synth
    from commit_amd.operator.operator import LinearOperator
    dev = torch.device("cuda", 0)
    D, K, raw = synth.make_problem("brain200k", seed=1003)
    K = dict(K)
    K["dirs"] = raw["dirs"]
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    from commit_amd import autotune
    tuned = autotune.tune(A, nrhs=0, reps=3, multi=False)          # launch options of the products
    print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    xt = torch.from_numpy(synth.x_true(A.shape[1], 1004)).to(dev, torch.float32)
    truth = (1.9e-3, 0.5e-3)
    tt = lambda v: torch.tensor(v, dtype=torch.float64, device=dev)
    with torch.no_grad():
        tab = autograd.szb_tables(tt(truth[0]), tt(truth[1]), K["bvals"], K["bvecs"], K["dirs"])
        y = autograd.apply(A, xt, *[t.float() for t in tab]).clone()
    tab0 = [t.float().requires_grad_(True) for t in
            autograd.szb_tables(tt(1.7e-3), tt(0.6e-3), K["bvals"], K["bvecs"], K["dirs"])]
    x = xt.clone().requires_grad_(True)

    def fb():
        r = autograd.apply(A, x, *tab0) - y
        loss = 0.5 * torch.dot(r, r)
        loss.backward()

    fb()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        fb()
    torch.cuda.synchronize()
    ms_fb = (time.perf_counter() - t0) * 1e3 / 5
    hist = []
    t0 = time.perf_counter()
    dp, dq = autograd.fit_diffusivities(A, xt, y, K, (1.7e-3, 0.6e-3), iters=40, history=hist)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    print(f"[bench] autograd: forward+backward {ms_fb:.2f} ms; fit d_par {dp:.6e} d_perp {dq:.6e} "
          f"in {len(hist)} evaluations, {t_fit:.2f}s", file=sys.stderr, flush=True)
    line = {"metric": "differentiable operator: forward+backward per second (200k fibres, nS=288)",
            "value": round(1e3 / ms_fb, 3), "unit": "forward+backward/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32 vectors, f64 accumulation and table gradients", "data": "synthetic",
            "config": {"workload": f"{D['IC']['nSTR']} fibres, {D['IC']['n']} segments, {D['IC']['nVOX']} voxels",
                       "ms_forward_backward": round(ms_fb, 3),
                       "fit": {"start": [1.7e-3, 0.6e-3], "truth": list(truth), "recovered": [dp, dq],
                               "evaluations": len(hist), "s": round(t_fit, 3),
                               "loss_first": hist[0][0], "loss_last": hist[-1][0]},
                       "autotune": tuned}}
    print(json.dumps(line), flush=True)


def products_mode(args):
    """The single-vector products on the full-brain workload with the plan's
    slab option fixed (no autotuning, so a profiler sees only these kernels):
    --steps pairs in FP64, then in FP32 storage, each entry point and A^T pass
    timed with HIP events on its launch stream."""
    import ctypes

    import torch
    from commit_amd import _lib, autotune, synth
    from commit_amd.operator.operator import LinearOperator
    dev = torch.device("cuda", 0)
    D, K, _ = synth.make_problem(args.workload, seed=1002, nF=args.fibres)
    A = LinearOperator(D, K, None)
    At = A.T
    autotune.set_option(A, _lib.OPT_SLAB, args.slab)
    lib = _lib.load()
    h = A._plan.handle
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    g = torch.Generator(device=dev).manual_seed(1002)
    out = {"workload": f"{args.workload}: {D['IC']['nSTR']} fibres, {D['IC']['n']} segments, "
                       f"{D['IC']['nVOX']} voxels", "slab": args.slab, "steps": args.steps}
    for name, prec, dt in (("fp64", _lib.FP64, torch.float64), ("fp32", _lib.FP32, torch.float32)):
        A.set_precision(prec)
        x = torch.rand(A.shape[1], dtype=torch.float64, device=dev, generator=g).to(dt)
        y = torch.empty(A.shape[0], dtype=dt, device=dev)
        xo = torch.empty_like(x)
        ms = {}
        for key, fn in (("A", lambda: A.matvec(x, out=y)), ("At", lambda: At.matvec(y, out=xo)),
                        ("At_pass1", lambda: _lib.check(lib.commit_hip_At_pass(
                            h, ctypes.c_void_p(y.data_ptr()), None, ctypes.c_void_p(xo.data_ptr()), 1, sp), "pass 1")),
                        ("At_pass2", lambda: _lib.check(lib.commit_hip_At_pass(
                            h, ctypes.c_void_p(y.data_ptr()), None, ctypes.c_void_p(xo.data_ptr()), 2, sp), "pass 2"))):
            fn()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(args.steps):
                fn()
            ev1.record(stream)
            ev1.synchronize()
            ms[key] = round(ev0.elapsed_time(ev1) / args.steps, 4)
        out[name] = ms
        print(f"[bench] products {name}: {ms}", file=sys.stderr, flush=True)
    A.set_precision(_lib.FP64)
    print(json.dumps(out), flush=True)


def voxelize_mode(args):
    """Tractogram voxelisation (commit_hip_trk2dictionary) of the full-brain
    tractogram: 1M streamlines sampled every 0.625 mm (~1.8e8 points) on the
    145x174x145 grid at 1.25 mm with the ~600k-voxel mask, streamlines of
    20-250 mm kept.  Points and offsets resident in HBM; a step is one whole
    export (walks, sorts, merges, norms, TDI, EC and ISO).  CPU baseline: the
    oracle's C restatement on the host cores, the same tractogram."""
    import ctypes

    import torch
    from commit_amd import synth
    from commit_amd import trk2dictionary as t2d
    dev = torch.device("cuda", 0)
    cfg = synth.WORKLOADS["brain"]
    dim, pix = cfg["dim"], (cfg["pixdim"],) * 3
    nF = args.fibres or cfg["nF"]
    mask = synth.ellipsoid_mask(dim, cfg["mask"][1]).astype(np.float32)
    t0 = time.time()
    pts, ptr = synth.streamlines(1002, nF, dim, cfg["pixdim"], mask=mask, mode=1, step_vox=0.5)
    print(f"[bench] {nF} streamlines, {pts.shape[0]} points in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    H = t2d.default_hash_table(500)
    tp = torch.from_numpy(pts).to(dev)
    tq = torch.from_numpy(ptr).to(dev)
    stream = torch.cuda.current_stream(dev)
    v = t2d.Voxeliser(0)
    kw = dict(mask=torch.from_numpy(mask).to(dev), min_fiber_len=20.0, max_fiber_len=250.0)

    def step():
        v.run(tp, tq, dim, pix, H, stream=stream.cuda_stream, **kw)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    ev1.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = ev0.elapsed_time(ev1) / args.steps
    s = v.sizes()
    d = v.fetch()
    v.close()
    npts = int(pts.shape[0])
    # compulsory bytes of one export: the points and offsets read once, the
    # dictionary written once (IC 14 B per segment, 12 B + 1 B per kept and
    # input streamline, EC 6 B, ISO 4 B, TDI 4 B per voxel)
    nvox = int(np.prod(dim))
    alg = (12 * npts + 8 * (nF + 1) + 14 * s["n_ic"] + 12 * s["n_kept"] + nF + 6 * s["n_ec"] + 4 * s["n_iso"]
           + 4 * nvox + 4 * nvox)
    print(f"[bench] voxelise: {ms:.2f} ms per export (wall {wall * 1e3:.2f} ms), {s}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_pairs:
        from oracle import voxelize as vo
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        n = max(1, min(cores, 255))
    if args.cpu_pairs and vo.reference_available():
        # the reference's own trk2dictionary() (built from its source into
        # oracle/_ref) on a bounded sample: the first 250k streamlines, read
        # from a .trk file as the reference does, its files read back
        import shutil
        import tempfile
        ns = min(nF, 250_000)
        wd = tempfile.mkdtemp(prefix="vox_ref_")
        try:
            r = vo.reference_trk2dictionary(wd, pts[:ptr[ns]], ptr[:ns + 1], dim, pix, H, n_threads=n, **{
                "mask": mask, "min_fiber_len": 20.0, "max_fiber_len": 250.0})
        finally:
            shutil.rmtree(wd, ignore_errors=True)
        m = r["IC_v"].size
        same = all(np.array_equal(r[k].view(np.uint8), d[k][:m].view(np.uint8))
                   for k in ("IC_f", "IC_v", "IC_o", "IC_len")) and \
            all(np.array_equal(r[k].view(np.uint8), d[k][:r[k].size].view(np.uint8))
                for k in ("TRK_norm", "TRK_len", "TRK_lenTot", "TRK_kept"))
        print(f"[bench] reference trk2dictionary(): {r['seconds']:.2f}s for {ns} streamlines on {n} threads; "
              f"its segments equal the GPU's first {m}: {same}", file=sys.stderr, flush=True)
        cpu = {"value": round(ns / r["seconds"], 1), "unit": "streamlines/s", "cores": n, "kind": "reference",
               "sample": f"the first {ns} streamlines, the reference's trk2dictionary() compiled from its source "
                         f"(g++ -O2), {n} threads, reading them from a .trk file and writing its .dict files; "
                         f"its IC segments and streamline values bitwise equal to the GPU's: {same}"}
    elif args.cpu_pairs:
        t0 = time.perf_counter()
        o = vo.trk2dictionary(pts, ptr, dim, pix, H, n_threads=n, mask=mask, min_fiber_len=20.0,
                              max_fiber_len=250.0)
        tc = time.perf_counter() - t0
        same = all(np.array_equal(np.asarray(o[k]).view(np.uint8), np.asarray(d[k]).view(np.uint8))
                   for k in ("IC_f", "IC_v", "IC_o", "IC_len", "TRK_norm", "TRK_len", "TRK_kept", "ISO_v", "tdi"))
        print(f"[bench] cpu voxelise: {tc:.2f}s on {n} threads; outputs bitwise equal: {same}", file=sys.stderr,
              flush=True)
        cpu = {"value": round(nF / tc, 1), "unit": "streamlines/s", "cores": n, "kind": "port",
               "sample": f"the full tractogram once ({nF} streamlines), oracle C restatement of "
                         f"trk2dictionary_c.cpp, {n} threads; outputs bitwise equal to the GPU's: {same}"}
    line = {"metric": "tractogram voxelisation streamlines/s (1M streamlines, 145x174x145 at 1.25 mm)",
            "value": round(nF / (ms * 1e-3), 1), "unit": "streamlines/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
            "dtype": "f64 geometry, f32 lengths", "data": "synthetic",
            "config": {"workload": f"{nF} streamlines, {npts} points, {s['n_pieces']} pieces, {s['n_ic']} "
                                   f"segments, {s['n_kept']} kept", "wall_ms_per_step": round(wall * 1e3, 3)},
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes": alg, "scope": "whole export"},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


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
