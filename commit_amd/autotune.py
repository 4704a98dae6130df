"""Autotuner for the launch options of a device plan.

Times every variant of the plan's launch options on the plan's own
dictionary (HIP events on the launch stream, after a warm-up call) and keeps
the fastest in the plan:

* ``xcd_remap``       -- blocks renumbered so each of the 8 XCDs walks one
                         contiguous voxel range (1), or hardware order (0);
                         timed on the single-vector A x + A^T y pair;
* ``a_dense``         -- A x per segment (0) or per tile of 16 voxels as a
                         dense FP64 matrix-core product (1); timed with it;
* ``slab``            -- IC rows of A x and A^T y from L2 per entry (0), or the
                         IC table staged in LDS by orientation slabs with
                         work-groups of 8 (1) or 16 (2) wavefronts; timed with
                         them;
* ``multi_a_waves``   -- one or two wavefronts per (voxel, 16 right-hand
                         sides) in the multi-RHS A X; timed on A X.

Choices can be cached in a JSON file keyed by the problem's shape, so a
later process with the same shape skips the timing.
"""
import json
import os

import torch

from . import _lib


def _key(A, nrhs):
    p = A._plan
    return f"nS{p.nS}_nd{p.ndirs}_nF{p.nF}_n{p.n}_nE{p.nE}_nI{p.nV}_nV{p.nVOX}_R{nrhs}"


def _time(fn, reps, dev):
    st = torch.cuda.current_stream(dev)
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def set_option(A, option, value):
    lib = _lib.load()
    _lib.check(lib.commit_hip_plan_set_option(A._plan.handle, int(option), int(value)), "commit_hip_plan_set_option")


def get_option(A, option):
    lib = _lib.load()
    v = lib.commit_hip_plan_get_option(A._plan.handle, int(option))
    if v < 0:
        _lib.check(v, "commit_hip_plan_get_option")
    return int(v)


def tune(A, nrhs=16, reps=3, cache=None, multi=True):
    """Pick the fastest launch options for operator A; returns a report dict
    {option: {"choice": v, "ms": {variant: time}}}.  cache: JSON path or None."""
    key = _key(A, nrhs)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if key in known:
            for name, opt in (("slab", _lib.OPT_SLAB), ("xcd_remap", _lib.OPT_XCD_REMAP),
                              ("a_dense", _lib.OPT_A_DENSE), ("multi_a_waves", _lib.OPT_MULTI_A_WAVES)):
                if name in known[key]:
                    set_option(A, opt, known[key][name]["choice"])
            return known[key]
    dev = A.device
    At = A.T
    report = {}
    g = torch.Generator(device=dev).manual_seed(7)
    x = torch.rand(A.shape[1], dtype=A.dtype, device=dev, generator=g)
    y = torch.empty(A.shape[0], dtype=A.dtype, device=dev)
    xo = torch.empty_like(x)

    def pair():
        A.matvec(x, out=y)
        At.matvec(y, out=xo)

    ms = {}
    for slab in (0, 1, 2):
        set_option(A, _lib.OPT_SLAB, slab)
        for remap in (1, 0):
            for dense in (0, 1):
                set_option(A, _lib.OPT_XCD_REMAP, remap)
                set_option(A, _lib.OPT_A_DENSE, dense)
                ms[(remap, dense, slab)] = _time(pair, reps, dev)
    best = min(ms, key=ms.get)
    set_option(A, _lib.OPT_SLAB, best[2])
    set_option(A, _lib.OPT_XCD_REMAP, best[0])
    set_option(A, _lib.OPT_A_DENSE, best[1])
    report["xcd_remap"] = {"choice": best[0], "ms": {f"remap={k[0]},dense={k[1]},slab={k[2]}": round(t, 4)
                                                     for k, t in ms.items()}}
    report["a_dense"] = {"choice": best[1]}
    report["slab"] = {"choice": best[2]}
    del x, y, xo
    if multi and nrhs > 0:
        X = torch.rand((A.shape[1], nrhs), dtype=torch.float32, device=dev, generator=g)
        Y = torch.empty((A.shape[0], nrhs), dtype=torch.float32, device=dev)
        ms = {}
        for v in (2, 1):
            set_option(A, _lib.OPT_MULTI_A_WAVES, v)
            ms[v] = _time(lambda: A.matmat(X, out=Y), reps, dev)
        best = min(ms, key=ms.get)
        set_option(A, _lib.OPT_MULTI_A_WAVES, best)
        report["multi_a_waves"] = {"choice": best, "ms": {str(k): round(t, 4) for k, t in ms.items()}}
        del X, Y
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[key] = report
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(known, f, indent=1)
    return report
