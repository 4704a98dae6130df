"""commit_hip_fista_sums_noalloc (workspace in a fixed device array) against
commit_hip_fista_sums (caller's workspace), through the C ABI: two calls in a
row, the second on smaller inputs, must give bitwise equal sums; group counts
beyond the fixed workspace are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from commit_amd import _lib

pytestmark = pytest.mark.gpu


def _inputs(n, ngroups, seed, dev):
    rng = np.random.default_rng(seed)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    x, xhat, grad, prev = (t(rng.standard_normal(n)) for _ in range(4))
    w = t(rng.random(n) + 0.5)
    lab = np.sort(rng.integers(0, ngroups, n)) if ngroups else np.zeros(0, np.int64)
    gptr = np.zeros(ngroups + 1, np.int32)
    if ngroups:
        gptr[1:] = np.cumsum(np.bincount(lab, minlength=ngroups))
    gidx = rng.permutation(n).astype(np.int32)[: int(gptr[-1])]
    gw = rng.random(ngroups) + 0.1
    l1 = np.array([0, n // 3, n // 3, n // 2, n // 2, n], np.int64)
    l1w = np.array([1, 0, 1], np.int32)
    return dict(x=x, xhat=xhat, grad=grad, prev=prev, w=w, gptr=t(gptr), gidx=t(gidx), gw=t(gw), l1=l1, l1w=l1w,
                n=n, ngroups=ngroups)


def _call(lib, a, out, work=None):
    P = lambda v: ctypes.c_void_p(v.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [a["n"], P(a["x"]), P(a["xhat"]), P(a["grad"]), P(a["prev"]), _lib.ptr(a["l1"]), _lib.ptr(a["l1w"]),
            P(a["w"]), a["ngroups"], P(a["gptr"]), P(a["gidx"]), P(a["gw"]), P(out)]
    if work is None:
        return lib.commit_hip_fista_sums_noalloc(*args, st)
    return lib.commit_hip_fista_sums(*args, P(work), st)


def test_noalloc_sums_equal_workspace_sums_twice():
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    for n, ng, seed in ((400_000, 1000, 1), (3_000, 7, 2)):    # second call smaller
        a = _inputs(n, ng, seed, dev)
        print(f"noalloc sums: n={n} ngroups={ng}", flush=True)
        ref = torch.full((8,), float("nan"), dtype=torch.float64, device=dev)
        got = torch.full((8,), float("nan"), dtype=torch.float64, device=dev)
        work = torch.empty(_lib.sums_work(ng), dtype=torch.float64, device=dev)
        _lib.check(_call(lib, a, ref, work), "commit_hip_fista_sums")
        _lib.check(_call(lib, a, got), "commit_hip_fista_sums_noalloc")
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(r))
        assert np.array_equal(r, g), (r, g)
        # against the definition on the host
        x, xh, gr, pv, w = (a[k].cpu().numpy() for k in ("x", "xhat", "grad", "prev", "w"))
        assert np.isclose(g[0], np.dot(x - xh, gr), rtol=1e-12)
        assert np.isclose(g[2], np.dot(x - pv, x - pv), rtol=1e-12)
        assert np.isclose(g[4], np.abs(w[: n // 3] * x[: n // 3]).sum(), rtol=1e-12)
        gp, gi, gw = a["gptr"].cpu().numpy(), a["gidx"].cpu().numpy(), a["gw"].cpu().numpy()
        gsum = sum(gw[k] * np.linalg.norm(x[gi[gp[k]:gp[k + 1]]]) for k in range(ng))
        assert np.isclose(g[7], gsum, rtol=1e-12)


def test_noalloc_sums_without_groups_and_rejects_too_many():
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    a = _inputs(1000, 0, 3, dev)
    got = torch.full((8,), float("nan"), dtype=torch.float64, device=dev)
    _lib.check(_call(lib, a, got), "commit_hip_fista_sums_noalloc")
    torch.cuda.synchronize()
    assert got[7].item() == 0.0                 # the previous call's group partials do not leak in
    big = dict(a, ngroups=_lib.NOALLOC_MAX_GROUPS + 1)
    assert _call(lib, big, got) == -4
    assert b"groups exceed" in lib.commit_hip_last_error()
