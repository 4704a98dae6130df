"""Differentiable operator (commit_amd.autograd) on the GPU.

Bars: the LUT gradient kernel equals the oracle's restatement to 1e-12
relative in FP64 (1e-6 with float32 vectors); autograd's gradients with respect
to x and to the tables match central finite differences of
0.5 ||A(LUT) x - y||^2 (x: h = 1e-4 relative, rtol 1e-6; tables: h = 1e-3
relative on the float32-rounded entries, rtol 1e-4) and the dense A^T; the
diffusivity fit recovers (d_par, d_perp) to 1e-3 relative.
"""
import ctypes

import numpy as np
import pytest
import torch

from _problems import random_dictionary, rel
from commit_amd import _lib, autograd, synth
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu


def _lut_grad(A, x, gy, shapes):
    out = [torch.zeros(s, dtype=torch.float64, device="cuda") if np.prod(s) else None for s in shapes]
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.commit_hip_lut_grad(A._plan.handle, P(x), P(gy), P(out[0]), P(out[1]), P(out[2]), st),
               "commit_hip_lut_grad")
    return [np.zeros(s) if o is None else o.cpu().numpy() for o, s in zip(out, shapes)]


@pytest.mark.parametrize("ndirs", [3, 250, 251, 600])
@pytest.mark.parametrize("nS", [7, 64, 65, 288])
def test_lut_grad_vs_oracle(ndirs, nS):
    D, K = random_dictionary(ndirs + nS, nVOX=11, nF=17, n=260, nS=nS, ndirs=ndirs, nIC=2, nEC=1, nISO=2,
                             empty_voxels=(4,), ec_per_voxel=2)
    ev = np.ones(34, np.uint32)
    ev[[2, 20]] = 0
    D["IC"]["eval"] = ev
    rng = np.random.default_rng(ndirs * 7 + nS)
    for prec, tol in ((_lib.FP64, 1e-12), (_lib.FP32, 1e-6)):
        A = LinearOperator(D, K, None, precision=prec)
        dt = torch.float64 if prec == _lib.FP64 else torch.float32
        x = rng.standard_normal(A.shape[1])
        gy = rng.standard_normal(A.shape[0])
        xd, gyd = torch.from_numpy(x).to("cuda", dt), torch.from_numpy(gy).to("cuda", dt)
        g = _lut_grad(A, xd, gyd, (K["wmr"].shape, K["wmh"].shape, K["iso"].shape))
        go = oracle.lut_grad(D, K, xd.double().cpu().numpy(), gyd.double().cpu().numpy())
        for a, b in zip(g, go):
            assert rel(a.ravel(), b.ravel()) <= tol


def test_autograd_gradients_vs_finite_differences():
    D, K = random_dictionary(3, nVOX=10, nF=14, n=160, nS=30, ndirs=9, nIC=1, nEC=1, nISO=2)
    A = LinearOperator(D, K, None, precision=_lib.FP64)
    rng = np.random.default_rng(3)
    x0 = rng.standard_normal(A.shape[1])
    y0 = rng.standard_normal(A.shape[0])
    y = torch.from_numpy(y0).cuda()
    tabs = [torch.from_numpy(K[k]).cuda().requires_grad_(True) for k in ("wmr", "wmh", "iso")]
    x = torch.from_numpy(x0).cuda().requires_grad_(True)

    def loss_of(xv, wmr, wmh, iso):
        r = autograd.apply(A, xv, wmr, wmh, iso) - y
        return 0.5 * torch.dot(r, r)

    L = loss_of(x, *tabs)
    L.backward()
    dense = oracle.assemble_A(D, K)
    gx_ref = dense.T @ (dense @ x0 - y0)
    assert rel(x.grad.cpu().numpy(), gx_ref) <= 1e-12
    go = oracle.lut_grad(D, K, x0, dense @ x0 - y0)
    for t, b in zip(tabs, go):
        assert rel(t.grad.double().cpu().numpy().ravel(), b.ravel()) <= 1e-6     # float32 gradient tensors
    with torch.no_grad():
        # x: central differences
        for j in rng.choice(A.shape[1], 6, replace=False):
            h = 1e-4 * max(1.0, abs(x0[j]))
            e = torch.zeros_like(x)
            e[j] = h
            fd = (loss_of(x + e, *tabs) - loss_of(x - e, *tabs)).item() / (2 * h)
            assert abs(fd - x.grad[j].item()) <= 1e-6 * max(1.0, abs(fd)), (j, fd, x.grad[j].item())
        # tables: central differences on the float32-rounded entries
        for k, t in enumerate(tabs):
            flat = t.detach().reshape(-1)
            for j in rng.choice(flat.numel(), 4, replace=False):
                v = flat[j].item()
                h = 1e-3 * max(abs(v), 1e-3)
                up, dn = [tt.detach().clone() for tt in tabs], [tt.detach().clone() for tt in tabs]
                up[k].reshape(-1)[j] = v + h
                dn[k].reshape(-1)[j] = v - h
                heff = up[k].reshape(-1)[j].item() - dn[k].reshape(-1)[j].item()
                fd = (loss_of(x, *up) - loss_of(x, *dn)).item() / heff
                g = t.grad.reshape(-1)[j].item()
                assert abs(fd - g) <= 1e-4 * max(1e-3, abs(fd)), (k, j, fd, g)


def test_fit_diffusivities_recovers_truth():
    """Noise-free y generated at (1.9e-3, 0.5e-3); L-BFGS from (1.7e-3, 0.6e-3)."""
    D, K, raw = synth.make_problem("small", seed=1001, nF=3000)
    K = dict(K)
    K["dirs"] = raw["dirs"]
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    xt = torch.from_numpy(synth.x_true(A.shape[1], 2001)).cuda().float()
    truth = (1.9e-3, 0.5e-3)
    with torch.no_grad():
        t = autograd.szb_tables(torch.tensor(truth[0], dtype=torch.float64, device="cuda"),
                                torch.tensor(truth[1], dtype=torch.float64, device="cuda"),
                                K["bvals"], K["bvecs"], K["dirs"])
        y = autograd.apply(A, xt, t[0].float(), t[1].float(), t[2].float()).clone()
    hist = []
    dp, dq = autograd.fit_diffusivities(A, xt, y, K, (1.7e-3, 0.6e-3), iters=60, history=hist)
    print(f"diffusivity fit: d_par {dp:.6e} d_perp {dq:.6e} after {len(hist)} evaluations "
          f"(loss {hist[0][0]:.3e} -> {hist[-1][0]:.3e})", flush=True)
    assert abs(dp - truth[0]) <= 1e-3 * truth[0]
    assert abs(dq - truth[1]) <= 1e-3 * truth[1]


# ------------------------------------------ 200k fibres, nS=288, float32 --
@pytest.fixture(scope="module")
def brain200k():
    """The differentiable-operator workload: 200k fibres on the full-brain grid,
    nS = 288, float32 vectors, LDS-slab products."""
    from commit_amd import autotune
    D, K, raw = synth.make_problem("brain200k", seed=1003)
    K = dict(K)
    K["dirs"] = raw["dirs"]
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    print(f"brain200k: {D['IC']['n']} segments, {D['IC']['nVOX']} voxels", flush=True)
    return D, K, A


def _set_tables(A, wmr, wmh, iso):
    lib = _lib.load()
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.commit_hip_plan_set_lut(A._plan.handle, P(wmr), P(wmh), P(iso), st), "commit_hip_plan_set_lut")


def test_lut_grad_bilinear_identity_200k(brain200k):
    """A is linear in the tables, so for any perturbation delta of one table
    sum(G * delta) = <gy, A(delta) x> exactly (G the table gradient of
    <gy, A x>); checked per table with the other tables set to zero."""
    D, K, A = brain200k
    g = torch.Generator(device="cuda").manual_seed(21)
    x = torch.rand(A.shape[1], dtype=torch.float32, device="cuda", generator=g)
    gy = torch.randn(A.shape[0], dtype=torch.float32, device="cuda", generator=g)
    tabs = [torch.from_numpy(np.ascontiguousarray(K[k], dtype=np.float32)).cuda() for k in ("wmr", "wmh", "iso")]
    G = _lut_grad(A, x, gy, [t.shape for t in tabs])
    Gabs = _lut_grad(A, x, gy.abs(), [t.shape for t in tabs])     # x >= 0: sum |terms| = Gabs . |delta|
    for k in range(3):
        delta = torch.randn(tabs[k].shape, dtype=torch.float32, device="cuda", generator=g)
        parts = [torch.zeros_like(t) for t in tabs]
        parts[k] = delta
        _set_tables(A, *parts)
        rhs = float(torch.dot(gy.double(), A.dot(x).double()).item())
        lhs = float((G[k] * delta.double().cpu().numpy()).sum())
        # FP32 products: the error is relative to sum |terms|
        scale = float((Gabs[k] * np.abs(delta.double().cpu().numpy())).sum())
        print(f"table {k}: sum G*delta {lhs:.9e}  <gy, A(delta) x> {rhs:.9e}", flush=True)
        assert abs(lhs - rhs) <= 1e-6 * scale, (k, lhs, rhs, scale)
    _set_tables(A, *tabs)


def test_x_gradient_is_the_adjoint_200k(brain200k):
    """x.grad of <gy, A x> through autograd equals A^T gy: <g_x, v> = <A v, gy>."""
    D, K, A = brain200k
    g = torch.Generator(device="cuda").manual_seed(22)
    x = torch.rand(A.shape[1], dtype=torch.float32, device="cuda", generator=g).requires_grad_(True)
    gy = torch.randn(A.shape[0], dtype=torch.float32, device="cuda", generator=g)
    loss = torch.dot(autograd.apply(A, x), gy)
    loss.backward()
    for s in range(3):
        v = torch.rand(A.shape[1], dtype=torch.float32, device="cuda", generator=g)
        a = float(torch.dot(x.grad.double(), v.double()).item())
        b = float(torch.dot(A.dot(v).double(), gy.double()).item())
        scale = float(torch.dot(A.T.dot(gy.abs()).double(), v.double()).item())
        assert abs(a - b) <= 1e-6 * scale, (a, b, scale)


def test_diffusivity_gradient_vs_finite_differences_200k(brain200k):
    """d/d(d_par, d_perp) of 0.5 ||A(theta) x - y||^2 through the tables
    (szb_tables -> commit_hip_lut_grad) against central differences with
    h = 1e-4 relative, in FP64 vectors on the same plan.  The tables are
    float32, which bounds the differences' accuracy near 6e-8 / 1e-4."""
    D, K, A = brain200k
    A.set_precision(_lib.FP64)
    try:
        xt = torch.from_numpy(synth.x_true(A.shape[1], 1004)).cuda()
        tt = lambda v: torch.tensor(v, dtype=torch.float64, device="cuda")
        with torch.no_grad():
            y = autograd.apply(A, xt, *[t.float() for t in autograd.szb_tables(
                tt(1.9e-3), tt(0.5e-3), K["bvals"], K["bvecs"], K["dirs"])]).clone()

        def loss(dp, dq, grad=False):
            p = tt(dp).requires_grad_(grad), tt(dq).requires_grad_(grad)
            tabs = autograd.szb_tables(p[0], p[1], K["bvals"], K["bvecs"], K["dirs"])
            r = autograd.apply(A, xt, *[t.float() for t in tabs]) - y
            L = 0.5 * torch.dot(r, r)
            if grad:
                L.backward()
                return float(L.item()), float(p[0].grad.item()), float(p[1].grad.item())
            return float(L.item())

        theta = (1.7e-3, 0.6e-3)
        _, g0, g1 = loss(*theta, grad=True)
        for k, gk in ((0, g0), (1, g1)):
            h = 1e-4 * theta[k]
            up, dn = list(theta), list(theta)
            up[k] += h
            dn[k] -= h
            fd = (loss(*up) - loss(*dn)) / (2 * h)
            print(f"theta {k}: autograd {gk:.9e}  central difference {fd:.9e}", flush=True)
            assert abs(fd - gk) <= 5e-3 * abs(fd), (k, fd, gk)
    finally:
        A.set_precision(_lib.FP32)
        _set_tables(A, *[torch.from_numpy(np.ascontiguousarray(K[k], dtype=np.float32)).cuda()
                         for k in ("wmr", "wmh", "iso")])


def test_fit_diffusivities_200k(brain200k):
    """(d_par, d_perp) = (1.9e-3, 0.5e-3) recovered from (1.7e-3, 0.6e-3) to
    1e-3 relative; afterwards the plan holds the fitted tables (A.dot equals a
    fresh operator built on the updated KERNELS)."""
    D, K, A = brain200k
    xt = torch.from_numpy(synth.x_true(A.shape[1], 1004)).cuda().float()
    tt = lambda v: torch.tensor(v, dtype=torch.float64, device="cuda")
    with torch.no_grad():
        y = autograd.apply(A, xt, *[t.float() for t in autograd.szb_tables(
            tt(1.9e-3), tt(0.5e-3), K["bvals"], K["bvecs"], K["dirs"])]).clone()
    hist = []
    dp, dq = autograd.fit_diffusivities(A, xt, y, A.KERNELS, (1.7e-3, 0.6e-3), iters=40, history=hist)
    print(f"200k fit: d_par {dp:.6e} d_perp {dq:.6e} in {len(hist)} evaluations", flush=True)
    assert abs(dp - 1.9e-3) <= 1e-3 * 1.9e-3 and abs(dq - 0.5e-3) <= 1e-3 * 0.5e-3
    B = LinearOperator(D, A.KERNELS, None, precision=_lib.FP32)
    assert rel(A.dot(xt).double().cpu().numpy(), B.dot(xt).double().cpu().numpy()) <= 1e-6
