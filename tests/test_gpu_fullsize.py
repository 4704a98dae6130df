"""Parity at the headline size: the 1M-fibre head (84,376,486 segments,
607,680 voxels, nS=288, SZB 1 IC + 1 EC + 2 ISO), the device products
against the oracle's C operator on the same inputs (on all host cores), plus
size-independent properties.

Bars: FP64 products <= 1e-12 relative (vector norm) to the oracle, with the
slab kernels and with the per-voxel kernels; FP32 storage <= 1e-6 against the
oracle fed the same float32-rounded inputs; 16-RHS columns <= 1e-5 against the
FP64 products; adjointness and linearity <= 1e-12; the fused residual of the
batched fits equals the two-call result bitwise.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

from _problems import rel
from commit_amd import _lib, autotune, batched, synth
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def brain():
    D, K, _ = synth.make_problem("brain", seed=1001)
    A = LinearOperator(D, K, None)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    O = oracle.OracleOperator(D, K, oracle.set_threads(D, max(1, min(cores, 16))))
    rng = np.random.default_rng(11)
    x = rng.random(A.shape[1])
    x[rng.random(x.size) < 0.5] = 0.0
    y = rng.random(A.shape[0])
    return A, O, x, y, O.dot(x), O.T.dot(y)


@pytest.mark.parametrize("slab", [2, 0])
def test_fullsize_fp64_vs_oracle(brain, slab):
    A, O, x, y, Ax_o, Aty_o = brain
    autotune.set_option(A, _lib.OPT_SLAB, slab)
    e1, e2 = rel(A.dot(x), Ax_o), rel(A.T.dot(y), Aty_o)
    print(f"1M fibres, slab {slab}: A x {e1:.2e}  A^T y {e2:.2e}", flush=True)
    assert e1 <= TOL and e2 <= TOL


def test_fullsize_properties(brain):
    A, O, x, y, Ax_o, Aty_o = brain
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    g = torch.Generator(device="cuda").manual_seed(5)
    x1 = torch.rand(A.shape[1], dtype=torch.float64, device="cuda", generator=g)
    x2 = torch.rand(A.shape[1], dtype=torch.float64, device="cuda", generator=g)
    yd = torch.rand(A.shape[0], dtype=torch.float64, device="cuda", generator=g)
    Ax1 = A.dot(x1)
    lhs = torch.dot(Ax1, yd).item()
    rhs = torch.dot(x1, A.T.dot(yd)).item()
    assert abs(lhs - rhs) <= TOL * abs(lhs), (lhs, rhs)
    lin = A.dot(2.0 * x1 - 3.0 * x2)
    ref = 2.0 * Ax1 - 3.0 * A.dot(x2)
    assert (torch.linalg.norm(lin - ref) / torch.linalg.norm(ref)).item() <= TOL
    assert torch.equal(A.dot(x1), Ax1)                       # bitwise repeatable


def test_fullsize_fp32_vs_oracle(brain):
    A, O, x, y, Ax_o, Aty_o = brain
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    x32 = x.astype(np.float32).astype(np.float64)
    y32 = y.astype(np.float32).astype(np.float64)
    A.set_precision(_lib.FP32)
    try:
        e1, e2 = rel(A.dot(x32), O.dot(x32)), rel(A.T.dot(y32), O.T.dot(y32))
    finally:
        A.set_precision(_lib.FP64)
    print(f"1M fibres, FP32 storage: A x {e1:.2e}  A^T y {e2:.2e}", flush=True)
    assert e1 <= 1e-6 and e2 <= 1e-6


def test_fullsize_multi_rhs_and_fused_residual(brain):
    A, O, x, y, Ax_o, Aty_o = brain
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    n, m = A.shape[1], A.shape[0]
    g = torch.Generator(device="cuda").manual_seed(6)
    X = torch.rand((n, 16), dtype=torch.float32, device="cuda", generator=g)
    Yt = torch.rand((m, 16), dtype=torch.float32, device="cuda", generator=g)
    AX = A.matmat(X)
    AtY = A.T.matmat(Yt)
    for r in (0, 7, 15):
        e1 = rel(AX[:, r].double().cpu().numpy(), A.matvec(X[:, r].contiguous().double()).cpu().numpy())
        e2 = rel(AtY[:, r].double().cpu().numpy(), A.T.matvec(Yt[:, r].contiguous().double()).cpu().numpy())
        assert e1 <= 1e-5 and e2 <= 1e-5, (r, e1, e2)
    del AtY
    # the fused residual of the batched fits against the two-call path
    yv = torch.rand(m, dtype=torch.float32, device="cuda", generator=g)
    rp = {"startIC": 0, "sizeIC": n, "startEC": n, "sizeEC": 0, "startISO": n, "sizeISO": 0, "regIC": None,
          "regEC": None, "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True}
    bf = batched.BatchedFista(A, yv, rp, np.zeros((1, 16)), 16, max_iter=2)
    lib = _lib.load()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    two = torch.empty((m, 16), dtype=torch.float32, device="cuda")
    rsq2 = torch.empty(16, dtype=torch.float64, device="cuda")
    work = torch.empty(bf.L.nypieces * 16 + 1, dtype=torch.float64, device="cuda")
    _lib.check(lib.commit_hip_multi_residual(ctypes.byref(bf.L), P(AX), P(yv), 0, None, P(two), P(rsq2), P(work), st),
               "multi_residual")
    del AX
    rsq1 = torch.empty(16, dtype=torch.float64, device="cuda")
    bf._resid(X, bf.y, 0, bf.res, rsq1)
    assert torch.equal(bf.res, two)
    d = ((rsq1 - rsq2).abs() / rsq2).max().item()
    print(f"1M fibres, fused residual: max rel. difference of the 16 norms {d:.1e}", flush=True)
    assert d <= 1e-13


# --------------------------------------- fits at the headline size vs the oracle --
def _nn_rp(D, K):
    nF, nE, nI = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
    nIC, nEC, nISO = K["wmr"].shape[0], K["wmh"].shape[0], K["iso"].shape[0]
    return {"startIC": 0, "sizeIC": nF * nIC, "startEC": nF * nIC, "sizeEC": nE * nEC,
            "startISO": nF * nIC + nE * nEC, "sizeISO": nI * nISO, "regIC": None, "regEC": None, "regISO": None,
            "nnIC": True, "nnEC": True, "nnISO": True}


def _compare_fits(dg, do, xg, xo, tol_x):
    e = rel(xg, xo)
    print(f"1M-fibre fit: iterations {dg['iterations']}/{do['iterations']}, rel x {e:.2e}, "
          f"cost {dg['cost_function']:.12e} / {do['cost_function']:.12e}", flush=True)
    assert dg["iterations"] == do["iterations"]
    assert dg["stopping_criterion"] == do["stopping_criterion"]
    assert e <= tol_x
    return e


@pytest.fixture(scope="module")
def brain_signal(brain):
    A, O, x, y, Ax_o, Aty_o = brain
    return synth.rician(Ax_o, 30.0, 7)


def test_fullsize_fp64_nn_fit_vs_oracle(brain, brain_signal):
    """The 1M-fibre FP64 non-negative fit, 10 iterations, against the oracle's
    FISTA on the oracle operator (host cores): same iterations, x to 1e-9,
    the costs to 1e-9 relative."""
    from commit_amd import solvers
    A, O, x, y, Ax_o, Aty_o = brain
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    rp = _nn_rp(A.DICTIONARY, A.KERNELS)
    xg, dg = solvers.solve(brain_signal, A, A.T, tol_fun=0.0, tol_x=0.0, max_iter=10, verbose=0,
                           regularisation=solvers.init_regularisation(dict(rp)))
    xo, do = oracle.solve(brain_signal, O, O.T, tol_fun=0.0, tol_x=0.0, max_iter=10,
                          regularisation=oracle.init_regularisation(dict(rp)))
    _compare_fits(dg, do, xg, xo, 1e-9)
    for k in ("residual", "cost_function", "abs_cost"):
        assert dg[k] == pytest.approx(do[k], rel=1e-9), k


def test_fullsize_fp32_group_lasso_fit_vs_oracle(brain, brain_signal):
    """The mixed-precision solve (float32 vectors, FP64 step constant, stops and
    final residual) with group lasso over 1,000 random bundles and
    non-negativity, 10 iterations, against the oracle's FP64 fit: x to 1e-5."""
    from commit_amd import solvers
    A, O, x, y, Ax_o, Aty_o = brain
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    D, K = A.DICTIONARY, A.KERNELS
    nF = D["IC"]["nSTR"]
    rng = np.random.default_rng(13)
    lab = rng.integers(0, 1000, nF)
    order = np.argsort(lab, kind="stable")
    groups = np.empty(1000, dtype=object)
    for k, g in enumerate(np.split(order, np.cumsum(np.bincount(lab, minlength=1000))[:-1])):
        groups[k] = g.astype(np.int32)
    gw = np.sqrt([g.size for g in groups]).astype(np.float64)
    aty = O.T.dot(brain_signal)
    lam = 0.05 * oracle.lambda_max_group(aty, gw, groups)
    rp = dict(_nn_rp(D, K), regIC="group_lasso", lambdaIC=lam,
              dictIC_params={"group_idx_kept": groups, "group_weights": gw})
    xo, do = oracle.solve(brain_signal, O, O.T, tol_fun=0.0, tol_x=0.0, max_iter=10,
                          regularisation=oracle.init_regularisation(dict(rp)))
    A.set_precision(_lib.FP32)
    try:
        xg, dg = solvers.solve(brain_signal, A, A.T, tol_fun=0.0, tol_x=0.0, max_iter=10, verbose=0,
                               regularisation=solvers.init_regularisation(dict(rp), n=A.shape[1], device=A.device))
    finally:
        A.set_precision(_lib.FP64)
    _compare_fits(dg, do, xg, xo, 1e-5)
    assert dg["cost_function"] == pytest.approx(do["cost_function"], rel=1e-6)
