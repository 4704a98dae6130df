"""FP32 storage mode: float32 vectors with FP64 accumulation in the products,
FP32 FISTA vectors with FP64 stopping norms, Lipschitz estimate and final
residual.  Checked against the FP64 oracle.

Bars: single products <= 1e-6 relative (float32 rounding of inputs and
outputs only; the oracle is fed the same float32-rounded inputs); full fits
<= 1e-5 relative error in x against the oracle's FP64 fit with the same
number of iterations, and per-fibre error reported.
"""
import numpy as np
import pytest
import torch

from _problems import random_dictionary, rel
from commit_amd import _lib, solvers, synth
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL_PRODUCT = 1e-6
TOL_FIT = 1e-5


@pytest.mark.parametrize("nS", [1, 60, 65, 288])
@pytest.mark.parametrize("rf", [(1, 1, 2), (2, 0, 1)])
def test_fp32_products_vs_dense(nS, rf):
    nIC, nEC, nISO = rf
    D, K = random_dictionary(91 + nS, nVOX=9, nF=13, n=120, nS=nS, ndirs=5, nIC=nIC, nEC=nEC, nISO=nISO,
                             empty_voxels=(2,), ec_per_voxel=2)
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    rng = np.random.default_rng(nS)
    x = rng.standard_normal(A.shape[1]).astype(np.float32)
    y = rng.standard_normal(A.shape[0]).astype(np.float32)
    dense = oracle.assemble_A(D, K)
    ya = A.matvec(torch.from_numpy(x).cuda()).double().cpu().numpy()
    xa = A.T.matvec(torch.from_numpy(y).cuda()).double().cpu().numpy()
    ey = rel(ya, dense @ x.astype(np.float64))
    ex = rel(xa, dense.T @ y.astype(np.float64))
    print(f"fp32 nS={nS} rf={rf}: A x {ey:.2e}  A^T y {ex:.2e}", flush=True)
    assert ey <= TOL_PRODUCT and ex <= TOL_PRODUCT


def test_precision_switch_shares_plan():
    D, K = random_dictionary(5, nVOX=10, nF=20, n=200, nS=60, ndirs=6, nIC=1, nEC=1, nISO=2)
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    rng = np.random.default_rng(0)
    x = rng.random(A.shape[1])
    y32 = A.dot(x)                                     # numpy float64 in and out, float32 on the device
    A.set_precision(_lib.FP64)
    assert A.T.precision == _lib.FP64                  # .T shares the plan
    y64 = A.dot(x)
    yo = oracle.assemble_A(D, K) @ x
    assert rel(y64, yo) <= 1e-12
    assert rel(y32, yo) <= TOL_PRODUCT
    with pytest.raises(RuntimeError):
        A.matvec(torch.zeros(A.shape[1], dtype=torch.float32, device="cuda"))


@pytest.fixture(scope="module")
def problem():
    D, K, _ = synth.make_problem("small", seed=1001, nF=1500)
    T = oracle.set_threads(D, 4)
    R = oracle.OracleOperator(D, K, T)
    xt = synth.x_true(R.shape[1], 2001)
    y = synth.rician(R.dot(xt), 30.0, 3001)
    return D, K, T, R, y


def _rp(D, K, regIC=None, **extra):
    nF, nE, nI = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
    rp = {"startIC": 0, "sizeIC": nF * K["wmr"].shape[0]}
    rp["startEC"] = rp["sizeIC"]
    rp["sizeEC"] = nE * K["wmh"].shape[0]
    rp["startISO"] = rp["sizeIC"] + rp["sizeEC"]
    rp["sizeISO"] = nI * K["iso"].shape[0]
    rp.update(regIC=regIC, regEC=None, regISO=None, nnIC=True, nnEC=True, nnISO=True)
    rp.update(extra)
    return rp


def _fit32(problem, rp, iters):
    D, K, T, R, y = problem
    A = LinearOperator(D, K, T, precision=_lib.FP32)
    xg, dg = solvers.solve(y, A, A.T, tol_fun=0.0, tol_x=0.0, max_iter=iters, verbose=0,
                           regularisation=solvers.init_regularisation(dict(rp)))
    xo, do = oracle.solve(y, R, R.T, tol_fun=0.0, tol_x=0.0, max_iter=iters,
                          regularisation=oracle.init_regularisation(dict(rp)))
    e = rel(xg, xo)
    nF = D["IC"]["nSTR"]
    fib = np.abs(xg[:nF] - xo[:nF]) / (np.abs(xo[:nF]).max() + 1e-300)
    print(f"fp32 fit ({iters} it): rel x {e:.2e}, per-fibre |dx|/max|x| max {fib.max():.2e} "
          f"p99 {np.percentile(fib, 99):.2e}; cost {dg['cost_function']:.9e} vs {do['cost_function']:.9e}",
          flush=True)
    return xg, dg, xo, do, e


def test_fp32_nonnegative_fit_vs_fp64_oracle(problem):
    D, K = problem[0], problem[1]
    xg, dg, xo, do, e = _fit32(problem, _rp(D, K), 100)
    assert dg["iterations"] == do["iterations"] == 100
    assert dg["precision"] == "fp32"
    assert e <= TOL_FIT
    assert abs(dg["cost_function"] - do["cost_function"]) <= 1e-5 * abs(do["cost_function"])


def test_fp32_group_lasso_fit_vs_fp64_oracle(problem):
    D, K = problem[0], problem[1]
    nF = D["IC"]["nSTR"]
    rng = np.random.default_rng(4)
    lab = rng.integers(0, 20, nF)
    groups = np.empty(20, dtype=object)
    for g in range(20):
        groups[g] = np.nonzero(lab == g)[0].astype(np.int32)
    gw = np.sqrt([g.size for g in groups]).astype(np.float64)
    rp = _rp(D, K, regIC="group_lasso", lambdaIC=0.02,
             dictIC_params={"group_idx_kept": groups, "group_weights": gw})
    xg, dg, xo, do, e = _fit32(problem, rp, 60)
    assert e <= TOL_FIT


def test_power_iteration_lipschitz():
    """FP64 power iterations reach the largest eigenvalue of A^T W A (dense check)."""
    D, K = random_dictionary(8, nVOX=12, nF=15, n=150, nS=20, ndirs=6, nIC=1, nEC=1, nISO=1)
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    dense = oracle.assemble_A(D, K)
    dense = dense.toarray() if hasattr(dense, "toarray") else np.asarray(dense)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.5, 2.0, A.shape[0])
    M = dense.T @ (w[:, None] * dense)
    lam_max = np.linalg.eigvalsh(M)[-1]
    g = torch.from_numpy(rng.random(A.shape[1]).astype(np.float32)).cuda()
    sw = torch.from_numpy(np.sqrt(w).astype(np.float32)).cuda()
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    import ctypes
    L = solvers._lipschitz_power(A, lib, A._plan.handle, sw, g, 500, ctypes.c_void_p(st))
    assert A.precision == _lib.FP32                   # restored after the FP64 iterations
    assert abs(L - lam_max) <= 1e-6 * lam_max, (L, lam_max)


@pytest.mark.parametrize("mode", [0, 2])
def test_fp32_rows_unaligned(mode):
    """The float kernels move a lane's quad of y, sqrtw or residual samples as
    one 16-byte access only when the row is 16-byte aligned (nS a multiple of
    4 and the vector aligned); vectors at a 4-byte offset take the per-sample
    path.  Both give the same residual, norm and A^T, bitwise, on the
    per-voxel (mode 0) and the slab (mode 2) kernels."""
    import ctypes
    from commit_amd import autotune
    D, K = random_dictionary(12 + mode, nVOX=40, nF=30, n=600, nS=60, ndirs=8, nIC=1, nEC=1, nISO=2)
    A = LinearOperator(D, K, None, precision=_lib.FP32)
    autotune.set_option(A, _lib.OPT_SLAB, mode)
    m, n = A.shape
    rng = np.random.default_rng(mode)
    x = torch.from_numpy(rng.random(n).astype(np.float32)).cuda()
    vals = [rng.standard_normal(m).astype(np.float32), rng.uniform(0.5, 1.5, m).astype(np.float32)]
    lib, h = _lib.load(), A._plan.handle
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    out = []
    for off in (0, 1):
        bufs = [torch.zeros(m + 4, dtype=torch.float32, device="cuda") for _ in range(3)]
        y, w, r = (b[off:off + m] for b in bufs)
        y.copy_(torch.from_numpy(vals[0]))
        w.copy_(torch.from_numpy(vals[1]))
        assert (y.data_ptr() % 16 == 0) == (off == 0)
        rsq = torch.empty(1, dtype=torch.float64, device="cuda")
        g = torch.empty(n, dtype=torch.float32, device="cuda")
        _lib.check(lib.commit_hip_A_residual(h, P(x), P(y), P(w), P(r), P(rsq), st), "A_residual")
        _lib.check(lib.commit_hip_At_weighted(h, P(r), P(w), P(g), st), "At_weighted")
        torch.cuda.synchronize()
        out.append((r.cpu().numpy().copy(), rsq.item(), g.cpu().numpy().copy()))
    (r0, s0, g0), (r1, s1, g1) = out
    print(f"mode={mode}: residual max diff {np.abs(r0 - r1).max():.1e}, norm {s0:.17g} {s1:.17g}", flush=True)
    assert np.array_equal(r0, r1) and s0 == s1 and np.array_equal(g0, g1)
