"""GPU FISTA and proximal operators against the oracle's restatement of
solvers.py / proximals.pyx.

Bars: proximal operators and the regulariser value to 1e-14 relative (they
are elementwise or per-group, only summation order differs); full FISTA fits
run the same number of iterations with the same stopping criterion, and
x agrees to 1e-9 relative, the fit_details to the reference test's absolute
tolerance of 1e-4 (tests/test_demo.py:130-136).
"""
import numpy as np
import pytest
import torch

from _problems import rel
from commit_amd import solvers, synth
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu


def _rp(D, K, regIC=None, regEC=None, regISO=None, nn=(True, True, True), **extra):
    nF, nE, nI = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
    rp = {"startIC": 0, "sizeIC": nF * K["wmr"].shape[0]}
    rp["startEC"] = rp["sizeIC"]
    rp["sizeEC"] = nE * K["wmh"].shape[0]
    rp["startISO"] = rp["sizeIC"] + rp["sizeEC"]
    rp["sizeISO"] = nI * K["iso"].shape[0]
    rp.update(regIC=regIC, regEC=regEC, regISO=regISO, nnIC=nn[0], nnEC=nn[1], nnISO=nn[2])
    rp.update(extra)
    return rp


def _groups(nF, ng, seed):
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, ng, nF)
    groups = np.empty(ng, dtype=object)
    for g in range(ng):
        groups[g] = np.nonzero(lab == g)[0].astype(np.int32)
    keep = [g for g in range(ng) if groups[g].size]
    out = np.empty(len(keep), dtype=object)
    for i, g in enumerate(keep):
        out[i] = groups[g]
    return out


@pytest.fixture(scope="module")
def problem():
    D, K, _ = synth.make_problem("small", seed=1001, nF=1500)
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T)
    R = oracle.OracleOperator(D, K, T)
    xt = synth.x_true(A.shape[1], 2001)
    y = synth.rician(R.dot(xt), 30.0, 3001)
    return D, K, T, A, R, y


def _fit_both(problem, rp, max_iter=60, tol_fun=1e-12, tol_x=1e-12):
    D, K, T, A, R, y = problem
    reg_g = solvers.init_regularisation(dict(rp))
    reg_o = oracle.init_regularisation(dict(rp))
    xg, dg = solvers.solve(y, A, A.T, tol_fun=tol_fun, tol_x=tol_x, max_iter=max_iter, verbose=0,
                           regularisation=reg_g)
    xo, do = oracle.solve(y, R, R.T, tol_fun=tol_fun, tol_x=tol_x, max_iter=max_iter, regularisation=reg_o)
    print(f"fit: iters {dg['iterations']}/{do['iterations']}  rel x {rel(xg, xo):.2e}  "
          f"cost {dg['cost_function']:.9e}/{do['cost_function']:.9e}", flush=True)
    return xg, dg, xo, do


def _check(xg, dg, xo, do, tol_x=1e-9):
    assert dg["iterations"] == do["iterations"]
    assert dg["stopping_criterion"] == do["stopping_criterion"]
    assert rel(xg, xo) <= tol_x
    for k in ("residual", "regterm", "cost_function", "abs_cost", "rel_cost", "abs_x", "rel_x"):
        assert abs(dg[k] - do[k]) < 1e-4, k


def test_nonnegative_fit(problem):
    D, K = problem[0], problem[1]
    _check(*_fit_both(problem, _rp(D, K)))


def test_unconstrained_fit_converges_like_oracle(problem):
    D, K = problem[0], problem[1]
    _check(*_fit_both(problem, _rp(D, K, nn=(False, False, False)), max_iter=40))


def test_lasso_all_compartments(problem):
    D, K, T, A, R, y = problem
    rp = _rp(D, K, "lasso", "lasso", "lasso", lambdaIC=2e-3, lambdaEC=1e-3, lambdaISO=5e-4)
    _check(*_fit_both(problem, rp))


def test_weighted_lasso_quirks(problem):
    """Weighted IC and ISO lasso, including the reference's x[startISO:sizeISO]
    omega slice (solvers.py:183)."""
    D, K, T, A, R, y = problem
    rng = np.random.default_rng(4)
    rp = _rp(D, K, "lasso", None, "lasso", lambdaIC=2e-3, lambdaISO=5e-4,
             dictIC_params={"coeff_weights_kept": rng.uniform(0.5, 2, D["IC"]["nSTR"])},
             dictISO_params={"coeff_weights": rng.uniform(0.5, 2, 2 * D["ISO"]["n"])})
    _check(*_fit_both(problem, rp))


@pytest.mark.parametrize("kind", ["group_lasso", "sparse_group_lasso"])
def test_group_penalties(problem, kind):
    D, K, T, A, R, y = problem
    groups = _groups(D["IC"]["nSTR"], 64, 5)
    gw = np.sqrt([g.size for g in groups]).astype(np.float64)
    lam = 5e-3 if kind == "group_lasso" else (1e-3, 5e-3)
    rp = _rp(D, K, kind, None, None, lambdaIC=lam, dictIC_params={"group_idx_kept": groups, "group_weights": gw})
    _check(*_fit_both(problem, rp))


def test_overlapping_groups_prox_order():
    """Hierarchical groups are applied in the given order (proximals.pyx:116-134)."""
    rng = np.random.default_rng(8)
    n = 50
    groups = np.empty(3, dtype=object)
    groups[0] = np.arange(0, 10, dtype=np.int32)
    groups[1] = np.arange(0, 30, dtype=np.int32)      # contains group 0
    groups[2] = np.arange(25, 50, dtype=np.int32)     # overlaps group 1
    gw = np.array([1.0, 2.0, 1.5])
    rp = {"startIC": 0, "sizeIC": n, "startEC": n, "sizeEC": 0, "startISO": n, "sizeISO": 0,
          "regIC": "group_lasso", "regEC": None, "regISO": None, "nnIC": True, "nnEC": False, "nnISO": False,
          "lambdaIC": 0.3, "dictIC_params": {"group_idx_kept": groups, "group_weights": gw}}
    reg_g = solvers.init_regularisation(dict(rp))
    reg_o = oracle.init_regularisation(dict(rp))
    for _ in range(5):
        x = rng.standard_normal(n)
        a = reg_g["prox"](x.copy(), 0.7)
        b = reg_o["prox"](x.copy(), 0.7)
        assert rel(a, b) < 1e-14
        assert abs(reg_g["omega"](a) - reg_o["omega"](a)) <= 1e-13 * max(1.0, abs(reg_o["omega"](a)))


def test_prox_and_omega_match_oracle(problem):
    D, K, T, A, R, y = problem
    rng = np.random.default_rng(6)
    groups = _groups(D["IC"]["nSTR"], 40, 7)
    gw = rng.uniform(0.5, 2.0, len(groups))
    for kind, lam in (("lasso", 0.3), ("group_lasso", 0.2), ("sparse_group_lasso", (0.1, 0.2))):
        rp = _rp(D, K, kind, "lasso", "lasso", lambdaIC=lam, lambdaEC=0.05, lambdaISO=0.07,
                 dictIC_params={"group_idx_kept": groups, "group_weights": gw})
        reg_g = solvers.init_regularisation(dict(rp))
        reg_o = oracle.init_regularisation(dict(rp))
        x = rng.standard_normal(A.shape[1])
        a = reg_g["prox"](x.copy(), 0.9)
        b = reg_o["prox"](x.copy(), 0.9)
        assert rel(a, b) < 1e-14, kind
        assert abs(reg_g["omega"](x) - reg_o["omega"](x)) <= 1e-12 * abs(reg_o["omega"](x)), kind


def test_confidence_weights(problem):
    D, K, T, A, R, y = problem
    conf = np.random.default_rng(9).uniform(0.2, 1.0, y.size)
    rp = _rp(D, K)
    reg_g = solvers.init_regularisation(dict(rp))
    reg_o = oracle.init_regularisation(dict(rp))
    xg, dg = solvers.solve(y, A, A.T, tol_fun=1e-12, tol_x=1e-12, max_iter=40, verbose=0, regularisation=reg_g,
                           confidence_array=conf)
    xo, do = oracle.solve(y, R, R.T, tol_fun=1e-12, tol_x=1e-12, max_iter=40, regularisation=reg_o,
                          confidence_array=conf)
    _check(xg, dg, xo, do)


def test_fit_is_deterministic(problem):
    D, K, T, A, R, y = problem
    reg = solvers.init_regularisation(_rp(D, K))
    x1, d1 = solvers.solve(y, A, A.T, max_iter=30, verbose=0, regularisation=reg)
    x2, d2 = solvers.solve(y, A, A.T, max_iter=30, verbose=0, regularisation=reg)
    assert np.array_equal(x1, x2) and d1 == d2


@pytest.mark.parametrize("fp32", [False, True])
def test_residual_combine(problem, fp32):
    """commit_hip_residual_combine (the momentum point's residual from r(x) and
    r(prev_x)): r = a r1 + b r2 elementwise, ||r||^2 to 1e-13; and, the
    residual being affine in x, it equals sqrtw (A xhat - y) to the FP64 bar
    (1e-6 with float32 storage)."""
    import ctypes
    from commit_amd import _lib
    D, K, T, _, _, y = problem
    A = LinearOperator(D, K, T)
    if fp32:
        A.set_precision(_lib.FP32)
    dt = torch.float32 if fp32 else torch.float64
    lib = _lib.load()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(3)
    m, n = A.shape
    x, xp = rng.random(n), rng.random(n)
    w = rng.uniform(0.5, 1.5, m)
    mom = 0.61803
    xh = x + mom * (x - xp)
    dev = lambda a: torch.from_numpy(a).to(device="cuda", dtype=dt)
    yd, wd = dev(y), dev(w)
    r1, r2, rh, r = (torch.empty(m, dtype=dt, device="cuda") for _ in range(4))
    rsq = torch.empty(3, dtype=torch.float64, device="cuda")
    for v, out, k in ((x, r1, 0), (xp, r2, 1), (xh, rh, 2)):
        _lib.check(lib.commit_hip_A_residual(A._plan.handle, P(dev(v)), P(yd), P(wd), P(out), P(rsq[k:k + 1]), st),
                   "A_residual")
    _lib.check(lib.commit_hip_residual_combine(A._plan.handle, 1.0 + mom, P(r1), -mom, P(r2), P(r), P(rsq[2:3]), st),
               "residual_combine")
    r_np = r.double().cpu().numpy()
    expect = (1.0 + mom) * r1.double().cpu().numpy() - mom * r2.double().cpu().numpy()
    tol_e = 1e-6 if fp32 else 1e-15
    assert np.max(np.abs(r_np - expect)) <= tol_e * np.max(np.abs(expect))
    assert abs(rsq[2].item() - expect @ expect) <= (1e-6 if fp32 else 1e-13) * (expect @ expect)
    e = rel(r_np, rh.double().cpu().numpy())
    print(f"fp32={fp32}: combined vs direct residual {e:.2e}", flush=True)
    assert e <= (1e-6 if fp32 else 1e-12)


def test_c1_nonnegative_fit_100_iterations():
    """The CPU-sized workload (32x32x20 grid, 10,000 fibres, nS = 60, SZB):
    the 100-iteration FP64 non-negative fit against the oracle's, with the
    reference test's bar (tests/test_demo.py:120-141)."""
    D, K, _ = synth.make_problem("small", seed=1001)
    T = oracle.set_threads(D, 8)
    A = LinearOperator(D, K, T)
    R = oracle.OracleOperator(D, K, T)
    y = synth.rician(R.dot(synth.x_true(A.shape[1], 2001)), 30.0, 3001)
    rp = _rp(D, K)
    xg, dg = solvers.solve(y, A, A.T, tol_fun=0.0, tol_x=0.0, max_iter=100, verbose=0,
                           regularisation=solvers.init_regularisation(dict(rp)))
    xo, do = oracle.solve(y, R, R.T, tol_fun=0.0, tol_x=0.0, max_iter=100,
                          regularisation=oracle.init_regularisation(dict(rp)))
    print(f"C1 fit: {D['IC']['nSTR']} fibres, {D['IC']['n']} segments, iterations {dg['iterations']}, "
          f"rel x {rel(xg, xo):.2e}", flush=True)
    assert dg["iterations"] == 100
    _check(xg, dg, xo, do)
