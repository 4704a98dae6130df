"""Batched fits on the GPU: several tractograms of uneven size stacked into one
block-diagonal plan, each with a lambda sweep of group-lasso fits, advanced
together by the device-side FISTA control (commit_amd.batched).

Bars: every (tractogram, lambda) fit runs the reference's iteration count and
its x agrees with the oracle's FP64 fit of that tractogram alone to 1e-5
relative (the batched products are float32 matrix-core products with FP32
accumulation per voxel); the graph-replayed run equals the eager run bitwise.
Both are checked on small tractograms and on the full batch of eight
tractograms of 50k-400k fibres, nS = 288, 16 lambdas each.
"""
import numpy as np
import pytest
import torch

from _problems import rel
from commit_amd import batched, synth
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL_BATCHED = 1e-5
SIZES = (600, 1200, 2000)
LAMS = (1e-4, 1e-3, 1e-2, 5e-2)


def _groups(nF, ng, seed):
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, ng, nF)
    out = [np.nonzero(lab == g)[0].astype(np.int32) for g in range(ng)]
    return [g for g in out if g.size]


@pytest.fixture(scope="module")
def stack():
    probs = []
    for k, nF in enumerate(SIZES):
        D, K, _ = synth.make_problem("small", seed=1100 + k, nF=nF)
        R = oracle.OracleOperator(D, K, oracle.set_threads(D, 2))
        y = synth.rician(R.dot(synth.x_true(R.shape[1], 1200 + k)), 30.0, 1300 + k)
        probs.append((D, K, R, y, _groups(nF, 8, 1400 + k)))
    K = probs[0][1]
    C = batched.Combined([p[0] for p in probs], K)
    y = np.zeros(C.m)
    groups, gw = [], []
    for s, p in enumerate(probs):
        y[C.rows(s)] = p[3]
        for g in p[4]:
            groups.append((g + C.fo[s]).astype(np.int32))
            gw.append(np.sqrt(g.size))
    gidx = np.empty(len(groups), dtype=object)
    for i, g in enumerate(groups):
        gidx[i] = g
    return probs, C, y, gidx, np.asarray(gw)


def _rp(nIC, F, nEC, E, nISO, I, gidx, gw, lam=0.0):
    rp = {"startIC": 0, "sizeIC": nIC * F, "startEC": nIC * F, "sizeEC": nEC * E,
          "startISO": nIC * F + nEC * E, "sizeISO": nISO * I, "regIC": "group_lasso", "regEC": None,
          "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True, "lambdaIC": lam,
          "dictIC_params": {"group_idx_kept": gidx, "group_weights": gw}}
    return rp


def _batched(stack, iters, use_graph):
    probs, C, y, gidx, gw = stack
    A = C.operator()
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    lam = np.tile(np.asarray(LAMS), (C.S, 1))
    bf = batched.BatchedFista(A, y, rp, lam, len(LAMS), layout=C, tol_fun=0.0, tol_x=0.0, max_iter=iters)
    bf.initialise()
    steps = bf.run(use_graph=use_graph)
    return bf, steps


def test_batched_group_lasso_sweep_vs_oracle(stack):
    probs, C, y, gidx, gw = stack
    iters = 40
    bf, steps = _batched(stack, iters, use_graph=True)
    X = bf.solution()
    det = bf.details()
    worst = 0.0
    for s, (D, K, R, ys, groups) in enumerate(probs):
        ga = np.empty(len(groups), dtype=object)
        for i, g in enumerate(groups):
            ga[i] = g
        gws = np.sqrt([g.size for g in groups])
        nF, nE, nI = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
        cols = C.columns(s)
        for r, lam in enumerate(LAMS):
            rp = _rp(K["wmr"].shape[0], nF, K["wmh"].shape[0], nE, K["iso"].shape[0], nI, ga, gws, lam)
            xo, do = oracle.solve(ys, R, R.T, tol_fun=0.0, tol_x=0.0, max_iter=iters,
                                  regularisation=oracle.init_regularisation(rp))
            e = rel(X[cols, r], xo)
            worst = max(worst, e)
            assert det[s][r]["iterations"] == do["iterations"], (s, r, det[s][r], do)
            assert e <= TOL_BATCHED, (s, r, e)
    print(f"batched {len(SIZES)} tractograms x {len(LAMS)} lambdas, {steps} trial steps: worst rel x {worst:.2e}",
          flush=True)


def test_graph_replay_equals_eager(stack):
    a, sa = _batched(stack, 15, use_graph=True)
    b, sb = _batched(stack, 15, use_graph=False)
    assert torch.equal(a.x, b.x)
    da, db = a.details(), b.details()
    for s in range(len(da)):
        for r in range(len(da[s])):
            assert da[s][r] == db[s][r]


@pytest.mark.parametrize("waves", [1, 2])
@pytest.mark.parametrize("R", [4, 16])
def test_fused_residual_equals_two_calls(stack, waves, R):
    """commit_hip_A_multi_residual (A X and the residual in one pass) against
    commit_hip_A_multi followed by commit_hip_multi_residual: res bitwise
    equal, the per-fit squared norms to 1e-13 (FP64, another summation order);
    y shared or per column, weighted or not, norms only, y = 0."""
    import ctypes
    from commit_amd import _lib, autotune
    probs, C, y, gidx, gw = stack
    A = C.operator()
    autotune.set_option(A, _lib.OPT_MULTI_A_WAVES, waves)
    lib = _lib.load()
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    lam = np.tile(np.linspace(1e-4, 1e-2, R), (C.S, 1))
    rng = np.random.default_rng(10 * waves + R)
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    X = torch.from_numpy(rng.standard_normal((C.n, R)).astype(np.float32)).cuda()
    AX = A.matmat(X)
    for ypc, weighted in ((0, False), (0, True), (1, True)):
        yv = y if ypc == 0 else rng.random((C.m, R))
        w = rng.uniform(0.5, 1.5, C.m) if weighted else None
        bf = batched.BatchedFista(A, yv, rp, lam, R, layout=C, sqrt_W=w, max_iter=2)
        nq = C.S * R
        for ydev in (bf.y, None):
            two = torch.empty((C.m, R), dtype=torch.float32, device="cuda")
            rsq2 = torch.empty(nq, dtype=torch.float64, device="cuda")
            work = torch.empty(bf.L.nypieces * R + 1, dtype=torch.float64, device="cuda")
            yz = ydev if ydev is not None else torch.zeros(C.m, dtype=torch.float32, device="cuda")
            _lib.check(lib.commit_hip_multi_residual(ctypes.byref(bf.L), P(AX), P(yz), ypc if ydev is not None else 0,
                                                     P(bf.sw), P(two), P(rsq2), P(work), st), "multi_residual")
            one = torch.full((C.m, R), 7.0, dtype=torch.float32, device="cuda")
            rsq1 = torch.empty(nq, dtype=torch.float64, device="cuda")
            bf._resid(X, ydev, ypc if ydev is not None else 0, one, rsq1)
            assert torch.equal(one, two), (ypc, weighted)
            a, b = rsq1.cpu().numpy(), rsq2.cpu().numpy()
            assert np.max(np.abs(a - b) / b) <= 1e-13, (ypc, weighted, np.max(np.abs(a - b) / b))
            rsq0 = torch.empty(nq, dtype=torch.float64, device="cuda")
            bf._resid(X, ydev, ypc if ydev is not None else 0, None, rsq0)      # norms only
            assert torch.equal(rsq0, rsq1)
            r64 = two.double().cpu().numpy()
            for s in range(C.S):
                rows = C.rows(s)
                ref = (r64[rows] ** 2).sum(axis=0)
                assert np.allclose(a[s * R:(s + 1) * R], ref, rtol=1e-6)


def _set_cols(bf, adv, done, coef):
    cols = bf._download_cols()
    for q in range(bf.S * bf.R):
        cols[q].advance, cols[q].done, cols[q].coef = int(adv[q]), int(done[q]), float(coef[q])
    bf._cols_host = cols
    bf._upload_cols()


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("sel", [0, 1])
def test_momentum_At_vs_At_of_combined_residual(stack, weighted, sel):
    """commit_hip_At_multi_momentum against the plain 16-RHS A^T of the
    residual combined on the host: rhat = (1 + coef) r[1 - sel] - coef r[sel]
    (FP64, rounded to float32) in the columns of fits that advanced.  Those
    gradient columns agree to FP32 rounding (1e-6 of the column norm) and their
    ||rhat||^2 to 1e-12; the other fits' gradient columns keep their values."""
    import ctypes
    from commit_amd import _lib
    probs, C, y, gidx, gw = stack
    A = C.operator()
    lib = _lib.load()
    R = 16
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    lam = np.tile(np.logspace(-4, -2, R), (C.S, 1))
    rng = np.random.default_rng(40 + 2 * sel + weighted)
    w = rng.uniform(0.5, 1.5, C.m) if weighted else None
    bf = batched.BatchedFista(A, y, rp, lam, R, layout=C, sqrt_W=w, max_iter=2)
    nq = C.S * R
    adv = rng.random(nq) < 0.6
    adv[0], adv[1] = True, False
    coef = rng.uniform(0.0, 0.95, nq)
    _set_cols(bf, adv, np.zeros(nq, bool), coef)
    r = [rng.standard_normal((C.m, R)).astype(np.float32) for _ in range(2)]
    for k in range(2):
        bf.r[k].copy_(torch.from_numpy(r[k]))
    bf.sel.fill_(sel)
    g0 = torch.from_numpy(rng.standard_normal((C.n, R)).astype(np.float32)).cuda()
    bf.grad.copy_(g0)
    rsq = torch.zeros(nq, dtype=torch.float64, device="cuda")
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.commit_hip_At_multi_momentum(A._plan.handle, ctypes.byref(bf.L), P(bf.r[0]), P(bf.r[1]),
                                                P(bf.sel), P(bf.sw), P(bf.col), P(bf.grad), P(rsq), P(bf.rwork),
                                                st), "At_multi_momentum")
    got = bf.grad.cpu().numpy()
    rx, rpv = r[1 - sel].astype(np.float64), r[sel].astype(np.float64)
    yh = np.zeros((C.m, R), np.float32)
    rsq_ref = np.zeros(nq)
    for s_ in range(C.S):
        rows = C.rows(s_)
        for c in range(R):
            q = s_ * R + c
            if adv[q]:
                u = (1.0 + coef[q]) * rx[rows, c] - coef[q] * rpv[rows, c]
                yh[rows, c] = u.astype(np.float32)
                rsq_ref[q] = np.sum(u * u)
    ref = A.T.matmat(torch.from_numpy(yh).cuda(), sqrtw=bf.sw).cpu().numpy()
    g0n = g0.cpu().numpy()
    rs = rsq.cpu().numpy()
    for s_ in range(C.S):
        cols = C.columns(s_)
        for c in range(R):
            q = s_ * R + c
            if adv[q]:
                e = np.linalg.norm(got[cols, c] - ref[cols, c]) / np.linalg.norm(ref[cols, c])
                assert e <= 1e-6, (s_, c, e)
                assert abs(rs[q] - rsq_ref[q]) <= 1e-12 * rsq_ref[q], (s_, c, rs[q], rsq_ref[q])
            else:
                assert np.array_equal(got[cols, c], g0n[cols, c]), (s_, c)


@pytest.mark.parametrize("case", ["all", "mixed", "backtrack", "finished"])
def test_residual_roles(stack, case):
    """commit_hip_fista_multi_roles: afterwards r(prev_x) of every unfinished
    fit is in r[sel].  all: every unfinished fit advanced, the roles swap;
    mixed: the advanced fits' columns are copied into r[sel]; backtrack: no fit
    advanced, nothing moves; finished: only finished fits, no buffer changes."""
    import ctypes
    from commit_amd import _lib
    probs, C, y, gidx, gw = stack
    A = C.operator()
    lib = _lib.load()
    R = 4
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    lam = np.tile(np.asarray(LAMS), (C.S, 1))
    bf = batched.BatchedFista(A, y, rp, lam, R, layout=C, max_iter=2)
    nq = C.S * R
    rng = np.random.default_rng(77)
    done = np.zeros(nq, bool)
    if case == "all":
        adv = np.ones(nq, bool)
        adv[3], done[3] = False, True
    elif case == "mixed":
        adv = rng.random(nq) < 0.5
        adv[0], adv[1] = True, False
        done[2], adv[2] = True, False
    elif case == "backtrack":
        adv = np.zeros(nq, bool)
    else:
        adv = np.zeros(nq, bool)
        done[:] = True
    _set_cols(bf, adv, done, np.zeros(nq))
    r = [rng.standard_normal((C.m, R)).astype(np.float32) for _ in range(2)]
    for k in range(2):
        bf.r[k].copy_(torch.from_numpy(r[k]))
    sel0 = 1
    bf.sel.fill_(sel0)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.commit_hip_fista_multi_roles(A._plan.handle, ctypes.byref(bf.L), P(bf.r[0]), P(bf.r[1]),
                                                P(bf.sel), P(bf.col), st), "fista_multi_roles")
    sel1 = int(bf.sel.item())
    out = [bf.r[k].cpu().numpy() for k in range(2)]
    if case in ("all", "finished"):
        assert sel1 == 1 - sel0
        assert np.array_equal(out[0], r[0]) and np.array_equal(out[1], r[1])
        return
    assert sel1 == sel0
    assert np.array_equal(out[1 - sel0], r[1 - sel0])
    exp = r[sel0].copy()
    if case == "mixed":
        for s_ in range(C.S):
            rows = C.rows(s_)
            for c in range(R):
                if adv[s_ * R + c]:
                    exp[rows, c] = r[1 - sel0][rows, c]
    assert np.array_equal(out[sel0], exp)


def test_combined_residual_equals_product_path(stack):
    """The fits with the momentum residual formed from r(x) and r(prev_x)
    against the fits that recompute A xhat: equal iteration counts and stopping
    tests, x to 1e-5, with early stops so that fits finish at different steps."""
    probs, C, y, gidx, gw = stack
    A = C.operator()
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    lam = np.tile(np.asarray(LAMS), (C.S, 1))
    out = []
    for combine in (True, False):
        bf = batched.BatchedFista(A, y, rp, lam, len(LAMS), layout=C, tol_fun=1e-3, tol_x=1e-4, max_iter=60,
                                  combine=combine)
        bf.initialise()
        bf.run(use_graph=True)
        out.append((bf.solution(), bf.details()))
    (xa, da), (xb, db) = out
    its = set()
    for s_ in range(C.S):
        cols = C.columns(s_)
        for r in range(len(LAMS)):
            assert da[s_][r]["iterations"] == db[s_][r]["iterations"], (s_, r)
            assert da[s_][r]["stopping_criterion"] == db[s_][r]["stopping_criterion"], (s_, r)
            assert rel(xa[cols, r], xb[cols, r]) <= 1e-5, (s_, r)
            its.add(da[s_][r]["iterations"])
    assert len(its) > 1, its                                   # fits stopped at different iterations


def test_batched_early_stops_vs_oracle(stack):
    """Fits with the reference's tolerances stop at different iterations; each
    stops at the oracle's iteration with the oracle's criterion, x to 1e-5."""
    probs, C, y, gidx, gw = stack
    A = C.operator()
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    lam = np.tile(np.asarray(LAMS), (C.S, 1))
    bf = batched.BatchedFista(A, y, rp, lam, len(LAMS), layout=C, tol_fun=1e-3, tol_x=1e-4, max_iter=60)
    bf.initialise()
    bf.run(use_graph=True)
    X, det = bf.solution(), bf.details()
    for s_, (D, K, R, ys, groups) in enumerate(probs):
        ga = np.empty(len(groups), dtype=object)
        for i, g in enumerate(groups):
            ga[i] = g
        nF, nE, nI = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
        for r, lamv in enumerate(LAMS):
            rpo = _rp(K["wmr"].shape[0], nF, K["wmh"].shape[0], nE, K["iso"].shape[0], nI, ga,
                      np.sqrt([g.size for g in groups]), lamv)
            xo, do = oracle.solve(ys, R, R.T, tol_fun=1e-3, tol_x=1e-4, max_iter=60,
                                  regularisation=oracle.init_regularisation(rpo))
            assert det[s_][r]["iterations"] == do["iterations"], (s_, r, det[s_][r], do)
            assert det[s_][r]["stopping_criterion"] == do["stopping_criterion"], (s_, r)
            assert rel(X[C.columns(s_), r], xo) <= TOL_BATCHED, (s_, r)


# ------------------------------------ 8 tractograms of 50k-400k, 16 lambdas --
SIZES8 = tuple(50_000 * (k + 1) for k in range(8))


@pytest.fixture(scope="module")
def stack8():
    """Eight tractograms of 50k..400k fibres on the 2.5 mm head (nS = 288),
    stacked block-diagonally; y = A x_true + Rician noise (SNR 30) rounded to
    float32, the input of both the batched fits and the oracle; 64 random
    bundles per tractogram; 16 lambdas log-spaced 1e-4..1e-1 x lambda_max."""
    from commit_amd import _lib
    probs = [synth.make_problem("brain25", seed=2000 + k, nF=nf) for k, nf in enumerate(SIZES8)]
    K = probs[0][1]
    C = batched.Combined([p[0] for p in probs], K)
    A = C.operator(precision=_lib.FP64)
    xt = torch.from_numpy(synth.x_true(C.n, 3001)).cuda()
    sgn = A.matvec(xt).cpu().numpy()
    y = synth.rician(sgn, 30.0, 3002).astype(np.float32).astype(np.float64)
    aty = A.T.dot(y)
    rng = np.random.default_rng(3000)
    groups, gw, lam_max, per = [], [], np.zeros(C.S), []
    for s_ in range(C.S):
        lab = rng.integers(0, 64, C.nF[s_])
        mine = []
        for b in range(64):
            gi = np.nonzero(lab == b)[0].astype(np.int32)
            if gi.size:
                mine.append(gi)
                groups.append(gi + C.fo[s_])
                gw.append(np.sqrt(gi.size))
                lam_max[s_] = max(lam_max[s_], np.linalg.norm(aty[gi + C.fo[s_]]) / gw[-1])
        per.append(mine)
    gidx = np.empty(len(groups), dtype=object)
    for k, gi in enumerate(groups):
        gidx[k] = gi
    lam = lam_max[:, None] * np.logspace(-4, -1, 16)[None, :]
    print(f"stack8: {C.F} fibres, {C.D['IC']['n']} segments, {C.m} rows", flush=True)
    return probs, C, y, gidx, np.asarray(gw), lam, per


def _batched8(stack8, iters, use_graph):
    probs, C, y, gidx, gw, lam, per = stack8
    A = C.operator()
    rp = _rp(C.nIC, C.F, C.nEC, C.E, C.nISO, C.I, gidx, gw)
    bf = batched.BatchedFista(A, y, rp, lam, 16, layout=C, tol_fun=0.0, tol_x=0.0, max_iter=iters)
    bf.initialise()
    steps = bf.run(use_graph=use_graph)
    return bf, steps


def test_batched_8x16_graph_replay_equals_eager(stack8):
    a, sa = _batched8(stack8, 6, use_graph=True)
    b, sb = _batched8(stack8, 6, use_graph=False)
    assert sa == sb and torch.equal(a.x, b.x)
    da, db = a.details(), b.details()
    assert all(da[s][r] == db[s][r] for s in range(len(da)) for r in range(16))


ITERS8 = 10


@pytest.fixture(scope="module")
def fit8(stack8):
    bf, steps = _batched8(stack8, ITERS8, use_graph=True)
    return bf.solution(), bf.details()


@pytest.mark.parametrize("s, r", [(0, 8), (7, 8), (7, 15)])
def test_batched_8x16_vs_oracle(stack8, fit8, s, r):
    """One lambda of the 50k and the 400k tractogram against the oracle's FP64
    fit of that tractogram alone, 10 iterations."""
    probs, C, y, gidx, gw, lam, per = stack8
    X, details = fit8
    iters = ITERS8
    D, K = probs[s][0], probs[s][1]
    R = oracle.OracleOperator(D, K, oracle.set_threads(D, 16))
    ga = np.empty(len(per[s]), dtype=object)
    for k, g in enumerate(per[s]):
        ga[k] = g
    rp = _rp(K["wmr"].shape[0], D["IC"]["nSTR"], K["wmh"].shape[0], D["EC"]["n"], K["iso"].shape[0],
             D["ISO"]["n"], ga, np.sqrt([g.size for g in per[s]]), float(lam[s, r]))
    xo, do = oracle.solve(y[C.rows(s)], R, R.T, tol_fun=0.0, tol_x=0.0, max_iter=iters,
                          regularisation=oracle.init_regularisation(rp))
    det = details[s][r]
    e = rel(X[C.columns(s), r], xo)
    print(f"stack8 tractogram {s} ({D['IC']['nSTR']} fibres) lambda {r}: iterations {det['iterations']}/"
          f"{do['iterations']}, rel x {e:.2e}", flush=True)
    assert det["iterations"] == do["iterations"]
    assert e <= TOL_BATCHED
