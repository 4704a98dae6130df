"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement (a dropped
compartment, a wrong x/LUT offset, a transposed index, a lost eval factor, a
wrong prox branch) fails it:
  * the C operator against a dense matrix assembled independently from the
    closed-form entry formula (brute force), for several RF counts, sample
    counts and thread partitions;
  * adjointness <Ax,y> = <x,A^T y> on the small synthetic workload;
  * single-segment and all-ones-LUT (ScalarMap incidence) closed forms;
  * the LUT-free operator against its own closed form;
  * the proximal operators against brute-force minimisation;
  * FISTA against scipy.optimize.nnls, KKT conditions and the orthogonal case.
"""
import numpy as np
import pytest
from scipy.optimize import minimize, nnls

from _problems import random_dictionary, rel
from commit_amd import synth
from oracle import oracle


@pytest.mark.parametrize("threads", [1, 2, 3])
@pytest.mark.parametrize("shape", [(1, 1, 2, 7), (2, 0, 1, 5), (3, 2, 0, 1), (1, 1, 1, 64)])
def test_operator_vs_dense(threads, shape):
    nIC, nEC, nISO, nS = shape
    D, K = random_dictionary(threads * 10 + nS, nVOX=9, nF=13, n=60, nS=nS, ndirs=5, nIC=nIC, nEC=nEC,
                             nISO=nISO, empty_voxels=(4,), ec_per_voxel=2)
    T = oracle.set_threads(D, threads)
    R = oracle.OracleOperator(D, K, T)
    M = oracle.assemble_A(D, K)
    assert R.shape == M.shape
    rng = np.random.default_rng(threads)
    x = rng.standard_normal(R.shape[1])
    y = rng.standard_normal(R.shape[0])
    assert rel(R.dot(x), M @ x) < 1e-14
    assert rel(R.T.dot(y), M.T @ y) < 1e-14


def test_dense_assembly_entry_formula():
    """Brute-force assembly checked entry by entry on a 2-segment example."""
    D, K = random_dictionary(1, nVOX=2, nF=2, n=2, nS=3, ndirs=2, nIC=1, nEC=1, nISO=1)
    M = oracle.assemble_A(D, K).toarray()
    nS, nF, nE = 3, 2, D["EC"]["n"]
    exp = np.zeros_like(M)
    for k in range(2):
        v, f, o, l = D["IC"]["vox"][k], D["IC"]["str"][k], D["IC"]["dir"][k], float(D["IC"]["len"][k])
        for s in range(nS):
            exp[v * nS + s, f] += l * float(K["wmr"][0, o, s])
    for e in range(nE):
        v, o = D["EC"]["vox"][e], D["EC"]["dir"][e]
        for s in range(nS):
            exp[v * nS + s, nF + e] += float(K["wmh"][0, o, s])
    for kk in range(D["ISO"]["n"]):
        v = D["ISO"]["vox"][kk]
        for s in range(nS):
            exp[v * nS + s, nF + nE + kk] += float(K["iso"][0, s])
    assert np.array_equal(M, exp)


def test_single_segment_closed_form():
    D, K = random_dictionary(2, nVOX=1, nF=1, n=1, nS=11, ndirs=3, nIC=1, nEC=0, nISO=0)
    T = oracle.set_threads(D, 1)
    R = oracle.OracleOperator(D, K, T)
    l, o = float(D["IC"]["len"][0]), int(D["IC"]["dir"][0])
    y = R.dot(np.array([2.5]))
    assert np.allclose(y, 2.5 * l * K["wmr"][0, o].astype(np.float64), rtol=0, atol=1e-15)
    yy = np.arange(11, dtype=np.float64)
    assert np.isclose(R.T.dot(yy)[0], l * float(K["wmr"][0, o].astype(np.float64) @ yy), rtol=1e-15)


def test_scalarmap_incidence():
    """All-ones LUT (ScalarMap, models.pyx:75-80): A is the length-weighted
    voxel x fibre incidence matrix, replicated over samples, and equals the
    LUT-free operator on sample 0."""
    D, K = random_dictionary(3, nVOX=7, nF=9, n=40, nS=4, ndirs=3, nIC=1, nEC=0, nISO=0)
    K["wmr"][:] = 1.0
    T = oracle.set_threads(D, 2)
    R = oracle.OracleOperator(D, K, T)
    x = np.random.default_rng(0).random(9)
    inc = np.zeros((7, 9))
    np.add.at(inc, (D["IC"]["vox"], D["IC"]["str"]), D["IC"]["len"].astype(np.float64))
    y = R.dot(x).reshape(7, 4)
    for s in range(4):
        assert rel(y[:, s], inc @ x) < 1e-15
    K1 = {"wmr": np.ones((1, 3, 1), np.float32), "wmh": np.zeros((0, 3, 1), np.float32),
          "iso": np.zeros((0, 1), np.float32)}
    Rn = oracle.OracleOperator(D, K1, T, nolut=True)
    assert rel(Rn.dot(x), inc @ x) < 1e-15
    yv = np.random.default_rng(1).random(7)
    assert rel(Rn.T.dot(yv), inc.T @ yv) < 1e-15


def test_nolut_with_iso_closed_form():
    D, K = random_dictionary(4, nVOX=5, nF=6, n=20, nS=1, ndirs=2, nIC=1, nEC=0, nISO=1)
    T = oracle.set_threads(D, 2)
    Rn = oracle.OracleOperator(D, K, T, nolut=True)
    x = np.random.default_rng(2).random(Rn.shape[1])
    exp = np.zeros(5)
    np.add.at(exp, D["IC"]["vox"], D["IC"]["len"] * x[D["IC"]["str"]])
    exp[D["ISO"]["vox"]] += x[6:]
    assert rel(Rn.dot(x), exp) < 1e-15


def test_adjointness_small_workload():
    D, K, _ = synth.make_problem("small", seed=1001, nF=3000)
    T = oracle.set_threads(D, 4)
    R = oracle.OracleOperator(D, K, T)
    rng = np.random.default_rng(9)
    x = rng.random(R.shape[1])
    y = rng.random(R.shape[0])
    lhs, rhs = R.dot(x) @ y, x @ R.T.dot(y)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_thread_count_invariance():
    """Partitions change only the summation order (core.pyx:586-724)."""
    D, K, _ = synth.make_problem("small", seed=1001, nF=2000)
    x = np.random.default_rng(3).random(D["IC"]["nSTR"] + D["EC"]["n"] + 2 * D["ISO"]["n"])
    outs = [oracle.OracleOperator(D, K, oracle.set_threads(D, n)).dot(x) for n in (1, 5)]
    assert rel(outs[0], outs[1]) < 1e-14


def test_set_threads_partition_invariants():
    D, K, _ = synth.make_problem("small", seed=1001, nF=2000)
    T = oracle.set_threads(D, 6)
    ic = T["IC"].astype(np.int64)
    assert ic[0] == 0 and ic[-1] == D["IC"]["n"] and np.all(np.diff(ic) > 0)
    vox = D["IC"]["vox"]
    for b in ic[1:-1]:                      # ranges start on voxel boundaries
        assert vox[b] != vox[b - 1]
    ict = T["ICt"]
    for f in np.unique(D["IC"]["str"])[:200]:   # each fibre has a single owner
        assert np.unique(ict[D["IC"]["str"] == f]).size == 1
    with pytest.raises(ValueError):
        oracle.set_threads(D, 0)
    with pytest.raises(ValueError):
        oracle.set_threads(D, 256)
    Dt, _ = random_dictionary(1, nVOX=3, nF=3, n=5)
    with pytest.raises(ValueError, match="Too many threads"):
        oracle.set_threads(Dt, 5)


# ---------------------------------------------------------------- proximals --
def _brute_prox(x, omega):
    f = lambda u: 0.5 * np.sum((u - x) ** 2) + omega(u)
    best = None
    for x0 in (x, np.zeros_like(x), np.maximum(x, 0)):
        r = minimize(f, x0, method="Nelder-Mead", options={"xatol": 1e-10, "fatol": 1e-14, "maxiter": 20000})
        if best is None or r.fun < best.fun:
            best = r
    return best.x


def test_soft_thresholding_is_l1_prox():
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.standard_normal(3)
        lam = 0.4
        got = oracle.soft_thresholding(x.copy(), lam, 0, 3)
        ref = _brute_prox(x, lambda u: lam * np.abs(u).sum())
        assert np.allclose(got, ref, atol=1e-5)
        w = rng.uniform(0.2, 2.0, 3)
        got = oracle.w_soft_thresholding(x.copy(), w, lam, 0, 3)
        ref = _brute_prox(x, lambda u: lam * np.abs(w * u).sum())
        assert np.allclose(got, ref, atol=1e-5)


def test_non_negativity_is_projection():
    x = np.array([-1.0, 0.0, 2.0, -0.0, 3.5])
    got = oracle.non_negativity(x.copy(), 1, 3)
    assert np.array_equal(got, [-1.0, 0.0, 2.0, 0.0, 3.5])


def test_group_lasso_prox_and_omega():
    rng = np.random.default_rng(1)
    gidx = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    gsz = np.array([2, 3], dtype=np.int32)
    gw = np.array([1.0, 1.7])
    lam = 0.6
    omega = lambda u: lam * (gw[0] * np.linalg.norm(u[:2]) + gw[1] * np.linalg.norm(u[2:]))
    for _ in range(4):
        x = rng.standard_normal(5)
        got = oracle.prox_group_lasso(x.copy(), gidx, gsz, gw, lam)
        ref = _brute_prox(x, omega)
        assert np.allclose(got, ref, atol=2e-5)
        assert np.isclose(oracle.omega_group_lasso(x, gidx, gsz, gw, lam), omega(x))
    # a group whose norm falls below its threshold is zeroed exactly
    x = np.array([0.1, 0.1, 5.0, 0.0, 0.0])
    got = oracle.prox_group_lasso(x.copy(), gidx, gsz, gw, lam)
    assert np.all(got[:2] == 0.0) and got[2] == pytest.approx(5.0 - lam * 1.7)


def test_sparse_group_lasso_omega_sums_whole_vector():
    """omega_sparse_group_lasso sums |x| over the whole vector (proximals.pyx:150)."""
    x = np.array([1.0, -2.0, 3.0, -4.0])
    gidx = np.array([0, 1], dtype=np.int32)
    gsz = np.array([2], dtype=np.int32)
    v = oracle.omega_sparse_group_lasso(x, gidx, gsz, np.array([1.0]), 0.5, 0.25)
    assert v == pytest.approx(0.5 * 10.0 + 0.25 * np.sqrt(5.0))


# -------------------------------------------------------------------- FISTA --
class _Dense:
    def __init__(self, M):
        self.M = M
        self.shape = M.shape

    def dot(self, v):
        return self.M @ v

    @property
    def T(self):
        return _Dense(self.M.T)


def _nn_regularisation(n):
    rp = {"startIC": 0, "sizeIC": n, "startEC": n, "sizeEC": 0, "startISO": n, "sizeISO": 0,
          "regIC": None, "regEC": None, "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True}
    return oracle.init_regularisation(rp)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fista_reaches_nnls(seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((40, 12))
    xt = np.maximum(rng.standard_normal(12), 0)
    y = M @ xt + 0.01 * rng.standard_normal(40)
    A = _Dense(M)
    x, det = oracle.solve(y, A, A.T, tol_fun=1e-14, tol_x=1e-14, max_iter=20000,
                          regularisation=_nn_regularisation(12))
    ref, _ = nnls(M, y)
    assert rel(x, ref) < 1e-6
    g = M.T @ (M @ x - y)                   # KKT for min 0.5||Mx-y||^2 s.t. x >= 0
    assert np.all(x >= 0)
    assert np.all(g[x == 0] >= -1e-6)
    assert np.all(np.abs(g[x > 1e-8]) < 1e-5)
    assert det["iterations"] >= 1


def test_fista_orthogonal_closed_form():
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 8)))
    y = rng.standard_normal(30)
    A = _Dense(Q)
    x, _ = oracle.solve(y, A, A.T, tol_fun=1e-15, tol_x=1e-15, max_iter=5000,
                        regularisation=_nn_regularisation(8))
    assert np.allclose(x, np.maximum(Q.T @ y, 0), atol=1e-8)


def test_fista_group_lasso_kkt():
    rng = np.random.default_rng(7)
    M = rng.random((50, 6))
    y = M @ np.array([1.0, 0.5, 0, 0, 2.0, 0.0]) + 0.01 * rng.standard_normal(50)
    groups = np.empty(3, dtype=object)
    for k in range(3):
        groups[k] = np.array([2 * k, 2 * k + 1], dtype=np.int32)
    rp = {"startIC": 0, "sizeIC": 6, "startEC": 6, "sizeEC": 0, "startISO": 6, "sizeISO": 0,
          "regIC": "group_lasso", "regEC": None, "regISO": None, "nnIC": False, "nnEC": False, "nnISO": False,
          "lambdaIC": 0.5, "dictIC_params": {"group_idx_kept": groups, "group_weights": np.ones(3)}}
    reg = oracle.init_regularisation(rp)
    A = _Dense(M)
    x, _ = oracle.solve(y, A, A.T, tol_fun=1e-15, tol_x=1e-15, max_iter=50000, regularisation=reg)
    g = M.T @ (M @ x - y)
    for gi in groups:                        # optimality of the group-lasso problem
        nx = np.linalg.norm(x[gi])
        if nx > 1e-9:
            assert np.allclose(g[gi], -0.5 * x[gi] / nx, atol=1e-5)
        else:
            assert np.linalg.norm(g[gi]) <= 0.5 + 1e-6


# ----------------------------------------------------------- load_dictionary --
def test_dictionary_preprocessing_matches_oracle():
    raw = synth.tractogram(11, 500, (16, 16, 12), mode=0, step_vox=0.4)
    synth.add_peaks_and_iso(raw)
    D = synth.to_dictionary(raw)
    mask = D["MASK"]
    R = oracle.sort_and_normalize(raw, mask)
    for c in ("IC", "EC", "ISO"):
        for key in ("vox", "dir", "str", "len"):
            if key in R[c]:
                assert np.array_equal(D[c][key], R[c][key]), (c, key)
    assert D["IC"]["nVOX"] == R["IC"]["nVOX"]


def test_dictionary_preprocessing_hand_example():
    """sort_and_normalize on a hand-made raw dictionary (core.pyx:509-525,
    571-580): stable sort by linear voxel id, length / TRK_norm[fibre], linear
    ids x + X(y + Y z) mapped to their rank among the mask voxels in Fortran
    order.  Mask (3, 2, 2) with voxels (1,0,0), (2,0,0), (0,1,0), (2,1,1):
    linear ids 1, 2, 3, 11 -> compact 0, 1, 2, 3."""
    mask = np.zeros((3, 2, 2), dtype=bool)
    for v in ((1, 0, 0), (2, 0, 0), (0, 1, 0), (2, 1, 1)):
        mask[v] = True
    raw = {"TRK_norm": np.array([2.0, 0.5, 4.0], np.float32),
           "IC_f": np.array([0, 0, 1, 1, 2, 2], np.uint32),
           "IC_v": np.array([11, 2, 3, 11, 2, 1], np.uint32),
           "IC_o": np.array([4, 1, 0, 2, 3, 0], np.uint16),
           "IC_len": np.array([0.5, 1.0, 0.25, 0.75, 0.5, 1.0], np.float32),
           "EC_v": np.array([3, 1, 11, 1], np.uint32), "EC_o": np.array([7, 8, 9, 10], np.uint16),
           "ISO_v": np.array([11, 2, 1, 3], np.uint32)}
    D = oracle.sort_and_normalize(raw, mask)
    # stable order of the segments by linear voxel: 5 | 1, 4 | 2 | 0, 3
    assert D["IC"]["vox"].tolist() == [0, 1, 1, 2, 3, 3]
    assert D["IC"]["dir"].tolist() == [0, 1, 3, 0, 4, 2]
    assert D["IC"]["str"].tolist() == [2, 0, 2, 1, 0, 1]
    assert np.array_equal(D["IC"]["len"], np.array([1.0 / 4, 1.0 / 2, 0.5 / 4, 0.25 / 0.5, 0.5 / 2, 0.75 / 0.5],
                                                    np.float32))
    assert D["IC"]["nVOX"] == 4 and D["IC"]["n"] == 6 and D["IC"]["nSTR"] == 3
    assert D["EC"]["vox"].tolist() == [0, 0, 2, 3] and D["EC"]["dir"].tolist() == [8, 10, 7, 9]
    assert D["ISO"]["vox"].tolist() == [0, 1, 2, 3]
    U = oracle.sort_and_normalize(raw, mask, normalize=False)
    assert np.array_equal(U["IC"]["len"], np.array([1.0, 1.0, 0.5, 0.25, 0.5, 0.75], np.float32))


def test_kernel_normalisation_hand_example():
    """normalize_kernels (core.pyx:412-431): every row of an IC/EC response is
    divided by the norm of its orientation-0 row (so only row 0 gets norm 1),
    each ISO response by its own norm."""
    K = {"wmr": np.array([[[3.0, 4.0], [6.0, 8.0]]], np.float32),
         "wmh": np.array([[[0.0, 2.0], [1.0, 1.0]]], np.float32),
         "iso": np.array([[3.0, 4.0], [0.0, 0.5]], np.float32)}
    N = oracle.normalize_kernels(K)
    assert np.allclose(N["wmr"], [[[0.6, 0.8], [1.2, 1.6]]], rtol=1e-7, atol=0)
    assert np.allclose(N["wmh"], [[[0.0, 1.0], [0.5, 0.5]]], rtol=1e-7, atol=0)
    assert np.allclose(N["iso"], [[0.6, 0.8], [0.0, 1.0]], rtol=1e-7, atol=0)
    assert N["wmr_norm"].tolist() == [5.0] and N["wmh_norm"].tolist() == [2.0]
    assert N["iso_norm"].tolist() == [5.0, 0.5]
    assert N["wmr"].dtype == np.float32


def test_kernel_normalisation_matches_oracle():
    bvals, bvecs = synth.scheme((1000, 2000), 30)
    K = synth.szb_kernels(bvals, bvecs, synth.hemisphere_directions(50))
    a, b = synth.normalize_kernels(K), oracle.normalize_kernels(K)
    for key in ("wmr", "wmh", "iso"):
        assert np.array_equal(a[key], b[key])
        assert np.allclose(a[key + "_norm"], b[key + "_norm"], rtol=0, atol=0)
    assert np.isclose(np.linalg.norm(a["wmr"][0, 0]), 1.0, rtol=1e-6)


def test_composite_prox_three_compartments():
    """prox = proxIC(proxEC(proxISO(x))) (solvers.py:200-201): each compartment
    thresholded with its own lambda, then clipped, against the closed form."""
    n1, n2, n3 = 7, 5, 6
    rp = {"startIC": 0, "sizeIC": n1, "startEC": n1, "sizeEC": n2, "startISO": n1 + n2, "sizeISO": n3,
          "regIC": "lasso", "regEC": "lasso", "regISO": "lasso", "nnIC": True, "nnEC": False, "nnISO": True,
          "lambdaIC": 0.3, "lambdaEC": 0.2, "lambdaISO": 0.1}
    reg = oracle.init_regularisation(rp)
    x = np.random.default_rng(3).standard_normal(n1 + n2 + n3)
    sc = 0.5
    st = lambda v, t: np.sign(v) * np.maximum(np.abs(v) - t, 0.0)
    exp = np.concatenate([np.maximum(st(x[:n1], 0.3 * sc), 0), st(x[n1:n1 + n2], 0.2 * sc),
                          np.maximum(st(x[n1 + n2:], 0.1 * sc), 0)])
    assert np.allclose(reg["prox"](x.copy(), sc), exp, rtol=0, atol=1e-15)
    om = 0.3 * np.abs(x[:n1]).sum() + 0.2 * np.abs(x[n1:n1 + n2]).sum() + 0.1 * np.abs(x[n1 + n2:]).sum()
    assert reg["omega"](x) == pytest.approx(om, rel=1e-15)


def test_lut_grad_matches_finite_differences():
    """oracle.lut_grad against central differences of 0.5||A(LUT) x - y||^2 on
    the dense assembly (every LUT entry of a tiny dictionary)."""
    D, K = random_dictionary(21, nVOX=5, nF=6, n=25, nS=4, ndirs=3, nIC=2, nEC=1, nISO=2, ec_per_voxel=2)
    ev = np.ones(12, np.uint32)
    ev[3] = 0
    D["IC"]["eval"] = ev
    rng = np.random.default_rng(2)
    A = oracle.assemble_A(D, K)
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])

    def loss(KK):
        r = oracle.assemble_A(D, KK) @ x - y
        return 0.5 * float(r @ r)

    gy = A @ x - y
    g = oracle.lut_grad(D, K, x, gy)
    for name, gk in zip(("wmr", "wmh", "iso"), g):
        base = K[name].astype(np.float64)
        fd = np.zeros_like(base)
        for idx in np.ndindex(base.shape):
            KK = {k: (v.astype(np.float64) if k in ("wmr", "wmh", "iso") else v) for k, v in K.items()}
            h = 1e-6
            KK[name] = base.copy()
            KK[name][idx] += h
            lp = loss(KK)
            KK[name][idx] -= 2 * h
            fd[idx] = (lp - loss(KK)) / (2 * h)
        assert np.allclose(gk, fd, rtol=1e-6, atol=1e-8), name


def test_oracle_matches_golden_fixture():
    """The committed loop-evaluated fixture (tests/golden/make_golden.py)."""
    from _problems import load_golden
    D, K, g = load_golden()
    for threads in (1, 4):
        R = oracle.OracleOperator(D, K, oracle.set_threads(D, threads))
        assert rel(R.dot(g["x"]), g["Ax"]) < 1e-14
        assert rel(R.T.dot(g["y"]), g["Aty"]) < 1e-14
    M = oracle.assemble_A(D, K)
    assert rel(M @ g["x"], g["Ax"]) < 1e-14
    gw = oracle.lut_grad(D, K, g["x"], g["y"])
    for a, b in zip(gw, (g["gwmr"], g["gwmh"], g["giso"])):
        assert rel(a.ravel(), b.ravel()) < 1e-13
