"""GPU parity of the multi-RHS products (commit_hip_A_multi / _At_multi,
FP32 storage, matrix-core accumulation) against the FP64 oracle.

Bar: for every right-hand side, ||Y_gpu[:, r] - Y_oracle[:, r]|| / ||Y_oracle[:, r]||
<= 1e-5 (the FP32 tolerance of the mixed-precision path), with the oracle fed
the same float32-rounded inputs in float64; results are bitwise repeatable.
"""
import numpy as np
import pytest
import torch

from _problems import random_dictionary, rel, runs_dictionary
from commit_amd import synth
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def _cols_rel(a, b):
    return max(rel(a[:, r], b[:, r]) for r in range(b.shape[1]))


def _dev32(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("nS", [1, 7, 16, 60, 65, 288, 320])
@pytest.mark.parametrize("rf", [(1, 1, 2), (2, 0, 1), (3, 2, 0)])
@pytest.mark.parametrize("R", [1, 16, 21])
def test_random_dictionaries_multi(nS, rf, R):
    nIC, nEC, nISO = rf
    D, K = random_dictionary(31 + nS + R, nVOX=9, nF=13, n=150, nS=nS, ndirs=5, nIC=nIC, nEC=nEC, nISO=nISO,
                             empty_voxels=(3,), ec_per_voxel=2)
    A = LinearOperator(D, K, None)
    rng = np.random.default_rng(nS * 100 + R)
    X = rng.standard_normal((A.shape[1], R)).astype(np.float32)
    Y = rng.standard_normal((A.shape[0], R)).astype(np.float32)
    dense = oracle.assemble_A(D, K)
    Ya = A.matmat(_dev32(X)).cpu().numpy()
    Xa = A.T.matmat(_dev32(Y)).cpu().numpy()
    ey = _cols_rel(Ya, dense @ X.astype(np.float64))
    ex = _cols_rel(Xa, dense.T @ Y.astype(np.float64))
    print(f"nS={nS} rf={rf} R={R}: A X {ey:.2e}  A^T Y {ex:.2e}", flush=True)
    assert ey <= TOL32 and ex <= TOL32


def test_multi_eval_mask_and_weights():
    D, K = random_dictionary(77, nVOX=12, nF=20, n=300, nS=60, ndirs=7, nIC=2, nEC=1, nISO=2)
    ev = np.ones(40, np.uint32)
    ev[[1, 5, 22, 39]] = 0
    D["IC"]["eval"] = ev
    A = LinearOperator(D, K, None)
    rng = np.random.default_rng(5)
    R = 16
    X = rng.random((A.shape[1], R)).astype(np.float32)
    Y = rng.standard_normal((A.shape[0], R)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, A.shape[0]).astype(np.float32)
    dense = oracle.assemble_A(D, K)          # uses DICTIONARY['IC']['eval']
    Ya = A.matmat(_dev32(X)).cpu().numpy()
    Xa = A.T.matmat(_dev32(Y), sqrtw=_dev32(w)).cpu().numpy()
    assert _cols_rel(Ya, dense @ X.astype(np.float64)) <= TOL32
    Yw = (Y.astype(np.float64) * w.astype(np.float64)[:, None])
    assert _cols_rel(Xa, dense.T @ Yw) <= TOL32


@pytest.fixture(scope="module")
def small():
    D, K, _ = synth.make_problem("small", seed=1001)
    return D, K


def test_small_workload_16_rhs_vs_oracle(small):
    """The CPU-sized workload (10k fibres, nS=60), 16 RHS: every column against
    the oracle's single-vector products."""
    D, K = small
    T = oracle.set_threads(D, 8)
    A = LinearOperator(D, K, T)
    R = 16
    rng = np.random.default_rng(16)
    X = rng.random((A.shape[1], R)).astype(np.float32)
    X[rng.random(X.shape) < 0.6] = 0.0                 # sparse, different per column
    Y = rng.random((A.shape[0], R)).astype(np.float32)
    Ya = A.matmat(_dev32(X)).cpu().numpy()
    Xa = A.T.matmat(_dev32(Y)).cpu().numpy()
    O = oracle.OracleOperator(D, K, T)
    Yo = np.stack([O.dot(X[:, r].astype(np.float64)) for r in range(R)], axis=1)
    Xo = np.stack([O.T.dot(Y[:, r].astype(np.float64)) for r in range(R)], axis=1)
    ey, ex = _cols_rel(Ya, Yo), _cols_rel(Xa, Xo)
    print(f"small workload, 16 RHS: A X {ey:.2e}  A^T Y {ex:.2e}", flush=True)
    assert ey <= TOL32 and ex <= TOL32


def test_multi_matches_single_fp64_columns(small):
    """Column r of the multi-RHS product equals the FP64 single-vector GPU
    product of column r (to the FP32 bar), and repeated calls are bitwise equal."""
    D, K = small
    A = LinearOperator(D, K, None)
    R = 5
    rng = np.random.default_rng(7)
    X = _dev32(rng.random((A.shape[1], R)))
    Y1 = A.matmat(X)
    Y2 = A.matmat(X)
    assert torch.equal(Y1, Y2)
    Yt = _dev32(rng.random((A.shape[0], R)))
    X1 = A.T.matmat(Yt)
    X2 = A.T.matmat(Yt)
    assert torch.equal(X1, X2)
    for r in range(R):
        y64 = A.matvec(X[:, r].contiguous().double())
        x64 = A.T.matvec(Yt[:, r].contiguous().double())
        assert rel(Y1[:, r].double().cpu().numpy(), y64.cpu().numpy()) <= TOL32
        assert rel(X1[:, r].double().cpu().numpy(), x64.cpu().numpy()) <= TOL32


def test_multi_rejects_bad_input(small):
    D, K = small
    A = LinearOperator(D, K, None)
    with pytest.raises(RuntimeError):
        A.matmat(torch.zeros((A.shape[1] + 1, 4), dtype=torch.float32, device="cuda"))
    with pytest.raises(RuntimeError):
        A.matmat(torch.zeros((A.shape[1], 4), dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("waves,remap,dense", [(1, 1, 0), (2, 0, 1), (1, 0, 1)])
def test_launch_options_keep_parity(waves, remap, dense):
    """Every variant the autotuner may pick computes the same products."""
    from commit_amd import _lib, autotune
    D, K = random_dictionary(401, nVOX=12, nF=15, n=220, nS=288, ndirs=6, nIC=1, nEC=1, nISO=2)
    A = LinearOperator(D, K, None)
    autotune.set_option(A, _lib.OPT_MULTI_A_WAVES, waves)
    autotune.set_option(A, _lib.OPT_XCD_REMAP, remap)
    autotune.set_option(A, _lib.OPT_A_DENSE, dense)
    assert autotune.get_option(A, _lib.OPT_MULTI_A_WAVES) == waves
    rng = np.random.default_rng(waves + 10 * remap)
    X = rng.standard_normal((A.shape[1], 16)).astype(np.float32)
    Y = rng.standard_normal((A.shape[0], 16)).astype(np.float32)
    dense = oracle.assemble_A(D, K)
    assert _cols_rel(A.matmat(_dev32(X)).cpu().numpy(), dense @ X.astype(np.float64)) <= TOL32
    assert _cols_rel(A.T.matmat(_dev32(Y)).cpu().numpy(), dense.T @ Y.astype(np.float64)) <= TOL32
    x = rng.standard_normal(A.shape[1])
    assert rel(A.dot(x), dense @ x) <= 1e-12


def test_autotune_sets_and_caches(small, tmp_path):
    from commit_amd import _lib, autotune
    D, K = small
    A = LinearOperator(D, K, None)
    cache = str(tmp_path / "tune.json")
    rep = autotune.tune(A, nrhs=16, reps=2, cache=cache)
    assert rep["xcd_remap"]["choice"] in (0, 1) and rep["multi_a_waves"]["choice"] in (1, 2)
    assert rep["slab"]["choice"] in (0, 1, 2)
    assert autotune.get_option(A, _lib.OPT_SLAB) == rep["slab"]["choice"]
    assert autotune.get_option(A, _lib.OPT_A_DENSE) == rep["a_dense"]["choice"]
    assert autotune.get_option(A, _lib.OPT_MULTI_A_WAVES) == rep["multi_a_waves"]["choice"]
    B = LinearOperator(D, K, None)
    rep2 = autotune.tune(B, nrhs=16, cache=cache)            # from the cache, no timing
    assert rep2 == rep
    assert autotune.get_option(B, _lib.OPT_XCD_REMAP) == rep["xcd_remap"]["choice"]
    assert autotune.get_option(B, _lib.OPT_SLAB) == rep["slab"]["choice"]


@pytest.mark.parametrize("rf", [(1, 1, 2), (2, 1, 0)])
def test_multi_runs_of_one_orientation(rf):
    """16-RHS A^T Y computes one row product per run of equal orientation in a
    voxel (runs of every length up to 200, across 16-row tiles and 64-row
    record windows) and pass 2 scales it by each entry's length: per column
    against the dense oracle, with an eval mask."""
    nIC, nEC, nISO = rf
    D, K = runs_dictionary(23, nS=288, nIC=nIC, nEC=nEC, nISO=nISO)
    ev = np.ones(D["IC"]["nSTR"] * nIC, np.uint32)
    ev[::5] = 0
    D["IC"]["eval"] = ev
    A = LinearOperator(D, K, None)
    rng = np.random.default_rng(5)
    R = 16
    X = rng.standard_normal((A.shape[1], R)).astype(np.float32)
    Y = rng.standard_normal((A.shape[0], R)).astype(np.float32)
    dense = oracle.assemble_A(D, K)
    Xa = A.T.matmat(_dev32(Y)).cpu().numpy()
    Ya = A.matmat(_dev32(X)).cpu().numpy()
    ex = _cols_rel(Xa, dense.T @ Y.astype(np.float64))
    ey = _cols_rel(Ya, dense @ X.astype(np.float64))
    print(f"rf={rf}: A X {ey:.2e}  A^T Y {ex:.2e}", flush=True)
    assert ey <= TOL32 and ex <= TOL32
