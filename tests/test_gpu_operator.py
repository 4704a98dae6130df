"""GPU parity of the HIP operator (through the C ABI) against the oracle.

Bar: relative error <= 1e-12 in FP64 against the oracle's FP64 products; the
oracle's summation order differs (it is sequential per sample), so results
are not bitwise, but the GPU path is deterministic run to run (checked).
"""
import numpy as np
import pytest
import torch

from _problems import random_dictionary, rel
from commit_amd import synth
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def small():
    D, K, _ = synth.make_problem("small", seed=1001)
    T = oracle.set_threads(D, 8)
    return D, K, T


def _pair(D, K, T, x, y, nolut=False):
    A = LinearOperator(D, K, T, nolut=nolut)
    R = oracle.OracleOperator(D, K, T, nolut=nolut)
    return A.dot(x), R.dot(x), A.T.dot(y), R.T.dot(y)


def test_small_workload_dense_x(small):
    D, K, T = small
    A = LinearOperator(D, K, T)
    rng = np.random.default_rng(1)
    x = rng.random(A.shape[1])
    y = rng.random(A.shape[0])
    ya, yr, xa, xr = _pair(D, K, T, x, y)
    print(f"small dense: A x {rel(ya, yr):.2e}  A^T y {rel(xa, xr):.2e}", flush=True)
    assert rel(ya, yr) <= TOL and rel(xa, xr) <= TOL


def test_small_workload_sparse_x(small):
    D, K, T = small
    A = LinearOperator(D, K, T)
    x = synth.x_true(A.shape[1], 2001)
    ya = A.dot(x)
    yr = oracle.OracleOperator(D, K, T).dot(x)
    assert rel(ya, yr) <= TOL


@pytest.mark.parametrize("nS", [1, 7, 60, 64, 65, 128, 288, 300])
@pytest.mark.parametrize("rf", [(1, 1, 2), (2, 0, 1), (3, 2, 0), (1, 0, 0)])
def test_random_dictionaries(nS, rf):
    nIC, nEC, nISO = rf
    D, K = random_dictionary(11 + nS, nVOX=9, nF=13, n=80, nS=nS, ndirs=5, nIC=nIC, nEC=nEC, nISO=nISO,
                             empty_voxels=(3,), ec_per_voxel=2)
    T = oracle.set_threads(D, 2)
    A = LinearOperator(D, K, T)
    rng = np.random.default_rng(nS)
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])
    dense = oracle.assemble_A(D, K)
    assert rel(A.dot(x), dense @ x) <= TOL
    assert rel(A.T.dot(y), dense.T @ y) <= TOL


def test_nolut_matches_oracle():
    D, K = random_dictionary(5, nVOX=8, nF=10, n=60, nS=1, ndirs=3, nIC=1, nEC=0, nISO=1)
    T = oracle.set_threads(D, 3)
    rng = np.random.default_rng(3)
    A = LinearOperator(D, K, T, nolut=True)
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])
    ya, yr, xa, xr = _pair(D, K, T, x, y, nolut=True)
    assert rel(ya, yr) <= TOL and rel(xa, xr) <= TOL


def test_eval_mask_and_skip():
    D, K = random_dictionary(9, nVOX=7, nF=12, n=70, nS=33, ndirs=6, nIC=2, nEC=1, nISO=1)
    T = oracle.set_threads(D, 2)
    A = LinearOperator(D, K, T)
    ev = (np.arange(12 * 2) % 3 != 0).astype(np.uint32)
    A.set_eval(ev)
    D2 = dict(D)
    D2["IC"] = dict(D["IC"], eval=ev)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(A.shape[1])
    x[::4] = 0.0
    y = rng.standard_normal(A.shape[0])
    dense = oracle.assemble_A(D2, K)
    assert rel(A.dot(x), dense @ x) <= TOL
    assert rel(A.T.dot(y), dense.T @ y) <= TOL


def test_size_check_and_dtype():
    D, K = random_dictionary(2)
    A = LinearOperator(D, K, None)
    with pytest.raises(RuntimeError, match="dimensions do not match"):
        A.dot(np.zeros(A.shape[1] + 1))
    with pytest.raises(RuntimeError, match="dimensions do not match"):
        A.T.dot(np.zeros(A.shape[1]))
    assert A.T.shape == (A.shape[1], A.shape[0])
    # a device output buffer of the wrong size, type or layout is refused
    # before any launch
    xd = torch.zeros(A.shape[1], dtype=torch.float64, device="cuda")
    for bad in (torch.empty(A.shape[0] - 1, dtype=torch.float64, device="cuda"),
                torch.empty(A.shape[0], dtype=torch.float32, device="cuda"),
                torch.empty((A.shape[0], 2), dtype=torch.float64, device="cuda")[:, 0]):
        with pytest.raises(RuntimeError, match="out must be"):
            A.matvec(xd, out=bad)
    ok = torch.full((A.shape[0],), 7.0, dtype=torch.float64, device="cuda")
    assert A.matvec(xd, out=ok) is ok and not ok.any()


def test_device_path_and_determinism(small):
    D, K, T = small
    A = LinearOperator(D, K, T)
    rng = np.random.default_rng(5)
    x = rng.random(A.shape[1])
    y = rng.random(A.shape[0])
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y).cuda()
    y1, y2 = A.dot(xd), A.dot(xd)
    x1, x2 = A.T.dot(yd), A.T.dot(yd)
    assert torch.equal(y1, y2) and torch.equal(x1, x2), "products must be bitwise repeatable"
    assert np.array_equal(y1.cpu().numpy(), A.dot(x))
    assert np.array_equal(x1.cpu().numpy(), A.T.dot(y))


def test_adjointness_brain_geometry():
    """Size-independent property at the brain geometry (200k fibres, nS=288)."""
    D, K, _ = synth.make_problem("brain200k", seed=1003)
    A = LinearOperator(D, K, None)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.rand(A.shape[1], dtype=torch.float64, device="cuda", generator=g)
    y = torch.rand(A.shape[0], dtype=torch.float64, device="cuda", generator=g)
    lhs = torch.dot(A.dot(x), y).item()
    rhs = torch.dot(x, A.T.dot(y)).item()
    print(f"adjointness: <Ax,y>={lhs:.15e} <x,A^T y>={rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_golden_fixture_all_paths():
    """The committed loop-evaluated fixture through every device path:
    FP64 (1e-12), FP32 storage (1e-6), multi-RHS (1e-5), table gradient (1e-12)."""
    import ctypes
    from _problems import load_golden
    from commit_amd import _lib
    D, K, g = load_golden()
    A = LinearOperator(D, K, None)
    assert rel(A.dot(g["x"]), g["Ax"]) <= TOL and rel(A.T.dot(g["y"]), g["Aty"]) <= TOL
    xd = torch.from_numpy(g["x"]).cuda()
    yd = torch.from_numpy(g["y"]).cuda()
    out = [torch.zeros(K[k].shape, dtype=torch.float64, device="cuda") for k in ("wmr", "wmh", "iso")]
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(_lib.load().commit_hip_lut_grad(A._plan.handle, P(xd), P(yd), P(out[0]), P(out[1]), P(out[2]),
                                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "lut_grad")
    for o, k in zip(out, ("gwmr", "gwmh", "giso")):
        assert rel(o.cpu().numpy().ravel(), g[k].ravel()) <= TOL
    A.set_precision(_lib.FP32)
    assert rel(A.dot(g["x"]), g["Ax"]) <= 1e-6 and rel(A.T.dot(g["y"]), g["Aty"]) <= 1e-6
    X = np.stack([g["x"] * (r + 1) for r in range(16)], axis=1).astype(np.float32)
    Y = np.stack([g["y"] * (r + 1) for r in range(16)], axis=1).astype(np.float32)
    Ym = A.matmat(torch.from_numpy(X).cuda()).cpu().numpy()
    Xm = A.T.matmat(torch.from_numpy(Y).cuda()).cpu().numpy()
    for r in range(16):
        assert rel(Ym[:, r], (r + 1) * g["Ax"]) <= 1e-5 and rel(Xm[:, r], (r + 1) * g["Aty"]) <= 1e-5


@pytest.mark.parametrize("ndirs", [3, 250, 500, 511])
@pytest.mark.parametrize("nS", [1, 17, 60, 288, 320])
def test_dense_tile_A(ndirs, nS):
    """The dense-tile matrix-core A x (plan option) against the dense assembly:
    plain, residual form with weights, eval mask, float32 storage."""
    import ctypes
    from commit_amd import _lib, autotune
    D, K = random_dictionary(ndirs + 3 * nS, nVOX=37, nF=23, n=700, nS=nS, ndirs=ndirs, nIC=2, nEC=1, nISO=2,
                             empty_voxels=(5, 17), ec_per_voxel=2)
    ev = np.ones(46, np.uint32)
    ev[[0, 9, 30]] = 0
    D["IC"]["eval"] = ev
    A = LinearOperator(D, K, None)
    autotune.set_option(A, _lib.OPT_A_DENSE, 1)
    rng = np.random.default_rng(ndirs + nS)
    x = rng.standard_normal(A.shape[1])
    x[rng.random(x.size) < 0.3] = 0.0
    dense = oracle.assemble_A(D, K)
    assert rel(A.dot(x), dense @ x) <= TOL
    # residual form: sqrtw (A x - y) and its squared norm
    yobs = rng.standard_normal(A.shape[0])
    w = rng.uniform(0.5, 1.5, A.shape[0])
    res = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    rsq = torch.empty(1, dtype=torch.float64, device="cuda")
    P = lambda a: ctypes.c_void_p(a.data_ptr())
    xd, yd, wd = (torch.from_numpy(a).cuda() for a in (x, yobs, w))
    _lib.check(_lib.load().commit_hip_A_residual(A._plan.handle, P(xd), P(yd), P(wd), P(res), P(rsq),
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "res")
    r_ref = w * (dense @ x - yobs)
    assert rel(res.cpu().numpy(), r_ref) <= TOL
    assert abs(rsq.item() - r_ref @ r_ref) <= 1e-12 * (r_ref @ r_ref)
    A.set_precision(_lib.FP32)
    x32 = x.astype(np.float32).astype(np.float64)
    assert rel(A.dot(x32), dense @ x32) <= 1e-6
