"""The reference's limit of 20 response functions per compartment
(setup_operator.py:70, 128, 183: range(0, 20)) on every device path: the
per-voxel and LDS-slab products in FP64 (1e-12 against the oracle) and
float32 vectors (1e-6), the 16-RHS products (1e-5 per column) and the table
gradient (1e-12).  21 RFs are rejected at plan creation (tests/test_abi.py)."""
import ctypes

import numpy as np
import pytest
import torch

from _problems import random_dictionary, rel
from commit_amd import _lib, autotune
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf20():
    D, K = random_dictionary(2020, nVOX=14, nF=11, n=150, nS=288, ndirs=120, nIC=20, nEC=20, nISO=20,
                             empty_voxels=(3,), ec_per_voxel=2)
    ev = np.ones(11 * 20, np.uint32)
    ev[[5, 77, 150]] = 0
    D["IC"]["eval"] = ev
    R = oracle.OracleOperator(D, K, oracle.set_threads(D, 3))
    rng = np.random.default_rng(20)
    return D, K, R, rng.standard_normal(R.shape[1]), rng.standard_normal(R.shape[0])


@pytest.mark.parametrize("slab", [0, 2])
def test_twenty_rfs_products(rf20, slab):
    D, K, R, x, y = rf20
    A = LinearOperator(D, K, None)
    autotune.set_option(A, _lib.OPT_SLAB, slab)
    e1, e2 = rel(A.dot(x), R.dot(x)), rel(A.T.dot(y), R.T.dot(y))
    print(f"20 RFs, slab {slab}: A x {e1:.1e}  A^T y {e2:.1e}", flush=True)
    assert e1 <= 1e-12 and e2 <= 1e-12
    A.set_precision(_lib.FP32)
    x32, y32 = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
    assert rel(A.dot(x32), R.dot(x32)) <= 1e-6 and rel(A.T.dot(y32), R.T.dot(y32)) <= 1e-6


def test_twenty_rfs_multi_rhs(rf20):
    D, K, R, x, y = rf20
    A = LinearOperator(D, K, None)
    rng = np.random.default_rng(21)
    X = rng.standard_normal((A.shape[1], 16)).astype(np.float32)
    Y = rng.standard_normal((A.shape[0], 16)).astype(np.float32)
    AX = A.matmat(torch.from_numpy(X).cuda()).double().cpu().numpy()
    AtY = A.T.matmat(torch.from_numpy(Y).cuda()).double().cpu().numpy()
    for r in range(16):
        assert rel(AX[:, r], R.dot(X[:, r].astype(np.float64))) <= 1e-5, r
        assert rel(AtY[:, r], R.T.dot(Y[:, r].astype(np.float64))) <= 1e-5, r


def test_twenty_rfs_lut_gradient(rf20):
    D, K, R, x, y = rf20
    A = LinearOperator(D, K, None)
    shapes = (K["wmr"].shape, K["wmh"].shape, K["iso"].shape)
    out = [torch.zeros(s, dtype=torch.float64, device="cuda") for s in shapes]
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    xd, gy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    _lib.check(_lib.load().commit_hip_lut_grad(A._plan.handle, P(xd), P(gy), P(out[0]), P(out[1]), P(out[2]), st),
               "commit_hip_lut_grad")
    for a, b in zip(out, oracle.lut_grad(D, K, x, y)):
        assert rel(a.cpu().numpy().ravel(), b.ravel()) <= 1e-12
