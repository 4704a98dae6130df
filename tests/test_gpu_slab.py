"""GPU parity of the LDS-slab products (plan option COMMIT_HIP_OPT_SLAB, modes 1
and 2: the IC table staged in LDS by orientation slabs) against the oracle.

Bar: relative error <= 1e-12 in FP64 against the oracle's FP64 products
(A x, A^T y, the FISTA residual sqrtw (A x - y) with its squared norm and the
weighted gradient A^T (sqrtw r)), <= 1e-6 with FP32 storage; bitwise repeatable
run to run; <= 1e-12 from the per-voxel kernels (slab off).  Cases cover
partial last slabs, several IC RFs, empty voxels, a heavy voxel whose
single-slab ranges span several 64-entry chunks, eval masks, the LUT-free
plan, and nS past the slab kernels' limit (per-voxel fallback).
"""
import ctypes

import numpy as np
import pytest
import torch

from _problems import random_dictionary, rel, runs_dictionary
from commit_amd import _lib, autotune
from commit_amd.operator.operator import LinearOperator
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dictionary(seed, nVOX, n, nS, ndirs, nIC, nEC, nISO, heavy=None, ec_per_voxel=2):
    """random_dictionary with ~n/nVOX entries per voxel; heavy = (voxel, count)
    moves count entries into one voxel (the arrays stay voxel-sorted)."""
    D, K = random_dictionary(seed, nVOX=nVOX, nF=max(8, n // 25), n=n, nS=nS, ndirs=ndirs, nIC=nIC, nEC=nEC,
                             nISO=nISO, empty_voxels=(1, nVOX - 2), ec_per_voxel=ec_per_voxel)
    if heavy is not None:
        hv, cnt = heavy
        v = D["IC"]["vox"].copy()
        rng = np.random.default_rng(seed + 1)
        pick = np.sort(rng.choice(v.size, cnt, replace=False))
        v[pick] = hv
        idx = np.argsort(v, kind="mergesort")
        for k in ("vox", "str", "dir", "len"):
            D["IC"][k] = D["IC"][k][idx] if k != "vox" else v[idx]
    ev = np.ones(D["IC"]["nSTR"] * nIC, np.uint32)
    ev[::7] = 0
    D["IC"]["eval"] = ev
    return D, K


def _products(A, x, y, w):
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _lib.load()
    xd, yd, wd = (torch.from_numpy(a).cuda() for a in (x, y, w))
    res = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    rsq = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.check(lib.commit_hip_A_residual(A._plan.handle, P(xd), P(yd), P(wd), P(res), P(rsq), st), "A_residual")
    g = torch.empty(A.shape[1], dtype=torch.float64, device="cuda")
    _lib.check(lib.commit_hip_At_weighted(A._plan.handle, P(res), P(wd), P(g), st), "At_weighted")
    return A.dot(x), A.T.dot(y), res.cpu().numpy(), rsq.item(), g.cpu().numpy()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("nS,ndirs", [(1, 3), (7, 500), (60, 130), (65, 500), (288, 500), (320, 37)])
@pytest.mark.parametrize("rf", [(1, 1, 2), (2, 0, 1), (3, 2, 0)])
def test_slab_vs_oracle(mode, nS, ndirs, rf):
    nIC, nEC, nISO = rf
    D, K = _dictionary(nS * 7 + ndirs + mode, nVOX=150, n=9000, nS=nS, ndirs=ndirs, nIC=nIC, nEC=nEC, nISO=nISO,
                       heavy=(3, 1500))
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T)
    autotune.set_option(A, _lib.OPT_SLAB, mode)
    assert autotune.get_option(A, _lib.OPT_SLAB) == mode
    O = oracle.OracleOperator(D, K, T)
    rng = np.random.default_rng(nS + ndirs)
    x = rng.standard_normal(A.shape[1])
    x[rng.random(x.size) < 0.3] = 0.0
    y = rng.standard_normal(A.shape[0])
    w = rng.uniform(0.5, 1.5, A.shape[0])
    Ax, Aty, res, rsq, g = _products(A, x, y, w)
    Ax_o, Aty_o = O.dot(x), O.T.dot(y)
    r_o = w * (Ax_o - y)
    g_o = O.T.dot(w * r_o)
    e = (rel(Ax, Ax_o), rel(Aty, Aty_o), rel(res, r_o), abs(rsq - r_o @ r_o) / (r_o @ r_o), rel(g, g_o))
    print(f"mode={mode} nS={nS} ndirs={ndirs} rf={rf}: " + " ".join(f"{v:.1e}" for v in e), flush=True)
    assert max(e) <= TOL
    # bitwise repeatable, and the per-voxel kernels agree
    Ax2, Aty2, res2, _, g2 = _products(A, x, y, w)
    assert np.array_equal(Ax, Ax2) and np.array_equal(Aty, Aty2) and np.array_equal(g, g2)
    autotune.set_option(A, _lib.OPT_SLAB, 0)
    assert rel(A.dot(x), Ax) <= TOL and rel(A.T.dot(y), Aty) <= TOL
    # FP32 storage
    autotune.set_option(A, _lib.OPT_SLAB, mode)
    A.set_precision(_lib.FP32)
    x32 = x.astype(np.float32).astype(np.float64)
    y32 = y.astype(np.float32).astype(np.float64)
    assert rel(A.dot(x32), O.dot(x32)) <= 1e-6 and rel(A.T.dot(y32), O.T.dot(y32)) <= 1e-6


def test_slab_nolut_and_fallback():
    """LUT-free plan with slabs on; nS = 400 (past the slab kernels) falls back
    to the per-voxel kernels and stays right."""
    D, K = _dictionary(5, nVOX=90, n=4000, nS=7, ndirs=9, nIC=1, nEC=0, nISO=1)
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T, nolut=True)
    autotune.set_option(A, _lib.OPT_SLAB, 2)
    O = oracle.OracleOperator(D, K, T, nolut=True)
    rng = np.random.default_rng(9)
    x, y = rng.random(A.shape[1]), rng.random(A.shape[0])
    assert rel(A.dot(x), O.dot(x)) <= TOL and rel(A.T.dot(y), O.T.dot(y)) <= TOL
    D, K = _dictionary(6, nVOX=40, n=2000, nS=400, ndirs=20, nIC=1, nEC=1, nISO=1)
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T)
    autotune.set_option(A, _lib.OPT_SLAB, 1)
    O = oracle.OracleOperator(D, K, T)
    x, y = rng.random(A.shape[1]), rng.random(A.shape[0])
    assert rel(A.dot(x), O.dot(x)) <= TOL and rel(A.T.dot(y), O.T.dot(y)) <= TOL


def test_slab_option_errors_and_toggling():
    D, K = _dictionary(8, nVOX=60, n=3000, nS=60, ndirs=40, nIC=1, nEC=1, nISO=1)
    A = LinearOperator(D, K, None)
    with pytest.raises(RuntimeError):
        autotune.set_option(A, _lib.OPT_SLAB, 3)
    rng = np.random.default_rng(2)
    x = rng.random(A.shape[1])
    ref = A.dot(x)
    for mode in (1, 2, 2, 0, 1, 0):
        autotune.set_option(A, _lib.OPT_SLAB, mode)
        assert autotune.get_option(A, _lib.OPT_SLAB) == mode
        assert rel(A.dot(x), ref) <= TOL


@pytest.mark.parametrize("mode", [1, 2])
def test_slab_many_ec_iso_items(mode):
    """Voxels with more EC/ISO items (peaks x RFs) than one lane group holds
    (5 peaks x 2 EC RFs + 1 ISO RF = 11) take the per-entry path."""
    D, K = _dictionary(17 + mode, nVOX=80, n=4000, nS=288, ndirs=60, nIC=1, nEC=2, nISO=1, ec_per_voxel=5)
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T)
    autotune.set_option(A, _lib.OPT_SLAB, mode)
    O = oracle.OracleOperator(D, K, T)
    rng = np.random.default_rng(mode)
    x, y = rng.standard_normal(A.shape[1]), rng.standard_normal(A.shape[0])
    assert rel(A.dot(x), O.dot(x)) <= TOL and rel(A.T.dot(y), O.T.dot(y)) <= TOL


@pytest.mark.parametrize("option,value", [(_lib.OPT_SLAB, 0), (_lib.OPT_SLAB, 2), (_lib.OPT_A_DENSE, 1)])
def test_residual_norm_only(option, value):
    """commit_hip_A_residual with res NULL (a FISTA trial point needs only the
    norm): the squared norm is bitwise the one of the call that stores res,
    and res is left untouched; FP64 and FP32 storage."""
    D, K = _dictionary(23, nVOX=120, n=6000, nS=288, ndirs=60, nIC=1, nEC=1, nISO=2)
    A = LinearOperator(D, K, None)
    autotune.set_option(A, option, value)
    lib = _lib.load()
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(4)
    for prec, dt in ((_lib.FP64, torch.float64), (_lib.FP32, torch.float32)):
        A.set_precision(prec)
        x = torch.from_numpy(rng.standard_normal(A.shape[1])).to("cuda", dt)
        y = torch.from_numpy(rng.standard_normal(A.shape[0])).to("cuda", dt)
        w = torch.from_numpy(rng.uniform(0.5, 1.5, A.shape[0])).to("cuda", dt)
        for yo, wo in ((y, w), (y, None), (None, w)):
            res = torch.empty(A.shape[0], dtype=dt, device="cuda")
            r1 = torch.empty(1, dtype=torch.float64, device="cuda")
            r2 = torch.empty(1, dtype=torch.float64, device="cuda")
            _lib.check(lib.commit_hip_A_residual(A._plan.handle, P(x), P(yo), P(wo), P(res), P(r1), st), "res")
            keep = res.clone()
            _lib.check(lib.commit_hip_A_residual(A._plan.handle, P(x), P(yo), P(wo), None, P(r2), st), "norm")
            assert torch.equal(r1, r2) and torch.equal(res, keep)
            ref = (res.double() ** 2).sum().item()           # (FP32: from the rounded res)
            assert abs(r1.item() - ref) <= (1e-12 if prec == _lib.FP64 else 1e-6) * ref
    A.set_precision(_lib.FP64)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("zero_frac", [0.0, 0.3, 0.85])
def test_slab_runs_of_one_orientation(mode, zero_frac):
    """Entries of equal orientation in a voxel share one dot product (A^T) and
    one row application (A x merges the run's nonzero coefficients): every run
    length, zeros inside runs, against the oracle at the FP64 bar."""
    D, K = runs_dictionary(11 + mode)
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T)
    autotune.set_option(A, _lib.OPT_SLAB, mode)
    O = oracle.OracleOperator(D, K, T)
    rng = np.random.default_rng(mode)
    x = rng.uniform(0.1, 1.0, A.shape[1])
    x[rng.random(x.size) < zero_frac] = 0.0
    y = rng.standard_normal(A.shape[0])
    w = rng.uniform(0.5, 1.5, A.shape[0])
    Ax, Aty, res, rsq, g = _products(A, x, y, w)
    r_o = w * (O.dot(x) - y)
    e = (rel(Ax, O.dot(x)), rel(Aty, O.T.dot(y)), rel(res, r_o), rel(g, O.T.dot(w * r_o)))
    print(f"mode={mode} zeros={zero_frac}: " + " ".join(f"{v:.1e}" for v in e), flush=True)
    assert max(e) <= TOL


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_At_combined(mode):
    """commit_hip_At_combined: A^T (sqrtw .* (a r1 + b r2)) and ||a r1 + b r2||^2
    in one call.  On the slab paths (modes 1 and 2: 8 and 16 wavefronts per
    work-group) the combination is formed while
    A^T loads its rows: against the oracle's A^T of the combined residual
    (1e-12 in FP64, 1e-6 with FP32 storage) and against the two calls it
    replaces (commit_hip_residual_combine, commit_hip_At_weighted).  Slabs off
    (mode 0): the two calls themselves, bitwise."""
    D, K = _dictionary(31 + mode, nVOX=150, n=9000, nS=288, ndirs=300, nIC=1, nEC=1, nISO=2, heavy=(5, 900))
    T = oracle.set_threads(D, 4)
    A = LinearOperator(D, K, T)
    autotune.set_option(A, _lib.OPT_SLAB, mode)
    O = oracle.OracleOperator(D, K, T)
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib, h = _lib.load(), A._plan.handle
    rng = np.random.default_rng(mode + 3)
    m, n = A.shape
    a, b = 1.37, -0.37
    for prec, dt, tol in ((_lib.FP64, torch.float64, TOL), (_lib.FP32, torch.float32, 1e-6)):
        A.set_precision(prec)
        for weighted in (True, False):
            r1 = torch.from_numpy(rng.standard_normal(m)).to("cuda", dt)
            r2 = torch.from_numpy(rng.standard_normal(m)).to("cuda", dt)
            w = torch.from_numpy(rng.uniform(0.5, 1.5, m)).to("cuda", dt) if weighted else None
            rs, rsq = torch.empty(m, dtype=dt, device="cuda"), torch.empty(1, dtype=torch.float64, device="cuda")
            g = torch.empty(n, dtype=dt, device="cuda")
            _lib.check(lib.commit_hip_At_combined(h, a, P(r1), b, P(r2), P(w), P(rs), P(g), P(rsq), st), "combined")
            r, rsq2 = torch.empty(m, dtype=dt, device="cuda"), torch.empty(1, dtype=torch.float64, device="cuda")
            g2 = torch.empty(n, dtype=dt, device="cuda")
            _lib.check(lib.commit_hip_residual_combine(h, a, P(r1), b, P(r2), P(r), P(rsq2), st), "combine")
            _lib.check(lib.commit_hip_At_weighted(h, P(r), P(w), P(g2), st), "At_weighted")
            u = a * r1.double().cpu().numpy() + b * r2.double().cpu().numpy()
            wn = np.ones(m) if w is None else w.double().cpu().numpy()
            g_o = O.T.dot(wn * u.astype(np.float32).astype(np.float64) if prec == _lib.FP32 else wn * u)
            e = (rel(g.double().cpu().numpy(), g_o), abs(rsq.item() - u @ u) / (u @ u),
                 rel(g.double().cpu().numpy(), g2.double().cpu().numpy()), abs(rsq.item() - rsq2.item()) / rsq2.item())
            print(f"mode={mode} prec={prec} weighted={weighted}: " + " ".join(f"{v:.1e}" for v in e), flush=True)
            assert e[0] <= tol and e[1] <= 1e-13 and e[2] <= tol and e[3] <= 1e-13
            if mode == 0:
                assert torch.equal(g, g2) and rsq.item() == rsq2.item()
    A.set_precision(_lib.FP64)
