"""Batched fits on the GPU: several tractograms of uneven size stacked into one
block-diagonal plan, each with a lambda sweep of group-lasso fits, advanced
together by the device-side FISTA control (commit_amd.batched).

Bars: every (tractogram, lambda) fit runs the reference's iteration count and
its x agrees with the oracle's FP64 fit of that tractogram alone to 1e-4
relative (the batched products are float32 matrix-core products with FP32
accumulation per voxel); the graph-replayed run equals the eager run bitwise.
"""
import numpy as np
import pytest
import torch

from _problems import rel
from commit_amd import batched, synth
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL_BATCHED = 1e-4
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
