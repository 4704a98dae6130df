"""Pins the oracle's restatements of the Evaluation pieces around the fit
(core.pyx) on hand-computed examples: lambda_max (core.pyx:961-976), the
group structure and weights (core.pyx:1123-1192), the debias masks
(core.pyx:1476-1501), the confidence values (core.pyx:1433-1457), the fit
error maps and statistics (core.pyx:1614-1707), the compartment maps
(core.pyx:1709-1756) and the set_wLasso_ISO weights (core.pyx:828-843).
Each expected value is worked out by hand in the comments."""
from math import sqrt

import numpy as np
import pytest

from oracle import oracle


def test_lambda_max_lasso_and_group():
    aty = np.array([3.0, -4.0, 1.0, 10.0])
    # |3|/1, |-4|/2, |1|/0.5 over the slice [0, 3); the 10 lies outside it
    assert oracle.lambda_max_lasso(aty, 0, 3, np.array([1.0, 2.0, 0.5])) == 3.0
    aty = np.array([3.0, 4.0, 1.0, 0.0, 2.0])
    groups = [np.array([0, 1]), np.array([2, 3, 4])]
    # ||(3,4)|| / 1 = 5, ||(1,0,2)|| / 2 = sqrt(5)/2
    assert oracle.lambda_max_group(aty, np.array([1.0, 2.0]), groups) == 5.0
    assert oracle.lambda_max_group(aty, np.array([10.0, 0.5]), groups) == pytest.approx(2 * sqrt(5.0), rel=1e-15)


def test_group_weights_with_dropped_streamlines():
    # streamline 2 is not kept: kept indices map 0,1,3,4,5 -> 0,1,2,3,4
    kept = np.array([1, 1, 0, 1, 1, 1], np.uint8)
    group_idx = [np.array([0, 1]), np.array([], dtype=np.int32), np.array([2, 3, 4]), np.array([5]), np.array([2])]
    extra = [2.0, 3.0, 5.0, 7.0]          # one per non-empty group
    x_nnls = np.array([3.0, 4.0, 0.0, 1.0, 2.0])
    g, w = oracle.group_weights(group_idx, kept, 5, extra=extra, cardinality=True, adaptive=True, x_nnls=x_nnls)
    # the empty group goes; [2] maps to nothing and goes with its weight 7
    assert [list(a) for a in g] == [[0, 1], [2, 3], [4]]
    # weights * sqrt(|g|) / (||x_nnls[g]|| + 1e-12): ||(3,4)|| = 5, ||(0,1)|| = 1, ||(2)|| = 2
    exp = [2 * sqrt(2) / (5 + 1e-12), 3 * sqrt(2) / (1 + 1e-12), 5 * 1 / (2 + 1e-12)]
    assert np.allclose(w, exp, rtol=1e-15, atol=0)
    g, w = oracle.group_weights(group_idx, np.ones(6, np.uint8), 6, cardinality=False, adaptive=False)
    assert [list(a) for a in g] == [[0, 1], [2, 3, 4], [5], [2]] and w.tolist() == [1.0, 1.0, 1.0, 1.0]
    g, w = oracle.group_weights(group_idx, np.ones(6, np.uint8), 6, cardinality=True, adaptive=False)
    assert np.allclose(w, [sqrt(2), sqrt(3), 1.0, 1.0], rtol=1e-15)


class _Dense:
    def __init__(self, M):
        self.M = np.asarray(M, dtype=np.float64)

    def dot(self, v):
        return self.M @ v


def test_debias_masks():
    x = np.array([0.5, 1e-16, 0.0, 2e-15, -1.0, 3.0])      # 5 IC columns, then one EC column
    mask = oracle.debias_eval(x, 5)
    assert mask.tolist() == [1, 0, 0, 1, 0] and mask.dtype == np.uint32
    M = [[1, 1, 0, 0, 0, 1],        # 0.5 (+ the EC column, zeroed)
         [0, 0, 1, 0, 0, 0],        # masked column only
         [0, 0, 0, 0, 1, 0],        # masked column only
         [0, 0, 0, -1, 0, 0]]       # -2e-15 < 0
    assert oracle.debias_rows(_Dense(M), x, mask, 5).tolist() == [1.0, 0.0, 0.0, 0.0]


def test_confidence_values():
    a, ch = oracle.confidence(np.array([1.0, np.nan, 3.0, 5.0]), rescale=False)
    assert a.tolist() == [1.0, 0.0, 3.0, 5.0] and ch
    a, ch = oracle.confidence(np.array([1.0, np.nan, 3.0, 5.0]), rescale=True)
    assert np.allclose(a, [0.2, 0.0, 0.6, 1.0], rtol=1e-7) and ch          # min 0 after the NaN
    a, ch = oracle.confidence(np.array([2.0, 2.0, 2.0]), rescale=True)
    assert a is None and not ch
    a, ch = oracle.confidence(np.array([1.0, 2.0]), rescale=False)
    assert a.tolist() == [1.0, 2.0] and not ch


def test_fit_errors_plain_adjusted_and_debias():
    y = np.array([1.0, 2.0, 3.0, 4.0])           # 2 voxels x 2 samples
    est = np.array([1.0, 1.0, 3.0, 5.0])
    e = oracle.fit_errors(y, y, est, 2, 2, conf=np.array([1.0, 0.0, 2.0, 2.0]))
    # voxel 0: residual (0, 1): RMSE sqrt(1/2), NRMSE sqrt(1/5); voxel 1: (0, -1): sqrt(1/2), sqrt(1/25)
    assert np.allclose(e["RMSE"], [sqrt(0.5), sqrt(0.5)], rtol=1e-7)
    assert np.allclose(e["NRMSE"], [sqrt(0.2), 0.2], rtol=1e-7)
    assert e["RMSE_stats"][0] == pytest.approx(sqrt(0.5), rel=1e-7) and e["RMSE_stats"][1] == pytest.approx(0, abs=1e-7)
    # weighted: voxel 0 sqrt((1*0 + 0*1)/1) = 0, voxel 1 sqrt((2*0 + 2*1)/4)
    assert np.allclose(e["RMSE_adjusted"], [0.0, sqrt(0.5)], rtol=1e-7)
    # voxel 0 sqrt(0 / (1*1 + 0*4)) = 0, voxel 1 sqrt(2 / (2*9 + 2*16))
    assert np.allclose(e["NRMSE_adjusted"], [0.0, 0.2], rtol=1e-7)
    # debias: only voxel 0 in the row mask; the fit saw y * mask
    dm = np.array([1.0, 1.0, 0.0, 0.0])
    e = oracle.fit_errors(y, y * dm, est, 2, 2, debias_mask=dm)
    assert e["RMSE_stats"] == pytest.approx((sqrt(0.5), 0.0), rel=1e-12, abs=1e-15)
    assert e["NRMSE_stats"] == pytest.approx((sqrt(0.2), 0.0), rel=1e-12, abs=1e-15)
    # maps over every voxel against get_y(): voxel 1 compares (0, 0) with (3, 5)
    assert np.allclose(e["RMSE"], [sqrt(0.5), sqrt(17.0)], rtol=1e-7)
    assert np.allclose(e["NRMSE"], [sqrt(0.2), 0.0], rtol=1e-7)      # zero signal -> 0


def _maps_dictionary():
    D = {"IC": {"vox": np.array([0, 1, 1], np.uint32), "str": np.array([0, 0, 1], np.uint32),
                "len": np.array([0.5, 0.25, 1.0], np.float32), "nSTR": 2, "n": 3},
         "EC": {"vox": np.array([0, 1], np.uint32), "n": 2},
         "ISO": {"vox": np.array([0, 1], np.uint32), "n": 2},
         "TRK": {"norm": np.array([0.5, 2.0], np.float32)}}
    K = {"wmr": np.zeros((1, 1, 1)), "wmh": np.zeros((1, 1, 1)), "iso": np.zeros((2, 1)),
         "wmr_norm": np.array([2.0]), "wmh_norm": np.array([4.0]), "iso_norm": np.array([10.0, 20.0])}
    x = np.array([2.0, 4.0, 1.0, 3.0, 5.0, 6.0, 7.0, 8.0])   # IC | EC | ISO RF 0 | ISO RF 1
    return D, K, x


def test_compartment_maps():
    D, K, x = _maps_dictionary()
    ic, ec, iso = oracle.compartment_maps(x, D, K, 2, normalize_kernels=False)
    # IC: fibre 0 (x=2) has 0.5 in voxel 0 and 0.25 in voxel 1, fibre 1 (x=4) 1.0 in voxel 1
    assert ic.tolist() == [1.0, 4.5] and ec.tolist() == [1.0, 3.0]
    assert iso.tolist() == [5.0, 6.0]                # first ISO RF only, not 5+7, 6+8
    ic, ec, iso = oracle.compartment_maps(x, D, K, 2, normalize_kernels=False, normalize_maps=True)
    assert np.allclose(ic, [1 / 7, 4.5 / 13.5], rtol=1e-6) and np.allclose(iso, [5 / 7, 6 / 13.5], rtol=1e-6)
    # renormalised: IC x / (2 * TRK_norm) = (2, 1), spread with norm_fib: (1, 2);
    # EC / 4, ISO / 10
    ic, ec, iso = oracle.compartment_maps(x, D, K, 2, normalize_kernels=True)
    assert np.allclose(ic, [0.5, 2.25], rtol=1e-7) and np.allclose(ec, [0.25, 0.75], rtol=1e-7)
    assert np.allclose(iso, [0.5, 0.6], rtol=1e-7)


def test_iso_weights():
    img = np.zeros((2, 2, 1), np.float32)
    img[0, 0, 0], img[1, 0, 0], img[1, 1, 0] = 2.0, np.nan, 3.0
    D = {"MASK_ix": np.array([0, 1, 1]), "MASK_iy": np.array([0, 0, 1]), "MASK_iz": np.array([0, 0, 0]),
         "ISO": {"vox": np.array([2, 0, 1], np.uint32), "n": 3}}
    # mask values (2, nan -> 1, 3), indexed by the ISO voxels (2, 0, 1), tiled over 2 RFs
    assert oracle.iso_weights(img, D, 2).tolist() == [3.0, 2.0, 1.0, 3.0, 2.0, 1.0]
    img[0, 0, 0] = 0.5
    with pytest.raises(ValueError):
        oracle.iso_weights(img, D, 1)
