"""End-to-end drop-in flow on files: trk2dictionary-format dictionary, NIfTI
DWI and scheme on disk -> Evaluation -> fit -> save_results, checked against
the oracle's restatements on the same inputs:
  * the dictionary read back from disk equals the restated load_dictionary
    preprocessing of the arrays that were written (core.pyx:505-580);
  * fits: the oracle FISTA on the oracle operator (same bar as
    tests/test_demo.py: fit_details within 1e-4 absolute, identical
    iteration count; x to 1e-9);
  * lambda_max, group structure and adaptive weights (core.pyx:961-976,
    1123-1192), with streamlines left out of the dictionary;
  * the debias masks and refit (core.pyx:1476-1508);
  * the error maps, adjusted maps, compartment maps and streamline weights
    written by save_results (core.pyx:1614-1790), with a confidence map that
    holds NaNs and is rescaled (core.pyx:1433-1457);
  * set_wLasso_ISO weights and its weighted-lasso fit (core.pyx:768-858)."""
import os
import pickle

import numpy as np
import pytest

from _problems import rel
from commit_amd import nifti, synth
from commit_amd.core import Evaluation
from oracle import oracle

pytestmark = pytest.mark.gpu

DIM = (32, 32, 20)
DROPPED = (3, 50, 51, 700)        # streamline ids left out of the dictionary (TRK_kept = 0)


@pytest.fixture(scope="module")
def study(tmp_path_factory):
    root = tmp_path_factory.mktemp("study")
    raw = synth.tractogram(1001, 1500, DIM, mode=0, step_vox=0.4)
    synth.add_peaks_and_iso(raw)
    kept = np.ones(raw["TRK_norm"].size + len(DROPPED), np.uint8)
    kept[list(DROPPED)] = 0
    raw["TRK_kept"] = kept
    synth.write_dictionary(raw, str(root / "COMMIT"), n_threads=4)
    bvals, bvecs = synth.scheme((1000, 2000), 30, n_b0=2)
    synth.write_scheme(str(root / "DWI.scheme"), bvals, bvecs)
    # signal: the SZB forward model of a sparse x_true through the oracle, Rician noise
    D = synth.to_dictionary(raw)
    K = synth.normalize_kernels(synth.szb_kernels(bvals, bvecs, raw["dirs"], 1.7e-3, 0.51e-3, (1.7e-3, 3.0e-3)))
    T = oracle.set_threads(D, 4)
    R = oracle.OracleOperator(D, K, T)
    xt = synth.x_true(R.shape[1], 5)
    y = synth.rician(R.dot(xt), 30.0, 6).reshape(D["IC"]["nVOX"], -1)
    y[:, :2] = 1.0                                   # b0 volumes, so the signal normalisation is a no-op
    dwi = np.zeros(DIM + (bvals.size,), np.float32)
    dwi[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :] = y
    nifti.save(nifti.Nifti1Image(dwi), str(root / "DWI.nii.gz"))
    # confidence map with NaNs (set to 0) and a range that is rescaled to [0, 1]
    rng = np.random.default_rng(8)
    conf = rng.uniform(0.5, 3.0, DIM).astype(np.float32)
    conf[rng.random(DIM) < 0.05] = np.nan
    nifti.save(nifti.Nifti1Image(conf), str(root / "conf.nii.gz"))
    # ISO weights image (>= 1, with NaNs that become 1)
    wimg = rng.uniform(1.0, 4.0, DIM).astype(np.float32)
    wimg[rng.random(DIM) < 0.05] = np.nan
    nifti.save(nifti.Nifti1Image(wimg), str(root / "iso_w.nii.gz"))
    return root, raw


def _flow(study, max_iter=50, fit=True, **cfg):
    root = study[0]
    mit = Evaluation(str(root), ".")
    for k, v in cfg.items():
        mit.set_config(k, v)
    mit.load_data("DWI.nii.gz", "DWI.scheme")
    mit.set_model("StickZeppelinBall")
    mit.model.set(1.7e-3, [0.51e-3], [1.7e-3, 3.0e-3])
    mit.generate_kernels(regenerate=True)
    mit.load_kernels()
    mit.load_dictionary()
    mit.set_threads()
    mit.build_operator()
    if fit:
        mit.fit(tol_fun=1e-3, max_iter=max_iter)
    return mit


def _oracle_op(mit):
    return oracle.OracleOperator(mit.DICTIONARY, mit.KERNELS, oracle.set_threads(mit.DICTIONARY, 4))


def _rp(mit):
    return {k: mit.regularisation_params[k] for k in ("startIC", "sizeIC", "startEC", "sizeEC", "startISO",
                                                       "sizeISO", "regIC", "regEC", "regISO", "nnIC", "nnEC", "nnISO")}


def _same_fit(dg, do, xg, xo, tol_x=1e-9):
    print("evaluation fit", dg["iterations"], do["iterations"], rel(xg, xo), flush=True)
    assert dg["iterations"] == do["iterations"]
    assert dg["stopping_criterion"] == do["stopping_criterion"]
    for k in ("residual", "regterm", "cost_function", "abs_cost", "rel_cost", "abs_x", "rel_x"):
        assert abs(dg[k] - do[k]) < 1e-4, k
    assert rel(xg, xo) < tol_x


def _at_mask(mit, path):
    D = mit.DICTIONARY
    return nifti.load(path).data[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"]]


def test_dictionary_round_trip(study):
    """The dictionary read from disk is the restated preprocessing of the
    arrays that were written, over the mask that was written."""
    mit = _flow(study, fit=False)
    raw = study[1]
    D = mit.DICTIONARY
    tdi = np.bincount(raw["IC_v"], weights=raw["IC_len"].astype(np.float64), minlength=int(np.prod(DIM)))
    assert np.array_equal(D["MASK"].astype(bool), np.reshape(tdi, DIM, order="F") > 0)
    R = oracle.sort_and_normalize(raw, D["MASK"])
    for c, keys in (("IC", ("vox", "dir", "str", "len")), ("EC", ("vox", "dir")), ("ISO", ("vox",))):
        for k in keys:
            assert np.array_equal(D[c][k], R[c][k]), (c, k)
    assert D["IC"]["nVOX"] == R["IC"]["nVOX"] and D["IC"]["nSTR"] == raw["TRK_norm"].size
    assert np.array_equal(D["TRK"]["kept"], raw["TRK_kept"])
    for k in ("norm", "len", "lenTot"):
        assert np.array_equal(D["TRK"][k], raw["TRK_" + k])


def test_flow_matches_oracle_fit(study):
    mit = _flow(study)
    R = _oracle_op(mit)
    xo, do = oracle.solve(mit.get_y(), R, R.T, tol_fun=1e-3, tol_x=1e-6, max_iter=50,
                          regularisation=oracle.init_regularisation(_rp(mit)))
    _same_fit(mit.CONFIG["optimization"]["fit_details"], do, mit.x, xo)
    # x_nnls: the unregularised fit's coefficients, mapped back to every streamline
    xic, _, _ = mit.get_coeffs(get_normalized=False)
    assert np.array_equal(mit.x_nnls, xic) and xic.size == mit.DICTIONARY["TRK"]["kept"].size
    assert np.all(xic[list(DROPPED)] == 0)


def _groups(n_all, ng, seed):
    """ng random groups covering every streamline, then one empty group."""
    lab = np.random.default_rng(seed).integers(0, ng, n_all)
    groups = np.empty(ng + 1, dtype=object)
    for g in range(ng):
        groups[g] = np.nonzero(lab == g)[0].astype(np.int32)
    groups[ng] = np.zeros(0, np.int32)
    return groups


def test_lambda_max_and_group_weights_match_reference(study):
    mit = _flow(study, max_iter=20)
    D = mit.DICTIONARY
    kept = D["TRK"]["kept"]
    groups = _groups(kept.size, 10, 1)
    extra = np.linspace(0.5, 2.0, 10)
    mit.set_regularisation(regularisers=("group_lasso", None, "lasso"), lambdas=(0.02, None, 0.05),
                           params=({"group_idx": groups, "group_weights_extra": extra}, None, None))
    rp = mit.regularisation_params
    R = _oracle_op(mit)
    aty = R.T.dot(mit.get_y())
    g_o, w_o = oracle.group_weights(groups, kept, D["IC"]["nSTR"], extra=extra, cardinality=True, adaptive=True,
                                    x_nnls=mit.x_nnls)
    got_g = rp["dictIC_params"]["group_idx_kept"]
    assert len(got_g) == len(g_o) and all(np.array_equal(a, b) for a, b in zip(got_g, g_o))
    assert np.allclose(rp["dictIC_params"]["group_weights"], w_o, rtol=1e-14, atol=0)
    assert rp["lambdaIC_max"] == pytest.approx(oracle.lambda_max_group(aty, w_o, g_o), rel=1e-12)
    assert rp["lambdaIC"] == pytest.approx(0.02 * oracle.lambda_max_group(aty, w_o, g_o), rel=1e-12)
    lm_iso = oracle.lambda_max_lasso(aty, rp["startISO"], rp["sizeISO"], np.ones(rp["sizeISO"]))
    assert rp["lambdaISO_max"] == pytest.approx(lm_iso, rel=1e-12)
    # the regularised fit against the oracle's
    mit.fit(tol_fun=1e-3, max_iter=30)
    rpo = dict(_rp(mit), lambdaIC=rp["lambdaIC"], lambdaISO=rp["lambdaISO"],
               dictIC_params={"group_idx_kept": g_o, "group_weights": w_o})
    xo, do = oracle.solve(mit.get_y(), R, R.T, tol_fun=1e-3, tol_x=1e-6, max_iter=30,
                          regularisation=oracle.init_regularisation(rpo))
    _same_fit(mit.CONFIG["optimization"]["fit_details"], do, mit.x, xo)


def _regularised(study, lam=0.05, **cfg):
    mit = _flow(study, max_iter=20, **cfg)
    groups = _groups(mit.DICTIONARY["TRK"]["kept"].size, 10, 2)
    mit.set_regularisation(regularisers=("group_lasso", None, "lasso"), lambdas=(lam, None, lam),
                           params=({"group_idx": groups[:10]}, None, None))
    return mit


def test_debias_matches_reference(study):
    # the regularised fit's x, from a flow without debias (the fits are deterministic)
    ref = _regularised(study)
    ref.fit(tol_fun=1e-3, max_iter=20)
    x_reg = ref.x.copy()
    mit = _regularised(study)
    mit.fit(tol_fun=1e-3, max_iter=20, debias=True)
    D, K = mit.DICTIONARY, mit.KERNELS
    offset = D["IC"]["nSTR"] * K["wmr"].shape[0]
    mask = oracle.debias_eval(x_reg, offset)
    assert np.array_equal(D["IC"]["eval"], mask) and 0 < mask.sum() < offset
    Dm = dict(D, IC=dict(D["IC"], eval=mask))
    Rm = oracle.OracleOperator(Dm, K, oracle.set_threads(Dm, 4))
    ymask = oracle.debias_rows(Rm, x_reg, mask, offset)
    assert np.array_equal(mit.debias_mask, ymask) and 0 < ymask.sum() < ymask.size
    # the refit: get_y() times the row mask, no penalty, the masked operator
    y_o = mit.niiDWI_img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :].flatten().astype(np.float64) * ymask
    assert np.array_equal(mit.get_y(), y_o)
    xo, do = oracle.solve(y_o, Rm, Rm.T, tol_fun=1e-3, tol_x=1e-6, max_iter=20,
                          regularisation=oracle.init_regularisation(_rp(mit)))
    assert mit.regularisation_params["regIC"] is None and mit.regularisation_params["regISO"] is None
    _same_fit(mit.CONFIG["optimization"]["fit_details"], do, mit.x, xo)


def _check_maps(mit, conf_values=None):
    out = mit.get_config("RESULTS_path")
    D, K = mit.DICTIONARY, mit.KERNELS
    nVOX, nS = D["IC"]["nVOX"], mit.niiDWI_img.shape[3]
    R = _oracle_op(mit)
    y_dwi = mit.niiDWI_img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :].flatten()
    e = oracle.fit_errors(y_dwi, mit.get_y(), R.dot(mit.x), nVOX, nS, debias_mask=mit.debias_mask, conf=conf_values)
    stats = mit.temp_data["fit_stats"]
    for name in ("RMSE", "NRMSE"):
        got = _at_mask(mit, os.path.join(out, f"fit_{name}.nii.gz"))
        assert np.allclose(got, e[name], rtol=1e-5, atol=1e-7), name
        assert stats[name][0] == pytest.approx(e[name + "_stats"][0], rel=1e-5)
        assert stats[name][1] == pytest.approx(e[name + "_stats"][1], rel=1e-4, abs=1e-7)
    if conf_values is not None:
        for name in ("RMSE_adjusted", "NRMSE_adjusted"):
            got = _at_mask(mit, os.path.join(out, f"fit_{name}.nii.gz"))
            assert np.allclose(got, e[name], rtol=1e-5, atol=1e-7), name
    maps = oracle.compartment_maps(mit.x, D, K, nVOX, normalize_kernels=bool(mit.get_config("doNormalizeKernels")),
                                   normalize_maps=bool(mit.get_config("doNormalizeMaps")))
    for name, m in zip(("IC", "EC", "ISO"), maps):
        got = _at_mask(mit, os.path.join(out, f"compartment_{name}.nii.gz"))
        assert np.allclose(got, m, rtol=1e-6, atol=1e-12), name
    # streamline weights: renormalised IC coefficients summed over RFs, 0 for dropped streamlines
    # (core.pyx:1542-1548, 1763-1790)
    w = np.loadtxt(os.path.join(out, "streamline_weights.txt"))
    x_ren = mit.x[:D["IC"]["nSTR"] * K["wmr"].shape[0]] / np.kron(K["wmr_norm"], D["TRK"]["norm"])
    exp = np.zeros(D["TRK"]["kept"].size)
    exp[D["TRK"]["kept"] == 1] = x_ren.reshape(K["wmr"].shape[0], -1).sum(axis=0)
    assert w.shape == exp.shape and np.allclose(w, exp, rtol=1e-5, atol=1e-12)
    with open(os.path.join(out, "results.pickle"), "rb") as f:
        cfg, x_pre, x_post = pickle.load(f)
    assert np.array_equal(x_pre, mit.x)
    assert cfg["optimization"]["fit_details"] == mit.CONFIG["optimization"]["fit_details"]


def test_save_results_matches_reference(study):
    mit = _flow(study, max_iter=20, doNormalizeMaps=True)
    mit.save_results()
    _check_maps(mit)


def test_save_results_after_debias_with_confidence_map(study):
    mit = _regularised(study, lam=0.002)
    mit.fit(tol_fun=1e-3, max_iter=20, confidence_map_filename="conf.nii.gz", confidence_map_rescale=True,
            debias=True)
    assert 0 < mit.debias_mask.sum() < mit.debias_mask.size
    D = mit.DICTIONARY
    nS = mit.niiDWI_img.shape[3]
    raw_conf = nifti.load(os.path.join(str(study[0]), "conf.nii.gz")).data.astype(np.float32)
    raw_conf = np.repeat(raw_conf[:, :, :, None], nS, axis=3)[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :].flatten()
    conf, changed = oracle.confidence(raw_conf, rescale=True)
    assert changed
    # the processed values replace the map at the mask voxels
    got = mit.confidence_map_img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :].flatten()
    assert np.array_equal(got, conf)
    mit.save_results()
    _check_maps(mit, conf_values=conf)


def test_set_wlasso_iso_matches_reference(study):
    mit = _flow(study, max_iter=20)
    path = os.path.join(str(study[0]), "iso_w.nii.gz")
    mit.set_wLasso_ISO(path, 0.1)
    rp = mit.regularisation_params
    D, K = mit.DICTIONARY, mit.KERNELS
    w = oracle.iso_weights(nifti.load(path).data, D, K["iso"].shape[0])
    assert np.array_equal(np.asarray(rp["dictISO_params"]["coeff_weights"], np.float32), w)
    R = _oracle_op(mit)
    aty = R.T.dot(mit.get_y())
    lm = oracle.lambda_max_lasso(aty, rp["startISO"], rp["sizeISO"], w)
    assert rp["lambdaISO_max"] == pytest.approx(lm, rel=1e-12)
    mit.fit(tol_fun=1e-3, max_iter=30)
    rpo = dict(_rp(mit), lambdaISO=rp["lambdaISO"], dictISO_params={"coeff_weights": w.astype(np.float64)})
    xo, do = oracle.solve(mit.get_y(), R, R.T, tol_fun=1e-3, tol_x=1e-6, max_iter=30,
                          regularisation=oracle.init_regularisation(rpo))
    _same_fit(mit.CONFIG["optimization"]["fit_details"], do, mit.x, xo)
