"""End-to-end drop-in flow on files: trk2dictionary-format dictionary, NIfTI
DWI and scheme on disk -> Evaluation -> fit -> save_results, checked against
the oracle FISTA on the same operator inputs (same bar as tests/test_demo.py:
fit_details within 1e-4 absolute, identical iteration count)."""
import os
import pickle

import numpy as np
import pytest

from _problems import rel
from commit_amd import nifti, synth
from commit_amd.core import Evaluation
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def study(tmp_path_factory):
    root = tmp_path_factory.mktemp("study")
    raw = synth.tractogram(1001, 1500, (32, 32, 20), mode=0, step_vox=0.4)
    synth.add_peaks_and_iso(raw)
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
    dwi = np.zeros((32, 32, 20, bvals.size), np.float32)
    dwi[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :] = y
    nifti.save(nifti.Nifti1Image(dwi), str(root / "DWI.nii.gz"))
    return root


def _flow(root, max_iter=50):
    mit = Evaluation(str(root), ".")
    mit.load_data("DWI.nii.gz", "DWI.scheme")
    mit.set_model("StickZeppelinBall")
    mit.model.set(1.7e-3, [0.51e-3], [1.7e-3, 3.0e-3])
    mit.generate_kernels(regenerate=True)
    mit.load_kernels()
    mit.load_dictionary()
    mit.set_threads()
    mit.build_operator()
    mit.fit(tol_fun=1e-3, max_iter=max_iter)
    return mit


def test_flow_matches_oracle_fit(study):
    mit = _flow(study)
    D, K = mit.DICTIONARY, mit.KERNELS
    T = oracle.set_threads(D, 4)
    R = oracle.OracleOperator(D, K, T)
    rp = {k: mit.regularisation_params[k] for k in ("startIC", "sizeIC", "startEC", "sizeEC", "startISO",
                                                      "sizeISO", "regIC", "regEC", "regISO", "nnIC", "nnEC", "nnISO")}
    xo, do = oracle.solve(mit.get_y(), R, R.T, tol_fun=1e-3, tol_x=1e-6, max_iter=50,
                          regularisation=oracle.init_regularisation(rp))
    dg = mit.CONFIG["optimization"]["fit_details"]
    print("evaluation fit", dg["iterations"], do["iterations"], rel(mit.x, xo), flush=True)
    assert dg["iterations"] == do["iterations"]
    for k in ("residual", "regterm", "cost_function", "abs_cost", "rel_cost", "abs_x", "rel_x"):
        assert abs(dg[k] - do[k]) < 1e-4, k
    assert rel(mit.x, xo) < 1e-9


def test_save_results_outputs(study):
    mit = _flow(study, max_iter=20)
    stats = mit.save_results()
    out = mit.get_config("RESULTS_path")
    for f in ("fit_RMSE.nii.gz", "fit_NRMSE.nii.gz", "compartment_IC.nii.gz", "compartment_EC.nii.gz",
              "compartment_ISO.nii.gz", "streamline_weights.txt", "results.pickle"):
        assert os.path.exists(os.path.join(out, f)), f
    w = np.loadtxt(os.path.join(out, "streamline_weights.txt"))
    xic, _, _ = mit.get_coeffs()
    assert w.shape == (mit.DICTIONARY["TRK"]["kept"].size,)
    assert np.allclose(w, xic, rtol=1e-5, atol=1e-12)
    with open(os.path.join(out, "results.pickle"), "rb") as f:
        cfg, x_pre, x_ren = pickle.load(f)
    assert np.array_equal(x_pre, mit.x)
    assert cfg["optimization"]["fit_details"]["iterations"] >= 1
    rm = nifti.load(os.path.join(out, "fit_RMSE.nii.gz")).data
    assert rm.shape == (32, 32, 20) and np.isfinite(rm).all()
    assert stats["NRMSE"][0] < 0.5


def test_regularised_fit_and_debias(study):
    mit = _flow(study, max_iter=20)
    groups = np.empty(10, dtype=object)
    lab = np.random.default_rng(1).integers(0, 10, mit.DICTIONARY["TRK"]["kept"].size)
    for g in range(10):
        groups[g] = np.nonzero(lab == g)[0]
    mit.set_regularisation(regularisers=("group_lasso", None, "lasso"), lambdas=(0.02, None, 0.05),
                           params=({"group_idx": groups}, None, None))
    rp = mit.regularisation_params
    assert rp["lambdaIC"] == pytest.approx(0.02 * rp["lambdaIC_max"])
    mit.fit(tol_fun=1e-3, max_iter=20, debias=True)
    assert mit.debias_mask is not None and mit.x.shape[0] == mit.A.shape[1]
