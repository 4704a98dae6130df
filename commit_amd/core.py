"""commit.Evaluation drop-in (core.pyx:60-1837) on top of the MI355X operator.

The flow and method names are the reference's:

    mit = Evaluation(study_path, subject, dictionary_path='COMMIT')
    mit.load_data(dwi_filename, scheme_filename)
    mit.set_model('StickZeppelinBall'); mit.model.set(d_par, d_perps, d_isos)
    mit.generate_kernels(regenerate=True); mit.load_kernels()
    mit.load_dictionary(); mit.set_threads(); mit.build_operator()
    mit.set_regularisation(...)            # optional
    mit.fit(tol_fun=1e-3, max_iter=1000)
    mit.save_results()

Differences, by design:
  * the operator and the FISTA solver run on the GPU (commit_amd.operator,
    commit_amd.solvers); ``set_threads`` is kept for the interface (the
    device plan builds its own work partition);
  * kernel generation uses closed-form Stick / Zeppelin / Ball response
    functions on a fixed set of hemisphere directions instead of dmri-amico's
    rotated high-resolution kernels (dmri-amico is not available);
  * NIfTI I/O goes through a minimal built-in reader/writer
    (commit_amd.nifti) instead of nibabel; errors raise ValueError /
    RuntimeError where the reference logs an error and exits.
"""
import glob
import os
import pickle
import time
from os.path import exists, isfile, join as pjoin

import numpy as np

from . import _lib, nifti, solvers, synth
from .operator import operator

VERSION = "commit_amd 0.1"


def setup(lmax=12):
    """Kept for interface compatibility (core.pyx:27-35); nothing to set up."""
    return None


def load_dictionary_info(filename):
    """dictionary_info.pickle written next to the dictionary (core.pyx:38-57)."""
    if not isfile(filename):
        raise RuntimeError('Dictionary is outdated or not found. Execute "trk2dictionary" script first.')
    with open(filename, "rb") as f:
        aux = pickle.load(f, fix_imports=True, encoding="bytes")
    return {(k.decode() if hasattr(k, "decode") else k): v for k, v in aux.items()}


class _Scheme:
    """Acquisition scheme: b-values, gradient directions, b0 / DWI indices."""

    def __init__(self, bvals, bvecs, b0_thr=0.0):
        self.b = np.asarray(bvals, dtype=np.float64)
        self.raw = np.column_stack([np.asarray(bvecs, dtype=np.float64), self.b])
        self.nS = self.b.size
        self.b0_idx = np.nonzero(self.b <= b0_thr)[0]
        self.dwi_idx = np.nonzero(self.b > b0_thr)[0]
        self.b0_count = self.b0_idx.size
        self.dwi_count = self.dwi_idx.size
        self.shells = [{"b": b, "idx": np.nonzero(np.isclose(self.b, b))[0]} for b in np.unique(self.b[self.dwi_idx])]

    @classmethod
    def from_file(cls, path, b0_thr=0.0):
        """Text scheme: optional 'VERSION: ...' header, then rows 'gx gy gz b'."""
        rows = []
        with open(path) as f:
            for line in f:
                s = line.strip()
                if not s or s.upper().startswith("VERSION"):
                    continue
                rows.append([float(v) for v in s.split()[:4]])
        a = np.array(rows)
        return cls(a[:, 3], a[:, :3], b0_thr)


class StickZeppelinBall:
    """Closed-form SZB model: Stick (IC), Zeppelin (EC, one per d_perp), Ball (ISO)."""

    def __init__(self):
        self.id = "StickZeppelinBall"
        self.name = "Stick-Zeppelin-Ball"
        self.d_par = 1.7e-3
        self.d_perps = [0.6e-3]
        self.d_isos = [1.7e-3, 3.0e-3]
        self.scheme = None

    def set(self, d_par, d_perps, d_isos):
        self.d_par = float(d_par)
        self.d_perps = list(d_perps)
        self.d_isos = list(d_isos)

    def get_params(self):
        return {"id": self.id, "name": self.name, "d_par": self.d_par, "d_perps": self.d_perps,
                "d_isos": self.d_isos}

    def kernels(self, dirs):
        sch = self.scheme
        K = synth.szb_kernels(sch.b, sch.raw[:, :3], dirs, self.d_par, None, self.d_isos)
        if self.d_perps:
            c2 = (dirs @ sch.raw[:, :3].T) ** 2
            b = sch.b[None, :]
            K["wmh"] = np.ascontiguousarray(np.stack(
                [np.exp(-b * (dp + (self.d_par - dp) * c2)) for dp in self.d_perps]).astype(np.float32))
        K["model"] = self.id
        return K


class ScalarMap:
    """All-ones IC kernel, one direction (models.pyx:33-82)."""

    def __init__(self):
        self.id = "ScalarMap"
        self.name = "ScalarMap"
        self.lesion_mask = False
        self.scheme = None

    def set(self):
        return None

    def get_params(self):
        return {"id": self.id, "name": self.name}

    def kernels(self, dirs):
        nS = self.scheme.nS
        return {"model": self.id, "wmr": np.ones((1, 1, nS), dtype=np.float32),
                "wmh": np.ones((0, 1, nS), dtype=np.float32),
                "iso": (-1 * np.ones((1, nS), dtype=np.float32)) if self.lesion_mask
                else np.ones((0, nS), dtype=np.float32)}


_MODELS = {"StickZeppelinBall": StickZeppelinBall, "ScalarMap": ScalarMap, "VolumeFractions": ScalarMap}


class Evaluation:
    """Holds the data and parameters of one COMMIT evaluation (core.pyx:60-135)."""

    def __init__(self, study_path=".", subject=".", dictionary_path="COMMIT"):
        self.niiDWI = None
        self.niiDWI_img = None
        self.scheme = None
        self.model = None
        self.KERNELS = None
        self.DICTIONARY = None
        self.dictionary_info = None
        self.THREADS = None
        self.A = None
        self.regularisation_params = None
        self.x = None
        self.confidence_map_img = None
        self.debias_mask = None
        self.x_nnls = None
        self.verbose = 3
        self.CONFIG = {}
        self.temp_data = {}
        self.set_config("version", VERSION)
        self.set_config("study_path", study_path)
        self.set_config("subject", subject)
        self.set_config("dictionary_path", dictionary_path)
        if study_path == subject:
            self.set_config("DATA_path", study_path)
            self.set_config("TRACKING_path", pjoin(study_path, dictionary_path))
        else:
            self.set_config("DATA_path", pjoin(study_path, subject))
            self.set_config("TRACKING_path", pjoin(study_path, subject, dictionary_path))
        info = pjoin(self.get_config("TRACKING_path"), "dictionary_info.pickle")
        self.dictionary_info = load_dictionary_info(info) if isfile(info) else \
            {"ndirs": 500, "n_threads": os.cpu_count() or 1, "iso_mask": False,
             "blur_gauss_extent": 0, "blur_core_extent": 0}
        self.set_config("doNormalizeSignal", True)
        self.set_config("doMergeB0", False)
        self.set_config("doNormalizeKernels", True)
        self.set_config("doDemean", False)
        self.set_config("doNormalizeMaps", False)
        self.set_config("sortDictionaryWith", "mergesort")
        self.set_config("sortThreadsWith", "mergesort")
        # vector storage of the device operator and FISTA: "fp64" (the
        # reference's float64) or "fp32" (float32 vectors, FP64 accumulation,
        # FP64 stopping norms and final residual); no counterpart in the reference
        self.set_config("precision", "fp64")

    # ------------------------------------------------------------- config --
    def set_verbose(self, verbose):
        self.verbose = verbose

    def set_config(self, key, value):
        self.CONFIG[key] = value
        self.temp_data[key] = value

    def get_config(self, key):
        return self.CONFIG.get(key)

    def get_temp_data(self, key):
        return self.temp_data.get(key)

    # --------------------------------------------------------------- data --
    def load_data(self, dwi_filename, scheme_filename=None, b0_thr=0, b0_min_signal=0, replace_bad_voxels=None):
        """DWI volume + scheme, with the reference's preprocessing
        (core.pyx:163-278): b0 normalisation, optional b0 merge and demean."""
        img = nifti.load(pjoin(self.get_config("DATA_path"), dwi_filename))
        if scheme_filename is not None:
            self.set_config("scheme_filename", scheme_filename)
            self.set_config("b0_thr", b0_thr)
            scheme = _Scheme.from_file(pjoin(self.get_config("DATA_path"), scheme_filename), b0_thr)
        else:
            scheme = _Scheme([1000.0], [[0.0, 0.0, 0.0]], 0)
        self.load_data_from_arrays(img.data, scheme, img.affine, img.pixdim, b0_min_signal, replace_bad_voxels,
                                   dwi_filename=dwi_filename)

    def load_data_from_arrays(self, dwi, scheme, affine=None, pixdim=(1.0, 1.0, 1.0), b0_min_signal=0,
                              replace_bad_voxels=None, dwi_filename="<memory>"):
        """In-memory load_data: dwi (X,Y,Z,nS) array, scheme (_Scheme or (bvals, bvecs))."""
        if not isinstance(scheme, _Scheme):
            scheme = _Scheme(scheme[0], scheme[1], self.get_config("b0_thr") or 0)
        self.scheme = scheme
        self.set_config("dwi_filename", dwi_filename)
        self.set_config("b0_min_signal", b0_min_signal)
        self.set_config("replace_bad_voxels", replace_bad_voxels)
        img = np.asarray(dwi).astype(np.float32)
        if img.ndim == 3:
            img = img[..., None]
        self.niiDWI = nifti.Nifti1Image(img[..., :1], affine, pixdim)
        self.set_config("dim", img.shape[:3])
        self.set_config("pixdim", tuple(float(p) for p in pixdim))
        if scheme.nS != img.shape[3]:
            raise ValueError("Scheme does not match with input data")
        if scheme.dwi_count == 0:
            raise ValueError("There are no DWI volumes in the data")
        if not np.isfinite(img).all():
            if replace_bad_voxels is None:
                raise ValueError('Nan or Inf values in the raw signal. Try using "replace_bad_voxels"')
            np.nan_to_num(img, copy=False, nan=replace_bad_voxels, posinf=replace_bad_voxels,
                          neginf=replace_bad_voxels)
        if self.get_config("doNormalizeSignal") and scheme.b0_count > 0:
            b0 = np.mean(img[:, :, :, scheme.b0_idx], axis=3)
            idx = b0 <= b0_min_signal * b0[b0 > 0].mean()
            b0[idx] = 1
            b0 = 1.0 / b0
            b0[idx] = 0
            img *= b0[..., None]
        if scheme.b0_count > 1 and self.get_config("doMergeB0"):
            mean = np.mean(img[:, :, :, scheme.b0_idx], axis=3, keepdims=True)
            img = np.concatenate((mean, img[:, :, :, scheme.dwi_idx]), axis=3)
            keep = np.concatenate(([scheme.b0_idx[0]], scheme.dwi_idx))
            self.scheme = _Scheme(scheme.b[keep], scheme.raw[keep, :3], self.get_config("b0_thr") or 0)
        if self.get_config("doDemean"):
            img = img - img.mean(axis=3, keepdims=True)
        self.niiDWI_img = img

    # ------------------------------------------------------------ kernels --
    def set_model(self, model_name):
        if model_name not in _MODELS:
            raise ValueError(f'Model "{model_name}" not recognized')
        self.model = _MODELS[model_name]()
        if self.dictionary_info.get("iso_mask") and self.model.id == "ScalarMap":
            self.model.lesion_mask = True
        self.set_config("ATOMS_path", pjoin(self.get_config("study_path"), "kernels", self.model.id))

    def generate_kernels(self, regenerate=False, lmax=12, ndirs=500):
        """Response functions on ndirs hemisphere directions (closed form)."""
        if self.scheme is None:
            raise RuntimeError('Scheme not loaded; call "load_data()" first')
        if self.model is None:
            raise RuntimeError('Model not set; call "set_model()" method first')
        if self.model.id == "ScalarMap":
            ndirs = 1
        self.set_config("lmax", lmax)
        self.set_config("ndirs", ndirs)
        self.set_config("model", self.model.get_params())
        self.model.scheme = self.scheme
        self._kernels_raw = self.model.kernels(synth.hemisphere_directions(ndirs))

    def load_kernels(self):
        """Normalisation / demean of the LUTs (core.pyx:364-433)."""
        if getattr(self, "_kernels_raw", None) is None:
            raise RuntimeError('Kernels not generated; call "generate_kernels()" first')
        K = {k: (np.ascontiguousarray(v, dtype=np.float32).copy() if isinstance(v, np.ndarray) else v)
             for k, v in self._kernels_raw.items()}
        nIC, nEC, nISO = K["wmr"].shape[0], K["wmh"].shape[0], K["iso"].shape[0]
        if self.get_config("doDemean"):
            K["wmr"] -= K["wmr"].mean(axis=2, keepdims=True)
            K["wmh"] -= K["wmh"].mean(axis=2, keepdims=True)
            K["iso"] -= K["iso"].mean(axis=1, keepdims=True)
        if self.get_config("doNormalizeKernels"):
            K["wmr_norm"] = np.array([np.linalg.norm(K["wmr"][i, 0, :]) for i in range(nIC)])
            for i in range(nIC):
                K["wmr"][i] /= K["wmr_norm"][i]
            K["wmh_norm"] = np.array([np.linalg.norm(K["wmh"][i, 0, :]) for i in range(nEC)])
            for i in range(nEC):
                K["wmh"][i] /= K["wmh_norm"][i]
            K["iso_norm"] = np.array([np.linalg.norm(K["iso"][i, :]) for i in range(nISO)])
            for i in range(nISO):
                K["iso"][i] /= K["iso_norm"][i]
        self.KERNELS = K

    def load_kernels_from_arrays(self, KERNELS, normalized=True):
        """Use precomputed LUTs (wmr, wmh, iso [+ *_norm when already normalised])."""
        if normalized:
            self.KERNELS = dict(KERNELS)
            self.set_config("ndirs", int(KERNELS["wmr"].shape[1]))
        else:
            self._kernels_raw = dict(KERNELS)
            self.set_config("ndirs", int(KERNELS["wmr"].shape[1]))
            self.load_kernels()

    # --------------------------------------------------------- dictionary --
    def load_dictionary(self, path=None, use_all_voxels_in_mask=False):
        """Read the trk2dictionary files (core.pyx:436-582): voxel-sorted IC/EC/ISO,
        lengths divided by TRK_norm, voxel ids compacted over the mask."""
        path = self.get_config("TRACKING_path")
        if self.get_config("ndirs") is not None and self.dictionary_info["ndirs"] != self.get_config("ndirs"):
            raise ValueError(f'"ndirs" of the dictionary ({self.dictionary_info["ndirs"]}) does not match '
                             f'with kernels ({self.get_config("ndirs")})')
        self.set_config("dictionary_mask", "mask" if use_all_voxels_in_mask else "tdi")
        mname = pjoin(path, f'dictionary_{self.get_config("dictionary_mask")}.nii')
        if not exists(mname):
            mname += ".gz"
            if not exists(mname):
                raise RuntimeError('Dictionary not found. Execute "trk2dictionary" script first')
        mask = (np.asarray(nifti.load(mname).data) > 0).astype(np.uint8)
        rd = lambda name, dt: np.fromfile(pjoin(path, f"dictionary_{name}.dict"), dtype=dt)
        raw = {"IC_f": rd("IC_f", np.uint32), "IC_v": rd("IC_v", np.uint32), "IC_o": rd("IC_o", np.uint16),
               "IC_len": rd("IC_len", np.float32), "TRK_norm": rd("TRK_norm", np.float32),
               "TRK_len": rd("TRK_len", np.float32), "TRK_lenTot": rd("TRK_lenTot", np.float32),
               "TRK_kept": rd("TRK_kept", np.uint8), "EC_v": rd("EC_v", np.uint32), "EC_o": rd("EC_o", np.uint16),
               "ISO_v": rd("ISO_v", np.uint32), "dim": mask.shape,
               "dirs": np.zeros((int(self.dictionary_info["ndirs"]), 3))}
        self.load_dictionary_from_arrays(raw, mask)

    def load_dictionary_from_arrays(self, raw, mask=None):
        """load_dictionary on in-memory trk2dictionary arrays (see synth.tractogram)."""
        t0 = time.time()
        if mask is None:
            D = synth.to_dictionary(raw, normalize=bool(self.get_config("doNormalizeKernels")))
        else:
            D = synth.to_dictionary(raw, normalize=bool(self.get_config("doNormalizeKernels")),
                                    use_all_voxels_in_mask=True, mask=mask)
        D["ndirs"] = int(self.dictionary_info.get("ndirs", raw["dirs"].shape[0]))
        D["n_threads"] = int(self.dictionary_info.get("n_threads", 1))
        self.DICTIONARY = D
        self.temp_data["load_dictionary_time"] = time.time() - t0

    def set_threads(self, n=None):
        """Interface parity with core.pyx:586-724; the device plan partitions
        the work itself, so only the argument check remains."""
        if n is None:
            n = self.DICTIONARY["n_threads"] if self.DICTIONARY else 1
        if n < 1 or n > 255:
            raise ValueError("Number of threads must be between 1 and 255")
        if self.DICTIONARY is None:
            raise RuntimeError('Dictionary not loaded; call "load_dictionary()" first')
        if self.KERNELS is None:
            raise RuntimeError('Response functions not generated; call "generate_kernels()" and "load_kernels()" first')
        self.THREADS = {"n": n}

    def build_operator(self):
        """Device-resident operator (core.pyx:727-748)."""
        if self.DICTIONARY is None:
            raise RuntimeError('Dictionary not loaded; call "load_dictionary()" first')
        if self.KERNELS is None:
            raise RuntimeError('Response functions not generated; call "generate_kernels()" and "load_kernels()" first')
        if self.THREADS is None:
            raise RuntimeError('Threads not set; call "set_threads()" first')
        if self.DICTIONARY["IC"]["nSTR"] <= 0:
            raise ValueError("No streamlines found in the dictionary; check your data")
        if self.DICTIONARY["EC"]["n"] <= 0 and self.KERNELS["wmh"].shape[0] > 0:
            raise ValueError("The selected model has EC compartments, but no peaks have been provided")
        self.DICTIONARY["IC"]["eval"] = np.ones(int(self.DICTIONARY["IC"]["nSTR"] * self.KERNELS["wmr"].shape[0]),
                                                dtype=np.uint32)
        nolut = getattr(self.model, "nolut", False) if self.model is not None else False
        prec = str(self.get_config("precision") or "fp64").lower()
        if prec not in ("fp64", "fp32"):
            raise ValueError('CONFIG["precision"] must be "fp64" or "fp32"')
        self.A = operator.LinearOperator(self.DICTIONARY, self.KERNELS, self.THREADS, nolut,
                                         precision=_lib.FP32 if prec == "fp32" else _lib.FP64)
        if self.get_config("autotune") is not False:
            # launch options of the products (LDS slabs, dense-tile A, XCD order) timed on this
            # dictionary; CONFIG["autotune"] = False keeps the defaults
            from . import autotune
            autotune.tune(self.A, nrhs=0, reps=2, multi=False)

    def get_y(self):
        """Measured signal at the mask voxels, voxel-major (core.pyx:751-766)."""
        if self.DICTIONARY is None:
            raise RuntimeError('Dictionary not loaded; call "load_dictionary()" first')
        if self.niiDWI_img is None:
            raise RuntimeError('Data not loaded; call "load_data()" first')
        D = self.DICTIONARY
        y = self.niiDWI_img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :].flatten().astype(np.float64)
        if self.debias_mask is not None:
            y *= self.debias_mask
        return y

    # ----------------------------------------------------- regularisation --
    def set_regularisation(self, regularisers=(None, None, None), lambdas=(None, None, None),
                           is_nonnegative=(True, True, True), params=(None, None, None)):
        """Penalties per compartment, lambdas as fractions of lambda_max
        (core.pyx:861-1350)."""
        for name in ("niiDWI_img", "DICTIONARY", "KERNELS", "THREADS", "A"):
            if getattr(self, name) is None:
                raise RuntimeError(f"{name} not set; run the previous steps of the flow first")
        D, K = self.DICTIONARY, self.KERNELS
        aty_cache = {}

        def aty():
            if "v" not in aty_cache:
                aty_cache["v"] = np.asarray(self.A.T.dot(self.get_y()))
            return aty_cache["v"]

        def lambda_max_lasso(start, size, w_coeff):   # Kim et al. 2007
            return np.max(np.abs(aty()[start:start + size]) / w_coeff)

        def lambda_max_group(w_group, idx_group):     # Yuan & Lin 2006
            a = aty()
            return np.max([np.sqrt(np.sum(a[idx_group[g]] ** 2)) / w_group[g] for g in range(w_group.size)])

        reg = {"startIC": 0, "sizeIC": int(D["IC"]["nSTR"] * K["wmr"].shape[0])}
        reg["startEC"] = reg["sizeIC"]
        reg["sizeEC"] = int(D["EC"]["n"] * K["wmh"].shape[0])
        reg["startISO"] = reg["sizeIC"] + reg["sizeEC"]
        reg["sizeISO"] = int(D["ISO"]["n"] * K["iso"].shape[0])
        reg["regIC"], reg["regEC"], reg["regISO"] = regularisers
        reg["nnIC"], reg["nnEC"], reg["nnISO"] = is_nonnegative
        dIC, dEC, dISO = params
        for r in regularisers:
            if r not in solvers.list_regularizers:
                raise ValueError(f'Regulariser "{r}" not recognised')
        kept = D["TRK"]["kept"]

        # IC
        rIC = reg["regIC"]
        if rIC is not None:
            lam = lambdas[0]
            if rIC == "sparse_group_lasso":
                if lam is None or len(lam) != 2:
                    raise ValueError("Regularisation parameters for the IC compartment are not exactly two")
                if lam[0] < 0 or lam[1] < 0:
                    raise ValueError("Regularisation parameters for the IC compartment must be greater than 0")
            else:
                if lam is None:
                    raise ValueError("Missing regularisation parameter for the IC compartment")
                if lam < 0:
                    raise ValueError("Regularisation parameter for the IC compartment must be greater than 0")
            reg["lambdaIC_perc"] = lam
            if rIC in ("group_lasso", "sparse_group_lasso") and (dIC is None or dIC.get("group_idx") is None):
                raise ValueError("Group structure for the IC compartment not provided")
        if rIC in ("lasso", "sparse_group_lasso") and dIC is not None and "coeff_weights" in dIC:
            cw = np.asarray(dIC["coeff_weights"], dtype=np.float64)
            if np.any(cw < 0):
                raise ValueError("All coefficients weights must be non-negative")
            if cw.size != kept.size:
                raise ValueError(f'"coeff_weights" must have the same size as the number of elements in the '
                                 f'IC compartment (got {cw.size} but {kept.size} expected)')
            dIC["coeff_weights_kept"] = cw[kept == 1]
        if rIC in ("group_lasso", "sparse_group_lasso"):
            dIC.setdefault("group_weights_cardinality", True)
            dIC.setdefault("group_weights_adaptive", True)
            for key in ("group_weights_cardinality", "group_weights_adaptive"):
                v = dIC[key]
                if isinstance(v, str):
                    if v.lower() not in ("true", "false"):
                        raise ValueError(f'"{key}" must be a boolean or a string')
                    dIC[key] = v.lower() == "true"
                elif not isinstance(v, (bool, np.bool_)):
                    raise ValueError(f'"{key}" must be a boolean or a string')
            groups = [np.asarray(g, dtype=np.int32) for g in dIC["group_idx"]]
            allin = np.sort(np.unique(np.concatenate(groups))) if groups else np.zeros(0, np.int32)
            if allin.size != kept.size or np.any(allin != np.arange(kept.size)):
                raise ValueError("Group indices must contain all the indices of the input streamlines")
            groups = [g for g in groups if g.size > 0]
            if "group_weights_extra" in dIC:
                wx = np.asarray(dIC["group_weights_extra"], dtype=np.float64)
                if np.any(wx < 0):
                    raise ValueError("All group weights must be non-negative")
                if wx.size != len(groups):
                    raise ValueError("Group weights and group indices must have the same size")
                wg = wx.copy()
            else:
                wg = np.ones(len(groups), dtype=np.float64)
            if 0 in kept:
                idx_in_kept = np.zeros(kept.size, dtype=np.int32) - 1
                idx_in_kept[kept == 1] = np.arange(D["IC"]["nSTR"])
                new_g, new_w = [], []
                for c, g in enumerate(groups):
                    g = idx_in_kept[g]
                    g = g[g != -1]
                    if g.size > 0:
                        new_g.append(g)
                        new_w.append(wg[c])
                groups, wg = new_g, np.array(new_w, dtype=np.float64)
            if dIC["group_weights_cardinality"]:
                wg = wg * np.sqrt(np.array([g.size for g in groups], dtype=np.int32))
            if dIC["group_weights_adaptive"]:
                if self.x_nnls is None:
                    raise RuntimeError("Group weights cannot be computed if the fit without regularisation "
                                       "has not been performed before")
                wg = wg / np.array([np.linalg.norm(self.x_nnls[g]) + 1e-12 for g in groups], dtype=np.float64)
            gk = np.empty(len(groups), dtype=object)
            for i, g in enumerate(groups):
                gk[i] = g
            dIC["group_idx_kept"] = gk
            dIC["group_weights"] = wg
        reg["dictIC_params"] = dIC
        if rIC == "lasso":
            w = dIC["coeff_weights_kept"] if dIC is not None and "coeff_weights_kept" in dIC \
                else np.ones(reg["sizeIC"])
            reg["lambdaIC_max"] = lambda_max_lasso(reg["startIC"], reg["sizeIC"], w)
            reg["lambdaIC"] = reg["lambdaIC_perc"] * reg["lambdaIC_max"]
        elif rIC == "group_lasso":
            reg["lambdaIC_max"] = lambda_max_group(dIC["group_weights"], dIC["group_idx_kept"])
            reg["lambdaIC"] = reg["lambdaIC_perc"] * reg["lambdaIC_max"]
        elif rIC == "sparse_group_lasso":
            w = dIC["coeff_weights_kept"] if "coeff_weights_kept" in dIC else np.ones(reg["sizeIC"])
            reg["lambdaIC_max"] = (lambda_max_lasso(reg["startIC"], reg["sizeIC"], w),
                                   lambda_max_group(dIC["group_weights"], dIC["group_idx_kept"]))
            reg["lambdaIC"] = (reg["lambdaIC_perc"][0] * reg["lambdaIC_max"][0],
                               reg["lambdaIC_perc"][1] * reg["lambdaIC_max"][1])
        # EC
        if reg["regEC"] in ("group_lasso", "sparse_group_lasso"):
            raise ValueError("Not yet implemented")
        if reg["regEC"] == "lasso":
            if reg["sizeEC"] == 0:
                raise ValueError("No extracellular compartment found in the dictionary.")
            if lambdas[1] is None or lambdas[1] < 0:
                raise ValueError("Missing or negative regularisation parameter for the EC compartment")
            reg["lambdaEC_perc"] = lambdas[1]
            reg["lambdaEC_max"] = lambda_max_lasso(reg["startEC"], reg["sizeEC"], np.ones(reg["sizeEC"]))
            reg["lambdaEC"] = reg["lambdaEC_perc"] * reg["lambdaEC_max"]
        reg["dictEC_params"] = dEC
        # ISO
        if reg["regISO"] in ("group_lasso", "sparse_group_lasso"):
            raise ValueError("Not yet implemented")
        if reg["regISO"] == "lasso":
            if reg["sizeISO"] == 0:
                raise ValueError("No isotropic compartment found in the dictionary.")
            if lambdas[2] is None or lambdas[2] < 0:
                raise ValueError("Missing or negative regularisation parameter for the ISO compartment")
            reg["lambdaISO_perc"] = lambdas[2]
            if dISO is not None and "coeff_weights" in dISO:
                if np.asarray(dISO["coeff_weights"]).size != reg["sizeISO"]:
                    raise ValueError('"coeff_weights" must have the same size as the number of elements in the '
                                     'ISO compartment')
                w = np.asarray(dISO["coeff_weights"], dtype=np.float64)
            else:
                w = np.ones(reg["sizeISO"])
            reg["lambdaISO_max"] = lambda_max_lasso(reg["startISO"], reg["sizeISO"], w)
            reg["lambdaISO"] = reg["lambdaISO_perc"] * reg["lambdaISO_max"]
        reg["dictISO_params"] = dISO
        self.regularisation_params = solvers.init_regularisation(reg, n=self.A.shape[1], device=self.A.device)

    def set_wLasso_ISO(self, img_weights_filename, lambda_perc_iso):
        """Weighted lasso on ISO from a weight image (core.pyx:768-858)."""
        img = nifti.load(img_weights_filename).data.astype(np.float32)
        if img.ndim != 3:
            raise ValueError("Weights image must be 3D dataset")
        if tuple(img.shape) != tuple(self.get_config("dim")):
            raise ValueError("Dataset does not have the same geometry as the DWI signal")
        img[np.isnan(img)] = 1.0
        D = self.DICTIONARY
        w = img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"]].flatten().astype(np.float32)[D["ISO"]["vox"]]
        if np.any(w < 1):
            raise ValueError("All weights must be greater or equal to 1")
        if self.KERNELS["iso"].shape[0] > 1:
            w = np.tile(w, self.KERNELS["iso"].shape[0])
        self.set_regularisation(regularisers=(None, None, "lasso"), lambdas=(None, None, lambda_perc_iso),
                                is_nonnegative=(True, True, True), params=(None, None, {"coeff_weights": w}))

    # ---------------------------------------------------------------- fit --
    def fit(self, tol_fun=1e-3, tol_x=1e-6, max_iter=100, x0=None, confidence_map_filename=None,
            confidence_map_rescale=False, debias=False):
        """FISTA fit on the GPU (core.pyx:1353-1519), optional confidence map and debias."""
        for name in ("niiDWI_img", "DICTIONARY", "KERNELS", "THREADS", "A"):
            if getattr(self, name) is None:
                raise RuntimeError(f"{name} not set; run the previous steps of the flow first")
        if self.regularisation_params is None:
            self.set_regularisation()
        self.confidence_map_img = None
        self.set_config("confidence_map_filename", None)
        confidence_array = None
        if confidence_map_filename is not None:
            cm = nifti.load(pjoin(self.get_config("DATA_path"), confidence_map_filename)).data.astype(np.float32)
            if cm.ndim not in (3, 4):
                raise ValueError("Confidence map must be 3D or 4D dataset")
            if cm.ndim == 3:
                cm = np.repeat(cm[:, :, :, None], self.niiDWI_img.shape[3], axis=3)
            if cm.shape != self.niiDWI_img.shape:
                raise ValueError("Dataset does not have the same geometry as the DWI signal")
            self.set_config("confidence_map_filename", confidence_map_filename)
            D = self.DICTIONARY
            confidence_array = cm[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :].flatten().astype(np.float32)
            changed = False
            if np.isnan(confidence_array).any():               # (core.pyx:1433-1437)
                confidence_array[np.isnan(confidence_array)] = 0.0
                changed = True
            cmax, cmin = confidence_array.max(), confidence_array.min()
            if cmin == cmax:                                    # a constant map is not used
                confidence_array = None
                self.set_config("confidence_map_filename", None)
            else:
                if confidence_map_rescale:
                    confidence_array = (confidence_array - cmin) / (cmax - cmin)
                    changed = True
                if changed:
                    # the processed values replace the map at the mask voxels, so that
                    # save_results weights its adjusted errors with them (core.pyx:1454-1457)
                    cm[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"], :] = np.reshape(
                        confidence_array, (D["IC"]["nVOX"], -1)).astype(np.float32)
                self.confidence_map_img = cm
        if x0 is not None and x0.shape[0] != self.A.shape[1]:
            raise ValueError("x0 dimension does not match the number of columns of the dictionary")
        opt = self.CONFIG["optimization"] = {"tol_fun": tol_fun, "tol_x": tol_x, "max_iter": max_iter,
                                            "regularisation": self.regularisation_params}
        t = time.time()
        self.x, details = solvers.solve(self.get_y(), self.A, self.A.T, tol_fun=tol_fun, tol_x=tol_x,
                                        max_iter=max_iter, verbose=self.verbose, x0=x0,
                                        regularisation=self.regularisation_params,
                                        confidence_array=confidence_array)
        rp = self.regularisation_params
        if (rp["regIC"] is not None or rp["regEC"] is not None or rp["regISO"] is not None) and debias:
            # debias (core.pyx:1476-1508): refit on the support, without penalties
            offset = self.DICTIONARY["IC"]["nSTR"] * self.KERNELS["wmr"].shape[0]
            mask = np.ones(offset, dtype=np.uint32)
            mask[self.x[:offset] < 1e-15] = 0
            self.A.set_eval(mask)
            self.set_regularisation()
            x_debias = self.x.copy()
            x_debias[:offset] *= mask
            x_debias[offset:] = 0
            y_mask = np.asarray(self.A.dot(x_debias))
            y_mask[y_mask < 0] = 0
            y_mask[y_mask > 0] = 1
            self.debias_mask = y_mask
            self.x, details = solvers.solve(self.get_y(), self.A, self.A.T, tol_fun=tol_fun, tol_x=tol_x,
                                            max_iter=max_iter, verbose=self.verbose, x0=x0,
                                            regularisation=self.regularisation_params,
                                            confidence_array=confidence_array)
        opt["fit_details"] = details
        opt["fit_time"] = time.time() - t
        if self.regularisation_params["regIC"] is None and self.x_nnls is None:
            self.x_nnls, _, _ = self.get_coeffs(get_normalized=False)

    def get_coeffs(self, get_normalized=True):
        """(xic, xec, xiso) of the original problem (core.pyx:1523-1560)."""
        if self.x is None:
            raise RuntimeError('Model not fitted to the data; call "fit()" first')
        D, K = self.DICTIONARY, self.KERNELS
        nSTR, nE, nV = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
        if get_normalized and self.get_config("doNormalizeKernels"):
            norm1 = np.repeat(K["wmr_norm"], nSTR)
            norm2 = np.repeat(K["wmh_norm"], nE)
            norm3 = np.repeat(K["iso_norm"], nV)
            norm_fib = np.kron(np.ones(K["wmr"].shape[0]), D["TRK"]["norm"])
            x = self.x / np.hstack((norm1 * norm_fib, norm2, norm3))
        else:
            x = self.x
        off1 = nSTR * K["wmr"].shape[0]
        off2 = off1 + nE * K["wmh"].shape[0]
        kept = np.tile(D["TRK"]["kept"], K["wmr"].shape[0])
        xic = np.zeros(kept.size)
        xic[kept == 1] = x[:off1]
        return xic, x[off1:off2], x[off2:]

    # ------------------------------------------------------------ results --
    def save_results(self, path_suffix=None, coeffs_format="%.5e", stat_coeffs="sum", save_est_dwi=False):
        """Fit errors, compartment maps, streamline_weights.txt and results.pickle
        (core.pyx:1563-1837)."""
        if self.x is None:
            raise RuntimeError('Model not fitted to the data; call "fit()" first')
        D, K = self.DICTIONARY, self.KERNELS
        name = "Results_" + (self.model.id if self.model is not None else K.get("model", "model"))
        if path_suffix:
            self.set_config("path_suffix", path_suffix)
            name += path_suffix
        out = pjoin(self.get_config("TRACKING_path"), name)
        os.makedirs(out, exist_ok=True)
        for f in glob.glob(pjoin(out, "*")):
            os.remove(f)
        self.set_config("RESULTS_path", out)
        nSTR, nE = D["IC"]["nSTR"], D["EC"]["n"]
        norm_fib = np.ones(nSTR)
        if self.get_config("doNormalizeKernels"):
            norm1 = np.repeat(K["wmr_norm"], nSTR)
            norm2 = np.repeat(K["wmh_norm"], nE)
            norm3 = np.repeat(K["iso_norm"], D["ISO"]["n"])
            norm_fib = np.kron(np.ones(K["wmr"].shape[0]), D["TRK"]["norm"])
            x = self.x / np.hstack((norm1 * norm_fib, norm2, norm3))
        else:
            x = self.x
        dim = self.get_config("dim")
        affine = self.niiDWI.affine if self.niiDWI is not None else np.eye(4)
        pix = self.get_config("pixdim") or (1.0, 1.0, 1.0)
        ix, iy, iz = D["MASK_ix"], D["MASK_iy"], D["MASK_iz"]
        nVOX = D["IC"]["nVOX"]
        y_est_all = np.asarray(self.A.dot(self.x))
        stats = {}

        def rmse_nrmse(ym, ye):
            r = np.sqrt(np.mean((ym - ye) ** 2, axis=1))
            den = np.sum(ym ** 2, axis=1)
            idx = den < 1e-12
            den[idx] = 1
            n = np.sqrt(np.sum((ym - ye) ** 2, axis=1) / den)
            n[idx] = 0
            return r, n

        if self.debias_mask is not None:
            # after a debiased fit the errors compare A x with the masked signal
            # the fit saw (get_y); the statistics cover the rows in the mask only
            # (core.pyx:1624-1658)
            y_fit = self.get_y()
            nS = self.niiDWI_img.shape[3]
            nV = int(np.sum(self.debias_mask) / nS)
            if nV == 0:
                raise ValueError("the debias mask holds no voxel: the regularised fit left every IC "
                                 "coefficient at zero")
            ind = np.nonzero(self.debias_mask > 0)[0]
            r, n = rmse_nrmse(np.reshape(y_fit[ind], (nV, -1)), np.reshape(y_est_all[ind], (nV, -1)))
            stats["RMSE"] = (float(r.mean()), float(r.std()))
            stats["NRMSE"] = (float(n.mean()), float(n.std()))
            y_mea = np.reshape(y_fit, (nVOX, -1))
            y_est = np.reshape(y_est_all, (nVOX, -1)).astype(np.float32)
            r, n = rmse_nrmse(y_mea, y_est)
        else:
            y_mea = np.reshape(self.niiDWI_img[ix, iy, iz, :].flatten().astype(np.float32), (nVOX, -1))
            y_est = np.reshape(y_est_all, (nVOX, -1)).astype(np.float32)
            r, n = rmse_nrmse(y_mea, y_est)
            stats["RMSE"] = (float(r.mean()), float(r.std()))
            stats["NRMSE"] = (float(n.mean()), float(n.std()))
        self._save_map(r, pjoin(out, "fit_RMSE.nii.gz"), dim, affine, pix)
        self._save_map(n, pjoin(out, "fit_NRMSE.nii.gz"), dim, affine, pix)
        if self.confidence_map_img is not None:
            ca = np.reshape(self.confidence_map_img[ix, iy, iz, :].flatten().astype(np.float32), (nVOX, -1))
            den = np.sum(ca, axis=1)
            idx = den < 1e-12
            den[idx] = 1
            tmp = np.sqrt(np.sum(ca * (y_mea - y_est) ** 2, axis=1) / den)
            tmp[idx] = 0
            self._save_map(tmp, pjoin(out, "fit_RMSE_adjusted.nii.gz"), dim, affine, pix)
            den = np.sum(ca * y_mea ** 2, axis=1)
            idx = den < 1e-12
            den[idx] = 1
            tmp = np.sqrt(np.sum(ca * (y_mea - y_est) ** 2, axis=1) / den)
            tmp[idx] = 0
            self._save_map(tmp, pjoin(out, "fit_NRMSE_adjusted.nii.gz"), dim, affine, pix)
        # voxelwise compartment contributions (core.pyx:1709-1756)
        maps = {}
        off = nSTR * K["wmr"].shape[0]
        if len(K["wmr"]) > 0:
            t = (x[:off].reshape((-1, nSTR)) * norm_fib.reshape((-1, nSTR))).sum(axis=0)
            maps["IC"] = np.bincount(D["IC"]["vox"], minlength=nVOX,
                                     weights=t[D["IC"]["str"]] * D["IC"]["len"]).astype(np.float32)
        else:
            maps["IC"] = np.zeros(nVOX, np.float32)
        if len(K["wmh"]) > 0:
            t = x[off:off + nE * len(K["wmh"])].reshape((-1, nE)).sum(axis=0)
            maps["EC"] = np.bincount(D["EC"]["vox"], weights=t, minlength=nVOX).astype(np.float32)
        else:
            maps["EC"] = np.zeros(nVOX, np.float32)
        if len(K["iso"]) > 0:
            # the reference's map takes the block of the first ISO response
            # function only (offset_iso = offset + ISO['n'], core.pyx:1737-1740),
            # also with several ISO RFs; kept, so the maps are the reference's
            o2 = off + nE * len(K["wmh"])
            t = x[o2:o2 + D["ISO"]["n"]].reshape((-1, D["ISO"]["n"])).sum(axis=0)
            maps["ISO"] = np.bincount(D["ISO"]["vox"], weights=t, minlength=nVOX).astype(np.float32)
        else:
            maps["ISO"] = np.zeros(nVOX, np.float32)
        tot = maps["IC"] + maps["EC"] + maps["ISO"] + 1e-16
        for c in ("IC", "EC", "ISO"):
            v = maps[c] / tot if self.get_config("doNormalizeMaps") else maps[c]
            self._save_map(v, pjoin(out, f"compartment_{c}.nii.gz"), dim, affine, pix)
        # streamline weights
        xic, _, _ = self.get_coeffs()
        self.set_config("stat_coeffs", stat_coeffs)
        if stat_coeffs != "all" and xic.size > 0:
            xic = np.reshape(xic, (-1, D["TRK"]["kept"].size))
            ops = {"sum": np.sum, "mean": np.mean, "median": np.median, "min": np.min, "max": np.max}
            if stat_coeffs not in ops:
                raise ValueError("Stat not allowed. Possible values: sum, mean, median, min, max, all")
            xic = ops[stat_coeffs](xic, axis=0)
        info = self.dictionary_info
        if info.get("blur_gauss_extent", 0) > 0 or info.get("blur_core_extent", 0) > 0:
            xic[D["TRK"]["kept"] == 1] *= D["TRK"]["lenTot"] / D["TRK"]["len"]
        np.savetxt(pjoin(out, "streamline_weights.txt"), xic, fmt=coeffs_format)
        opt = self.CONFIG.get("optimization", {})
        saved = {k: v for k, v in opt.get("regularisation", {}).items() if k not in ("omega", "prox", "_device")} \
            if isinstance(opt.get("regularisation"), dict) else None
        cfg = dict(self.CONFIG)
        cfg["optimization"] = dict(opt, regularisation=saved)
        with open(pjoin(out, "results.pickle"), "wb+") as fid:
            pickle.dump([cfg, self.x, x], fid, protocol=2)
        if save_est_dwi:
            img = np.zeros(tuple(dim) + (y_est.shape[1],), dtype=np.float32)
            img[ix, iy, iz, :] = y_est
            nifti.save(nifti.Nifti1Image(img, affine, pix), pjoin(out, "fit_signal_estimated.nii.gz"))
        self.temp_data["fit_stats"] = stats
        return stats

    def _save_map(self, values, path, dim, affine, pix):
        """Scatter per-voxel values into the volume and write a NIfTI map."""
        D = self.DICTIONARY
        img = np.zeros(dim, dtype=np.float32)
        img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"]] = values
        nifti.save(nifti.Nifti1Image(img, affine, pix, descrip=f"Created with {VERSION}"), path)
