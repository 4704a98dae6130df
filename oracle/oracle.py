"""ORACLE — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or, for the
baseline, as the thing timed on host cores).  The product path never imports it.

CPU restatement of the reference algorithm for the operator + FISTA hot path:

* ``set_threads``           <- Evaluation.set_threads        core.pyx:586-724
* ``assemble_A``            <- closed-form entries of A as computed by
                               COMMIT_A__block / COMMIT_At__block
                               setup_operator.py:50-218 / 370-563
* ``OracleOperator``        <- LinearOperator                 operator.pyx:10-184,
                               calling the C restatement ``operator_oracle.c``
* proximals                 <- proximals.pyx:13-186
* ``init_regularisation``   <- solvers.py:16-206
* ``fista`` / ``solve``     <- solvers.py:218-391
* ``lut_grad``              <- derivative of the entries above with respect to
                               the LUT values (no reference counterpart; the
                               terms are those of setup_operator.py:62-218)
* ``normalize_kernels``     <- Evaluation.load_kernels        core.pyx:412-431
* ``sort_and_normalize``    <- Evaluation.load_dictionary     core.pyx:509-525, 571-580
* ``lambda_max_lasso`` / ``lambda_max_group``  <- set_regularisation  core.pyx:961-976
* ``group_weights``         <- set_regularisation             core.pyx:1123-1192
* ``debias_eval`` / ``debias_rows``  <- fit (debias)          core.pyx:1476-1501
* ``confidence``            <- fit (confidence map)           core.pyx:1433-1457
* ``fit_errors`` / ``compartment_maps``  <- save_results      core.pyx:1614-1756
* ``iso_weights``           <- set_wLasso_ISO                 core.pyx:828-843

Pinning (tests/test_oracle.py): the C operator against the dense brute-force
assembly, adjointness, the single-segment closed form and the all-ones LUT
incidence case; the proximals against brute-force minimisation of
0.5||u-x||^2 + lam*Omega(u); FISTA against scipy.optimize.nnls on small dense
problems, KKT conditions, and the orthogonal-A closed form max(A^T y, 0);
the FISTA trajectory against hand-derived iterates, backtracking counts and
stopping cases (tests/test_oracle_fista.py); the Evaluation pieces against
hand-computed examples (tests/test_oracle_core.py).  No reference golden
vectors exist offline (their inputs are network-only).
"""
import ctypes
import os
from math import sqrt

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
eps = np.finfo(float).eps


# ----------------------------------------------------------------- library --
def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle`)")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    U = ctypes.c_uint32
    lib.oracle_A.argtypes = [P, P, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, U, U, U, U]
    lib.oracle_At.argtypes = [P, P, I, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, U, U, U, U]
    lib.oracle_A_nolut.argtypes = [P, P, I, P, P, P, P, P, P, P, U, U]
    lib.oracle_At_nolut.argtypes = [P, P, I, I, P, P, P, P, P, P, P, U, U]
    for f in (lib.oracle_A, lib.oracle_At, lib.oracle_A_nolut, lib.oracle_At_nolut):
        f.restype = ctypes.c_int
    return lib


_LIBS = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _LIBS:
        _LIBS[name] = _load(name)
    return _LIBS[name]


def _p(a):
    if a is None or a.size == 0:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------- set_threads --
def set_threads(DICTIONARY, n):
    """Host work partition, restated from core.pyx:586-724.

    Returns the THREADS dict: 'IC','EC','ISO' (u32[n+1] voxel-aligned ranges for
    A x), 'ICt' (u8 owner per segment for A^T y), 'ECt','ISOt' (u32[n+1]).
    Raises ValueError where the reference logs an error.
    """
    if n < 1 or n > 255:
        raise ValueError("Number of threads must be between 1 and 255")
    IC, EC, ISO = DICTIONARY["IC"], DICTIONARY["EC"], DICTIONARY["ISO"]
    T = {"n": n}
    if IC["n"] > 0:
        t_ic = np.zeros(n + 1, dtype=np.uint32)
        if n > 1:
            N = np.floor(IC["n"] / n)
            t, tot = 1, 0
            for c in np.bincount(IC["vox"]):
                tot += int(c)
                if tot >= N and t <= n:
                    t_ic[t] = t_ic[t - 1] + tot
                    t += 1
                    tot = 0
        t_ic[n] = IC["n"]
        if np.count_nonzero(np.diff(t_ic.astype(np.int32)) <= 0):
            raise ValueError("Too many threads for the IC compartments to evaluate")
        T["IC"] = t_ic
    else:
        T["IC"] = None
    if EC["n"] > 0:
        t_ec = np.zeros(n + 1, dtype=np.uint32)
        for i in range(n):
            t_ec[i] = np.searchsorted(EC["vox"], IC["vox"][T["IC"][i]])
        t_ec[n] = EC["n"]
        if np.count_nonzero(np.diff(t_ec.astype(np.int32)) <= 0):
            raise ValueError("Too many threads for the EC compartments to evaluate")
        T["EC"] = t_ec
    else:
        T["EC"] = None
    if ISO["n"] > 0:
        t_iso = np.zeros(n + 1, dtype=np.uint32)
        for i in range(n):
            t_iso[i] = np.searchsorted(ISO["vox"], IC["vox"][T["IC"][i]])
        t_iso[n] = ISO["n"]
        T["ISO"] = t_iso
    else:
        T["ISO"] = None
    if IC["n"] > 0:
        ict = np.full(IC["n"], n - 1, dtype=np.uint8)
        if n > 1:
            idx = np.argsort(IC["str"], kind="mergesort")
            C = np.bincount(IC["str"])
            t = tot = i1 = i2 = 0
            N = np.floor(IC["n"] / n)
            for c in C:
                c = int(c)
                i2 += c
                tot += c
                if tot >= N:
                    ict[i1:i2] = t
                    t += 1
                    if t == n - 1:
                        break
                    i1 = i2
                    tot = c
            ict[idx] = ict.copy()
        T["ICt"] = ict
    else:
        T["ICt"] = None
    for key, comp in (("ECt", EC), ("ISOt", ISO)):
        if comp["n"] > 0:
            tt = np.zeros(n + 1, dtype=np.uint32)
            N = np.floor(comp["n"] / n)
            for i in range(1, n):
                tt[i] = tt[i - 1] + N
            tt[n] = comp["n"]
            if np.count_nonzero(np.diff(tt.astype(np.int32)) <= 0):
                raise ValueError(f"Too many threads for the {key[:-1]} compartments to evaluate")
            T[key] = tt
        else:
            T[key] = None
    return T


# --------------------------------------------------------- dense assembly --
def assemble_A(DICTIONARY, KERNELS, nolut=False):
    """FP64 scipy.sparse CSR matrix of the operator, from the closed-form entries

        A[v*nS+s, i*nF+f]           += l * eval[i*nF+f] * LUT_IC[i,o,s]   (each IC segment)
        A[v*nS+s, nIC*nF + i*nE+e]  += LUT_EC[i,o_e,s]                    (each EC peak e)
        A[v*nS+s, nIC*nF+nEC*nE + i*nI+k] += LUT_ISO[i,s]                 (each ISO voxel k)

    (setup_operator.py:62-103, 111-165, 167-218; LUT-free: nS=1, LUT=1, no EC,
    setup_operator.py:719-752).
    """
    IC, EC, ISO = DICTIONARY["IC"], DICTIONARY["EC"], DICTIONARY["ISO"]
    wmr, wmh, iso = KERNELS["wmr"], KERNELS["wmh"], KERNELS["iso"]
    nIC, nEC, nISO = wmr.shape[0], wmh.shape[0], iso.shape[0]
    if wmr.size > 0:
        nS = wmr.shape[2]
    elif wmh.size > 0:
        nS = wmh.shape[2]
    else:
        nS = wmr.shape[1]
    if nolut:
        nS, nEC = 1, 0
        nIC = 1
    nF, nE, nI, nV = IC["nSTR"], EC["n"], ISO["n"], IC["nVOX"]
    rows, cols, vals = [], [], []
    s_idx = np.arange(nS)
    ev = IC.get("eval")
    if ev is None:
        ev = np.ones(nF * nIC, dtype=np.uint32)
    for k in range(IC["n"]):
        v, f, o, l = int(IC["vox"][k]), int(IC["str"][k]), int(IC["dir"][k]), float(IC["len"][k])
        for i in range(nIC):
            col = i * nF + f
            lut = np.ones(1) if nolut else wmr[i, o, :].astype(np.float64)
            rows.append(v * nS + s_idx)
            cols.append(np.full(nS, col))
            vals.append(l * float(ev[col]) * lut)
    for e in range(nE if nEC > 0 else 0):
        v, o = int(EC["vox"][e]), int(EC["dir"][e])
        for i in range(nEC):
            rows.append(v * nS + s_idx)
            cols.append(np.full(nS, nIC * nF + i * nE + e))
            vals.append(wmh[i, o, :].astype(np.float64))
    for k in range(nI if nISO > 0 else 0):
        v = int(ISO["vox"][k])
        for i in range(nISO):
            lut = np.ones(1) if nolut else iso[i, :].astype(np.float64)
            rows.append(v * nS + s_idx)
            cols.append(np.full(nS, nIC * nF + nEC * nE + i * nI + k))
            vals.append(lut)
    nrows = nV * nS
    ncols = nF * nIC + nE * nEC + nI * nISO
    if not rows:
        return sp.csr_matrix((nrows, ncols))
    return sp.csr_matrix(
        (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nrows, ncols)
    )


def lut_grad(DICTIONARY, KERNELS, x, gy):
    """Gradient of L(A x) with respect to the LUT values, given gy = dL/dy:

        gwmr[i,o,s] = sum_{IC segments (f,v,o,l)} l * eval[i*nF+f] * x[i*nF+f] * gy[v*nS+s]
        gwmh[i,o,s] = sum_{EC peaks e in voxel v with orientation o} x[nIC*nF + i*nE + e] * gy[v*nS+s]
        giso[i,s]   = sum_{ISO voxels k in voxel v} x[nIC*nF + nEC*nE + i*nI + k] * gy[v*nS+s]

    i.e. the coefficient of every LUT entry in the closed form of assemble_A.
    Pinned in tests/test_oracle.py by central finite differences of
    0.5 ||A(LUT) x - y||^2 on the dense assembly."""
    IC, EC, ISO = DICTIONARY["IC"], DICTIONARY["EC"], DICTIONARY["ISO"]
    wmr, wmh, iso = KERNELS["wmr"], KERNELS["wmh"], KERNELS["iso"]
    nIC, nEC, nISO = wmr.shape[0], wmh.shape[0], iso.shape[0]
    nS = wmr.shape[2] if wmr.size else (wmh.shape[2] if wmh.size else iso.shape[1])
    ndirs = wmr.shape[1] if wmr.ndim == 3 else wmh.shape[1]
    nF, nE, nI = IC["nSTR"], EC["n"], ISO["n"]
    x = np.asarray(x, dtype=np.float64)
    G = np.asarray(gy, dtype=np.float64).reshape(-1, nS)
    ev = IC.get("eval")
    if ev is None:
        ev = np.ones(nF * nIC)
    gwmr = np.zeros((nIC, ndirs, nS))
    gwmh = np.zeros((nEC, ndirs, nS))
    giso = np.zeros((nISO, nS))
    v = np.asarray(IC["vox"], dtype=np.int64)
    f = np.asarray(IC["str"], dtype=np.int64)
    o = np.asarray(IC["dir"], dtype=np.int64)
    l = np.asarray(IC["len"], dtype=np.float64)
    for i in range(nIC):
        c = l * x[i * nF + f] * np.asarray(ev, dtype=np.float64)[i * nF + f]
        np.add.at(gwmr[i], o, c[:, None] * G[v])
    ve, oe = np.asarray(EC["vox"], dtype=np.int64), np.asarray(EC["dir"], dtype=np.int64)
    for i in range(nEC):
        c = x[nIC * nF + i * nE + np.arange(nE)]
        np.add.at(gwmh[i], oe, c[:, None] * G[ve])
    vi = np.asarray(ISO["vox"], dtype=np.int64)
    for i in range(nISO):
        c = x[nIC * nF + nEC * nE + i * nI + np.arange(nI)]
        giso[i] = (c[:, None] * G[vi]).sum(axis=0)
    return gwmr, gwmh, giso


# ------------------------------------------------------- C-backed operator --
class OracleOperator:
    """LinearOperator restated (operator.pyx:40-184) on top of operator_oracle.c."""

    def __init__(self, DICTIONARY, KERNELS, THREADS, nolut=False, fast=False):
        self.D, self.K, self.THREADS = DICTIONARY, KERNELS, THREADS
        self.nolut = nolut
        self.fast = fast
        self.adjoint = 0
        IC, EC, ISO = DICTIONARY["IC"], DICTIONARY["EC"], DICTIONARY["ISO"]
        self.nIC, self.nEC, self.nISO = (KERNELS[k].shape[0] for k in ("wmr", "wmh", "iso"))
        if KERNELS["wmr"].size > 0:
            self.nS = KERNELS["wmr"].shape[2]
        elif KERNELS["wmh"].size > 0:
            self.nS = KERNELS["wmh"].shape[2]
        else:
            self.nS = KERNELS["wmr"].shape[1]
        self.ndirs = KERNELS["wmr"].shape[1]
        self.nROWS = IC["nVOX"] * self.nS
        self.nCOLS = IC["nSTR"] * self.nIC + EC["n"] * self.nEC + ISO["n"] * self.nISO

    @property
    def T(self):
        C = OracleOperator.__new__(OracleOperator)
        C.__dict__.update(self.__dict__)
        C.adjoint = 1 - self.adjoint
        return C

    @property
    def shape(self):
        return (self.nROWS, self.nCOLS) if not self.adjoint else (self.nCOLS, self.nROWS)

    def dot(self, v_in):
        v_in = np.ascontiguousarray(v_in, dtype=np.float64)
        if v_in.size != self.shape[1]:
            raise RuntimeError("A.dot(): dimensions do not match")
        out = np.zeros(self.shape[0], dtype=np.float64)
        D, K, T = self.D, self.K, self.THREADS
        IC, EC, ISO = D["IC"], D["EC"], D["ISO"]
        L = lib(self.fast)
        ev = IC["eval"]
        wmr = np.ascontiguousarray(K["wmr"], dtype=np.float32)
        wmh = np.ascontiguousarray(K["wmh"], dtype=np.float32)
        iso = np.ascontiguousarray(K["iso"], dtype=np.float32)
        if not self.adjoint:
            if self.nolut:
                rc = L.oracle_A_nolut(_p(v_in), _p(out), IC["nSTR"], _p(IC["str"]), _p(ev), _p(IC["vox"]),
                                      _p(IC["len"]), _p(ISO["vox"]), _p(T["IC"]), _p(T["ISO"]),
                                      self.nISO, T["n"])
            else:
                rc = L.oracle_A(_p(v_in), _p(out), self.nS, self.ndirs, IC["nSTR"], EC["n"], ISO["n"],
                                _p(IC["str"]), _p(ev), _p(IC["vox"]), _p(IC["dir"]), _p(IC["len"]),
                                _p(EC["vox"]), _p(EC["dir"]), _p(ISO["vox"]), _p(wmr), _p(wmh), _p(iso),
                                _p(T["IC"]), _p(T["EC"]), _p(T["ISO"]),
                                self.nIC, self.nEC, self.nISO, T["n"])
        else:
            if self.nolut:
                rc = L.oracle_At_nolut(_p(v_in), _p(out), IC["nSTR"], IC["n"], _p(IC["str"]), _p(ev),
                                       _p(IC["vox"]), _p(IC["len"]), _p(ISO["vox"]), _p(T["ICt"]),
                                       _p(T["ISOt"]), self.nISO, T["n"])
            else:
                rc = L.oracle_At(_p(v_in), _p(out), self.nS, self.ndirs, IC["nSTR"], IC["n"], EC["n"], ISO["n"],
                                 _p(IC["str"]), _p(ev), _p(IC["vox"]), _p(IC["dir"]), _p(IC["len"]),
                                 _p(EC["vox"]), _p(EC["dir"]), _p(ISO["vox"]), _p(wmr), _p(wmh), _p(iso),
                                 _p(T["ICt"]), _p(T["ECt"]), _p(T["ISOt"]),
                                 self.nIC, self.nEC, self.nISO, T["n"])
        if rc != 0:
            raise RuntimeError(f"oracle operator failed with code {rc}")
        return out


# ---------------------------------------------------------- preprocessing --
def normalize_kernels(KERNELS):
    """doNormalizeKernels branch of load_kernels (core.pyx:412-431): divide each
    RF by the norm of its orientation-0 row (ISO: by its own norm)."""
    K = dict(KERNELS)
    for key, nkey in (("wmr", "wmr_norm"), ("wmh", "wmh_norm")):
        a = np.ascontiguousarray(K[key], dtype=np.float32).copy()
        norms = np.zeros(a.shape[0])
        for i in range(a.shape[0]):
            norms[i] = np.linalg.norm(a[i, 0, :])
            for j in range(a.shape[1]):
                a[i, j, :] /= norms[i]
        K[key], K[nkey] = a, norms
    a = np.ascontiguousarray(K["iso"], dtype=np.float32).copy()
    norms = np.zeros(a.shape[0])
    for i in range(a.shape[0]):
        norms[i] = np.linalg.norm(a[i, :])
        a[i, :] /= norms[i]
    K["iso"], K["iso_norm"] = a, norms
    return K


def sort_and_normalize(raw, mask, normalize=True):
    """The hot-path part of load_dictionary (core.pyx:505-580) on in-memory
    arrays: stable mergesort by linear voxel id, segment length divided by the
    fibre's TRK_norm, then linear ids remapped to compact mask indices taken
    in Fortran order."""
    D = {"IC": {}, "EC": {}, "ISO": {}, "TRK": {}}
    D["TRK"]["norm"] = raw["TRK_norm"].astype(np.float32)
    D["TRK"]["kept"] = raw.get("TRK_kept", np.ones(raw["TRK_norm"].size, np.uint8))
    idx = np.argsort(raw["IC_v"], kind="mergesort")
    D["IC"]["vox"] = raw["IC_v"][idx]
    D["IC"]["dir"] = raw["IC_o"][idx]
    D["IC"]["str"] = raw["IC_f"][idx]
    D["IC"]["len"] = raw["IC_len"][idx].astype(np.float32).copy()
    D["IC"]["n"] = D["IC"]["vox"].size
    D["IC"]["nSTR"] = D["TRK"]["norm"].size
    D["IC"]["nVOX"] = int(mask.sum())
    if normalize:
        sl, tl, f = D["IC"]["len"], D["TRK"]["norm"], D["IC"]["str"]
        for s in range(sl.size):
            sl[s] /= tl[f[s]]
    idx = np.argsort(raw["EC_v"], kind="mergesort")
    D["EC"]["vox"], D["EC"]["dir"] = raw["EC_v"][idx], raw["EC_o"][idx]
    D["EC"]["n"] = D["EC"]["vox"].size
    idx = np.argsort(raw["ISO_v"], kind="mergesort")
    D["ISO"]["vox"] = raw["ISO_v"][idx]
    D["ISO"]["n"] = D["ISO"]["vox"].size
    lin = mask.ravel(order="F").nonzero()[0]
    lut = np.zeros(mask.size, dtype=np.uint32)
    for i in range(lin.size):
        lut[lin[i]] = i
    for c in ("IC", "EC", "ISO"):
        D[c]["vox"] = lut[D[c]["vox"]]
    D["MASK"] = mask
    return D


# -------------------------------------------------------------- proximals --
def non_negativity(x, start, size):                                   # proximals.pyx:13-22
    for i in range(start, start + size):
        if x[i] <= 0.0:
            x[i] = 0.0
    return x


def soft_thresholding(x, lam, start, size):                           # proximals.pyx:25-39
    for i in range(start, start + size):
        if x[i] > lam:
            x[i] = x[i] - lam
        elif x[i] < -lam:
            x[i] = x[i] + lam
        else:
            x[i] = 0.0
    return x


def w_soft_thresholding(x, w, lam, start, size):                      # proximals.pyx:42-56
    for i in range(start, start + size):
        if x[i] > w[i] * lam:
            x[i] = x[i] - w[i] * lam
        elif x[i] < -w[i] * lam:
            x[i] = x[i] + w[i] * lam
        else:
            x[i] = 0.0
    return x


def omega_group_lasso(x, group_idx, group_size, group_weight, lam):   # proximals.pyx:80-99
    om = 0.0
    if lam != 0:
        j = 0
        for k in range(group_size.size):
            N = int(group_size[k])
            g = 0.0
            for i in range(j, j + N):
                g += x[group_idx[i]] * x[group_idx[i]]
            om += group_weight[k] * sqrt(g)
            j += N
    return lam * om


def prox_group_lasso(x, group_idx, group_size, group_weight, lam):    # proximals.pyx:103-134
    if lam != 0:
        j = 0
        for k in range(group_size.size):
            N = int(group_size[k])
            g = 0.0
            for i in range(j, j + N):
                g += x[group_idx[i]] * x[group_idx[i]]
            g = sqrt(g)
            wl = group_weight[k] * lam
            if g <= wl:
                for i in range(j, j + N):
                    x[group_idx[i]] = 0.0
            else:
                wl = (g - wl) / g
                for i in range(j, j + N):
                    x[group_idx[i]] *= wl
            j += N
    return x


def omega_sparse_group_lasso(x, group_idx, group_size, group_weight, lam1, lam2):   # proximals.pyx:138-159
    o1 = o2 = 0.0
    if lam1 != 0 or lam2 != 0:
        o1 = float(np.sum(np.abs(x)))
        o2 = omega_group_lasso(x, group_idx, group_size, group_weight, 1.0)
    return lam1 * o1 + lam2 * o2


def omega_w_sparse_group_lasso(x, w, group_idx, group_size, group_weight, lam1, lam2):  # proximals.pyx:162-186
    o1 = o2 = 0.0
    if lam1 != 0 or lam2 != 0:
        o1 = float(np.sum(np.abs(w * x)))
        o2 = omega_group_lasso(x, group_idx, group_size, group_weight, 1.0)
    return lam1 * o1 + lam2 * o2


def _flatten_groups(groups):
    size = np.array([g.size for g in groups], dtype=np.int32)
    idx = np.concatenate([np.asarray(g, dtype=np.int32) for g in groups]) if len(groups) else np.zeros(0, np.int32)
    return idx, size


def init_regularisation(rp):
    """Restates solvers.init_regularisation (solvers.py:16-206), including its
    slicing quirks (the weighted-lasso omegas slice [start:size])."""
    sIC, zIC = rp.get("startIC"), rp.get("sizeIC")
    sEC, zEC = rp.get("startEC"), rp.get("sizeEC")
    sISO, zISO = rp.get("startISO"), rp.get("sizeISO")
    w_all = np.ones(zIC + zEC + zISO, dtype=np.float64)
    dIC = rp.get("dictIC_params")
    dISO = rp.get("dictISO_params")
    if dIC is not None and "coeff_weights_kept" in dIC:
        w_all[sIC:sIC + zIC] = dIC["coeff_weights_kept"]
    if dISO is not None and "coeff_weights" in dISO:
        w_all[sISO:sISO + zISO] = dISO["coeff_weights"]

    reg = rp["regIC"]
    if reg is None:
        omegaIC = lambda x: 0.0
        proxIC = (lambda x, _: non_negativity(x, sIC, zIC)) if rp.get("nnIC") else (lambda x, _: x)
    elif reg == "lasso":
        lam = rp["lambdaIC"]
        if dIC is not None and "coeff_weights_kept" in dIC:
            omegaIC = lambda x: lam * np.linalg.norm(w_all[sIC:zIC] * x[sIC:zIC], 1)
            baseIC = lambda x, sc: w_soft_thresholding(x, w_all, sc * lam, sIC, zIC)
        else:
            omegaIC = lambda x: lam * np.linalg.norm(x[sIC:zIC], 1)
            baseIC = lambda x, sc: soft_thresholding(x, sc * lam, sIC, zIC)
        proxIC = (lambda x, sc: non_negativity(baseIC(x, sc), sIC, zIC)) if rp.get("nnIC") else baseIC
    elif reg in ("group_lasso", "sparse_group_lasso"):
        gidx, gsz = _flatten_groups(dIC["group_idx_kept"])
        gw = dIC["group_weights"]
        if reg == "group_lasso":
            lam_g = rp["lambdaIC"]
            omegaIC = lambda x: omega_group_lasso(x, gidx, gsz, gw, lam_g)
            baseIC = lambda x, sc: prox_group_lasso(x, gidx, gsz, gw, sc * lam_g)
        else:
            lam1, lam_g = rp["lambdaIC"][0], rp["lambdaIC"][1]
            if "coeff_weights_kept" in dIC:
                omegaIC = lambda x: omega_w_sparse_group_lasso(x, w_all, gidx, gsz, gw, lam1, lam_g)
                baseIC = lambda x, sc: prox_group_lasso(w_soft_thresholding(x, w_all, sc * lam1, sIC, zIC),
                                                      gidx, gsz, gw, sc * lam_g)
            else:
                omegaIC = lambda x: omega_sparse_group_lasso(x, gidx, gsz, gw, lam1, lam_g)
                baseIC = lambda x, sc: prox_group_lasso(soft_thresholding(x, sc * lam1, sIC, zIC),
                                                      gidx, gsz, gw, sc * lam_g)
        proxIC = (lambda x, sc: non_negativity(baseIC(x, sc), sIC, zIC)) if rp.get("nnIC") else baseIC
    else:
        raise ValueError("Regularisation not in the list")

    reg = rp["regEC"]
    if reg is None:
        omegaEC = lambda x: 0.0
        proxEC = (lambda x, _: non_negativity(x, sEC, zEC)) if rp.get("nnEC") else (lambda x, _: x)
    elif reg == "lasso":
        lamE = rp.get("lambdaEC")
        omegaEC = lambda x: lamE * np.linalg.norm(x[sEC:sEC + zEC], 1)
        baseEC = lambda x, sc: soft_thresholding(x, sc * lamE, sEC, zEC)
        proxEC = (lambda x, sc: non_negativity(baseEC(x, sc), sEC, zEC)) if rp.get("nnEC") else baseEC
    else:
        raise ValueError("Regularisation not in the list")

    reg = rp["regISO"]
    if reg is None:
        omegaISO = lambda x: 0.0
        proxISO = (lambda x, _: non_negativity(x, sISO, zISO)) if rp.get("nnISO") else (lambda x, _: x)
    elif reg == "lasso":
        lamI = rp.get("lambdaISO")
        if dISO is not None and "coeff_weights" in dISO:
            omegaISO = lambda x: lamI * np.linalg.norm(w_all[sISO:zISO] * x[sISO:zISO], 1)
            baseISO = lambda x, sc: w_soft_thresholding(x, w_all, sc * lamI, sISO, zISO)
        else:
            omegaISO = lambda x: lamI * np.linalg.norm(x[sISO:sISO + zISO], 1)
            baseISO = lambda x, sc: soft_thresholding(x, sc * lamI, sISO, zISO)
        proxISO = (lambda x, sc: non_negativity(baseISO(x, sc), sISO, zISO)) if rp.get("nnISO") else baseISO
    else:
        raise ValueError("Regularisation not in the list")

    rp = dict(rp)
    rp["omega"] = lambda x: omegaIC(x) + omegaEC(x) + omegaISO(x)
    rp["prox"] = lambda x, sc: proxIC(proxEC(proxISO(x, sc), sc), sc)
    return rp


# ------------------------------------------------------------------ FISTA --
def solve(y, A, At, tol_fun=1e-4, tol_x=1e-6, max_iter=1000, x0=None, regularisation=None,
          confidence_array=None):
    """solvers.solve (solvers.py:218-241)."""
    if regularisation is None:
        omega = lambda x: 0.0
        prox = lambda x, _: x
    else:
        omega, prox = regularisation["omega"], regularisation["prox"]
    if x0 is None:
        x0 = np.zeros(A.shape[1])
    if confidence_array is not None:
        confidence_array = np.sqrt(confidence_array)
    return fista(y, A, At, omega, prox, confidence_array, tol_fun, tol_x, max_iter, x0)


def fista(y, A, At, omega, prox, sqrt_W=None, tol_fun=1e-4, tol_x=1e-6, max_iter=1000, x0=None):
    """FISTA with backtracking, restated step by step from solvers.py:244-391."""
    if x0 is None:
        x0 = np.zeros(A.shape[1])
    xhat = x0.copy()
    if sqrt_W is not None:
        res = sqrt_W * (A.dot(xhat) - y)
        grad = np.asarray(At.dot(sqrt_W * res))
    else:
        res = A.dot(xhat) - y
        grad = np.asarray(At.dot(res))
    prox(xhat, 1.0)
    reg_term = omega(xhat)
    prev_obj = 0.5 * np.linalg.norm(res) ** 2 + reg_term
    told = 1
    beta = 0.9
    prev_x = xhat.copy()
    qfval = prev_obj
    if sqrt_W is not None:
        L = (np.linalg.norm(sqrt_W * A.dot(grad)) / np.linalg.norm(grad)) ** 2
    else:
        L = (np.linalg.norm(A.dot(grad)) / np.linalg.norm(grad)) ** 2
    step_size = 1.9 / L
    it = 1
    while True:
        x = xhat - step_size * grad
        prox(x, step_size)
        reg_term_x = omega(x)
        tmp = x - xhat
        q = qfval + np.real(np.dot(tmp, grad)) + 0.5 / step_size * np.linalg.norm(tmp) ** 2 + reg_term_x
        res = sqrt_W * (A.dot(x) - y) if sqrt_W is not None else A.dot(x) - y
        res_norm = np.linalg.norm(res)
        curr_obj = 0.5 * res_norm ** 2 + reg_term_x
        while curr_obj > q:
            step_size = beta * step_size
            x = xhat - step_size * grad
            prox(x, step_size)
            reg_term_x = omega(x)
            tmp = x - xhat
            q = qfval + np.real(np.dot(tmp, grad)) + 0.5 / step_size * np.linalg.norm(tmp) ** 2 + reg_term_x
            res = sqrt_W * (A.dot(x) - y) if sqrt_W is not None else A.dot(x) - y
            res_norm = np.linalg.norm(res)
            curr_obj = 0.5 * res_norm ** 2 + reg_term_x
        abs_obj = abs(curr_obj - prev_obj)
        rel_obj = abs_obj / curr_obj
        abs_x = np.linalg.norm(x - prev_x)
        rel_x = abs_x / (np.linalg.norm(x) + eps)
        if abs_obj < eps:
            criterion = "Absolute tolerance on the objective"
            break
        elif rel_obj < tol_fun:
            criterion = "Relative tolerance on the objective"
            break
        elif abs_x < eps:
            criterion = "Absolute tolerance on the unknowns"
            break
        elif rel_x < tol_x:
            criterion = "Relative tolerance on the unknowns"
            break
        elif it >= max_iter:
            criterion = "Maximum number of iterations"
            break
        t = 0.5 * (1 + sqrt(1 + 4 * told ** 2))
        xhat = x + (told - 1) / t * (x - prev_x)
        if sqrt_W is not None:
            res = sqrt_W * (A.dot(xhat) - y)
            grad = np.asarray(At.dot(sqrt_W * res))
        else:
            res = A.dot(xhat) - y
            grad = np.asarray(At.dot(res))
        it += 1
        prev_obj = curr_obj
        prev_x = x.copy()
        told = t
        qfval = 0.5 * np.linalg.norm(res) ** 2
    details = {
        "residual": 0.5 * res_norm ** 2, "regterm": reg_term_x, "cost_function": curr_obj,
        "abs_cost": abs_obj, "rel_cost": rel_obj, "abs_x": abs_x, "rel_x": rel_x,
        "iterations": it, "stopping_criterion": criterion,
    }
    return x, details


# ------------------------------------------------ Evaluation (core.pyx) --
# Restatements of the pieces of commit.Evaluation that sit around the fit.
def lambda_max_lasso(aty, start, size, w_coeff):
    """compute_lambda_max_lasso (core.pyx:961-966): max |A^T y| / w over
    the compartment's slice [start, start + size)."""
    return np.max(np.abs(aty[start:start + size]) / w_coeff)


def lambda_max_group(aty, w_group, idx_group):
    """compute_lambda_max_group (core.pyx:968-976): max over groups of
    ||(A^T y)[g]|| / w_g."""
    norm_group = np.zeros(w_group.shape, dtype=np.float64)
    for g in range(w_group.size):
        norm_group[g] = np.sqrt(np.sum(aty[idx_group[g]] ** 2)) / w_group[g]
    return np.max(norm_group)


def group_weights(group_idx, kept, nSTR, extra=None, cardinality=True, adaptive=True, x_nnls=None):
    """Group structure and weights of set_regularisation (core.pyx:1123-1192):
    empty groups are dropped; with streamlines left out of the dictionary
    (kept == 0) each group's indices are mapped to the kept ones and groups
    left empty are dropped with their weight; then weights * sqrt(|g|) and
    / (||x_nnls[g]|| + 1e-12).  Returns (groups, weights)."""
    groups = [np.asarray(g) for g in group_idx if np.asarray(g).size > 0]
    w = np.array(extra, dtype=np.float64) if extra is not None else np.ones(len(groups), dtype=np.float64)
    if 0 in kept:
        idx_in_kept = np.zeros(kept.size, dtype=np.int32) - 1
        idx_in_kept[kept == 1] = list(range(nSTR))
        new_g, new_w = [], []
        for count, group in enumerate(groups):
            group = idx_in_kept[group]
            group = np.delete(group, np.where(group == -1)[0])
            if group.size > 0:
                new_g.append(group)
                new_w.append(w[count])
        groups, w = new_g, np.array(new_w, dtype=np.float64)
    if cardinality:
        w = w * np.sqrt(np.array([g.size for g in groups], dtype=np.int32))
    if adaptive:
        w = w / np.array([np.linalg.norm(x_nnls[g]) + 1e-12 for g in groups], dtype=np.float64)
    return groups, w


def debias_eval(x, offset):
    """The debias step's IC column mask (core.pyx:1482-1485): 0 where
    x < 1e-15."""
    mask = np.ones(offset, dtype=np.uint32)
    mask[x[:offset] < 0.000000000000001] = 0
    return mask


def debias_rows(A_masked, x, mask, offset):
    """The debias step's row mask (core.pyx:1494-1501): A x_debias binarised,
    x_debias being x with the masked IC columns and every EC/ISO column set
    to zero, A the operator built with the new eval mask."""
    x_debias = x.copy()
    x_debias[:offset] *= mask
    x_debias[offset:] = 0
    y_mask = np.asarray(A_masked.dot(x_debias))
    y_mask[y_mask < 0] = 0
    y_mask[y_mask > 0] = 1
    return y_mask


def confidence(values, rescale):
    """Confidence values of fit (core.pyx:1433-1457): NaN -> 0; a constant
    map is dropped (None); rescale maps [min, max] to [0, 1].  Returns
    (array or None, whether the values were changed)."""
    a = values.astype(np.float32).copy()
    changed = False
    if np.isnan(a).any():
        a[np.isnan(a)] = 0.0
        changed = True
    cmax, cmin = np.max(a), np.min(a)
    if cmin == cmax:
        return None, False
    if rescale:
        a = (a - cmin) / (cmax - cmin)
        changed = True
    return a, changed


def _rmse_nrmse(y_mea, y_est):
    rmse = np.sqrt(np.mean((y_mea - y_est) ** 2, axis=1))
    den = np.sum(y_mea ** 2, axis=1)
    idx = np.where(den < 1e-12)
    den[idx] = 1
    nrmse = np.sqrt(np.sum((y_mea - y_est) ** 2, axis=1) / den)
    nrmse[idx] = 0
    return rmse, nrmse


def fit_errors(y_dwi_mask, y_fit, y_est_all, nVOX, nS, debias_mask=None, conf=None):
    """Fit error maps and statistics of save_results (core.pyx:1614-1707).
    y_dwi_mask: the DWI at the mask voxels, voxel-major; y_fit: get_y() (the
    signal the fit saw, times the debias mask when set); y_est_all: A x.
    Without debias the maps and statistics use the DWI (as float32) against
    A x (float32).  With debias the statistics cover only the rows in the
    mask and the maps use get_y() against A x.  conf: the confidence values
    at the mask voxels (the RMSE/NRMSE adjusted maps).  Returns a dict of
    per-voxel maps and (mean, std) statistics."""
    out = {}
    if debias_mask is not None:
        nV = int(np.sum(debias_mask) / nS)
        ind = np.where(debias_mask > 0)[0]
        r, n = _rmse_nrmse(np.reshape(y_fit[ind], (nV, -1)), np.reshape(np.asarray(y_est_all)[ind], (nV, -1)))
        out["RMSE_stats"], out["NRMSE_stats"] = (r.mean(), r.std()), (n.mean(), n.std())
        y_mea = np.reshape(y_fit, (nVOX, -1))
        y_est = np.reshape(y_est_all, (nVOX, -1)).astype(np.float32)
        out["RMSE"], out["NRMSE"] = _rmse_nrmse(y_mea, y_est)
    else:
        y_mea = np.reshape(y_dwi_mask.astype(np.float32), (nVOX, -1))
        y_est = np.reshape(y_est_all, (nVOX, -1)).astype(np.float32)
        r, n = _rmse_nrmse(y_mea, y_est)
        out["RMSE"], out["NRMSE"] = r, n
        out["RMSE_stats"], out["NRMSE_stats"] = (r.mean(), r.std()), (n.mean(), n.std())
    if conf is not None:
        ca = np.reshape(conf.astype(np.float32), (nVOX, -1))
        den = np.sum(ca, axis=1)
        idx = np.where(den < 1e-12)
        den[idx] = 1
        t = np.sqrt(np.sum(ca * (y_mea - y_est) ** 2, axis=1) / den)
        t[idx] = 0
        out["RMSE_adjusted"] = t
        den = np.sum(ca * y_mea ** 2, axis=1)
        idx = np.where(den < 1e-12)
        den[idx] = 1
        t = np.sqrt(np.sum(ca * (y_mea - y_est) ** 2, axis=1) / den)
        t[idx] = 0
        out["NRMSE_adjusted"] = t
    return out


def compartment_maps(x, DICTIONARY, KERNELS, nVOX, normalize_kernels=True, normalize_maps=False):
    """Voxelwise compartment contributions of save_results (core.pyx:1709-1756)
    from the fitted x: renormalised by the kernel and fibre norms when the
    kernels were normalised (core.pyx:1595-1601), IC summed over RFs and
    spread over each fibre's segments by length, EC summed over RFs per peak,
    ISO from the first ISO response function's block only (the reference's
    slice x[offset : offset + nISOvoxels]).  normalize_maps divides each by
    the voxel's total + 1e-16.  Returns float32 (IC, EC, ISO)."""
    D, K = DICTIONARY, KERNELS
    nSTR, nE, nI = D["IC"]["nSTR"], D["EC"]["n"], D["ISO"]["n"]
    nIC, nEC = K["wmr"].shape[0], K["wmh"].shape[0]
    norm_fib = np.ones(nSTR)
    if normalize_kernels:
        norm1 = np.repeat(K["wmr_norm"], nSTR)
        norm2 = np.repeat(K["wmh_norm"], nE)
        norm3 = np.repeat(K["iso_norm"], nI)
        norm_fib = np.kron(np.ones(nIC), D["TRK"]["norm"])
        x = x / np.hstack((norm1 * norm_fib, norm2, norm3))
    ic = np.zeros(nVOX, np.float32)
    ec = np.zeros(nVOX, np.float32)
    iso = np.zeros(nVOX, np.float32)
    if nIC > 0:
        off = nSTR * nIC
        t = (x[:off].reshape((-1, nSTR)) * norm_fib.reshape((-1, nSTR))).sum(axis=0)
        ic = np.bincount(D["IC"]["vox"], minlength=nVOX, weights=t[D["IC"]["str"]] * D["IC"]["len"]).astype(np.float32)
    if nEC > 0:
        off = nSTR * nIC
        t = x[off:off + nE * nEC].reshape((-1, nE)).sum(axis=0)
        ec = np.bincount(D["EC"]["vox"], weights=t, minlength=nVOX).astype(np.float32)
    if K["iso"].shape[0] > 0:
        off = nSTR * nIC + nE * nEC
        t = x[off:off + nI].reshape((-1, nI)).sum(axis=0)
        iso = np.bincount(D["ISO"]["vox"], weights=t, minlength=nVOX).astype(np.float32)
    if normalize_maps:
        tot = ic + ec + iso + 1e-16
        return ic / tot, ec / tot, iso / tot
    return ic, ec, iso


def iso_weights(img, DICTIONARY, nISO):
    """Weights of set_wLasso_ISO (core.pyx:828-843): NaN -> 1, the image at
    the mask voxels (compact order) indexed by each ISO entry's voxel, tiled
    over the ISO response functions; every weight must be >= 1."""
    D = DICTIONARY
    w_img = img.astype(np.float32).copy()
    w_img[np.isnan(w_img)] = 1.0
    w = w_img[D["MASK_ix"], D["MASK_iy"], D["MASK_iz"]].flatten().astype(np.float32)
    w = w[D["ISO"]["vox"]]
    if w.size != D["ISO"]["n"]:
        raise ValueError("Number of voxels in the weights image does not match the number of voxels in the dictionary")
    if np.any(w < 1):
        raise ValueError("All weights must be greater or equal to 1")
    if nISO > 1:
        w = np.tile(w, nISO)
    return w
