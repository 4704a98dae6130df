"""Writes tests/golden/operator_small.npz: a small seeded dictionary, its
tables, input vectors and the expected A x, A^T y and table gradient,
evaluated here by explicit loops over the closed-form entries of the COMMIT
operator (setup_operator.py:62-218):

    A[v*nS+s, i*nF+f]            += l * eval[i*nF+f] * wmr[i,o,s]   per IC segment
    A[v*nS+s, nIC*nF + i*nE+e]   += wmh[i,o_e,s]                    per EC peak
    A[v*nS+s, nIC*nF+nEC*nE+i*nI+k] += iso[i,s]                     per ISO voxel

in float64, independently of the package and of the oracle.  Run from the
repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np


def build(seed=20261203, nVOX=30, nF=40, n=400, nS=60, ndirs=12, nIC=2, nEC=1, nISO=2):
    rng = np.random.default_rng(seed)
    v = np.sort(rng.integers(0, nVOX, n)).astype(np.uint32)           # voxel-sorted, as load_dictionary leaves it
    f = rng.integers(0, nF, n).astype(np.uint32)
    f[:nF] = np.arange(nF)                                              # every fibre has a segment
    o = rng.integers(0, ndirs, n).astype(np.uint16)
    l = rng.uniform(0.05, 1.2, n).astype(np.float32)
    ev = np.ones(nF * nIC, np.uint32)
    ev[rng.choice(nF * nIC, 5, replace=False)] = 0
    ecv = np.repeat(np.arange(nVOX, dtype=np.uint32), 2)
    eco = rng.integers(0, ndirs, ecv.size).astype(np.uint16)
    isov = np.arange(nVOX, dtype=np.uint32)
    wmr = rng.standard_normal((nIC, ndirs, nS)).astype(np.float32)
    wmh = rng.standard_normal((nEC, ndirs, nS)).astype(np.float32)
    iso = rng.standard_normal((nISO, nS)).astype(np.float32)
    nE, nI = ecv.size, isov.size
    ncols = nF * nIC + nE * nEC + nI * nISO
    x = rng.standard_normal(ncols)
    y = rng.standard_normal(nVOX * nS)
    Ax = np.zeros(nVOX * nS)
    Aty = np.zeros(ncols)
    gwmr = np.zeros(wmr.shape)
    gwmh = np.zeros(wmh.shape)
    giso = np.zeros(iso.shape)
    for k in range(n):
        for i in range(nIC):
            c = i * nF + int(f[k])
            for s in range(nS):
                a = float(l[k]) * float(ev[c]) * float(wmr[i, o[k], s])
                Ax[v[k] * nS + s] += a * x[c]
                Aty[c] += a * y[v[k] * nS + s]
                gwmr[i, o[k], s] += float(l[k]) * float(ev[c]) * x[c] * y[v[k] * nS + s]
    for e in range(nE):
        for i in range(nEC):
            c = nIC * nF + i * nE + e
            for s in range(nS):
                a = float(wmh[i, eco[e], s])
                Ax[ecv[e] * nS + s] += a * x[c]
                Aty[c] += a * y[ecv[e] * nS + s]
                gwmh[i, eco[e], s] += x[c] * y[ecv[e] * nS + s]
    for kk in range(nI):
        for i in range(nISO):
            c = nIC * nF + nEC * nE + i * nI + kk
            for s in range(nS):
                a = float(iso[i, s])
                Ax[isov[kk] * nS + s] += a * x[c]
                Aty[c] += a * y[isov[kk] * nS + s]
                giso[i, s] += x[c] * y[isov[kk] * nS + s]
    return dict(IC_vox=v, IC_str=f, IC_dir=o, IC_len=l, IC_eval=ev, EC_vox=ecv, EC_dir=eco, ISO_vox=isov,
                wmr=wmr, wmh=wmh, iso=iso, nVOX=np.int64(nVOX), nF=np.int64(nF), x=x, y=y, Ax=Ax, Aty=Aty,
                gwmr=gwmr, gwmh=gwmh, giso=giso)


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "operator_small.npz")
    np.savez_compressed(out, **build())
    print("wrote", out)
