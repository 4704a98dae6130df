"""Small hand-made dictionaries shared by the CPU and GPU tests."""
import numpy as np


def random_dictionary(seed, nVOX=6, nF=5, n=14, nS=7, ndirs=4, nIC=1, nEC=1, nISO=2, nE=None,
                      empty_voxels=(), ec_per_voxel=1):
    """Random dictionary, voxel-sorted as load_dictionary leaves it, with
    random (positive and negative) LUTs.  Every fibre has >= 1 segment."""
    rng = np.random.default_rng(seed)
    vox_pool = np.array([v for v in range(nVOX) if v not in set(empty_voxels)], dtype=np.uint32)
    f = np.concatenate([np.arange(nF), rng.integers(0, nF, max(0, n - nF))]).astype(np.uint32)[:n]
    v = rng.choice(vox_pool, n).astype(np.uint32)
    o = rng.integers(0, ndirs, n).astype(np.uint16)
    l = rng.uniform(0.05, 1.2, n).astype(np.float32)
    idx = np.argsort(v, kind="mergesort")
    D = {"IC": {"vox": v[idx], "str": f[idx], "dir": o[idx], "len": l[idx], "n": int(n), "nSTR": int(nF),
                "nVOX": int(nVOX)},
         "TRK": {"norm": np.ones(nF, np.float32), "kept": np.ones(nF, np.uint8)}}
    D["IC"]["eval"] = np.ones(nF * nIC, dtype=np.uint32)
    if nEC > 0:
        ev = np.repeat(vox_pool, ec_per_voxel).astype(np.uint32)
        D["EC"] = {"vox": ev, "dir": rng.integers(0, ndirs, ev.size).astype(np.uint16), "n": int(ev.size)}
    else:
        D["EC"] = {"vox": np.zeros(0, np.uint32), "dir": np.zeros(0, np.uint16), "n": 0}
    if nISO > 0:
        iv = vox_pool.copy()
        D["ISO"] = {"vox": iv, "n": int(iv.size)}
    else:
        D["ISO"] = {"vox": np.zeros(0, np.uint32), "n": 0}
    K = {"wmr": rng.standard_normal((nIC, ndirs, nS)).astype(np.float32),
         "wmh": rng.standard_normal((nEC, ndirs, nS)).astype(np.float32),
         "iso": rng.standard_normal((nISO, nS)).astype(np.float32)}
    return D, K


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def load_golden(name="operator_small"):
    """Committed fixture (tests/golden/make_golden.py): dictionary, tables,
    x, y and the expected A x, A^T y and table gradient (gy = y)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".npz"))
    nIC = g["wmr"].shape[0]
    D = {"IC": {"vox": g["IC_vox"], "str": g["IC_str"], "dir": g["IC_dir"], "len": g["IC_len"],
                "eval": g["IC_eval"], "n": int(g["IC_vox"].size), "nSTR": int(g["nF"]), "nVOX": int(g["nVOX"])},
         "EC": {"vox": g["EC_vox"], "dir": g["EC_dir"], "n": int(g["EC_vox"].size)},
         "ISO": {"vox": g["ISO_vox"], "n": int(g["ISO_vox"].size)},
         "TRK": {"norm": np.ones(int(g["nF"]), np.float32), "kept": np.ones(int(g["nF"]), np.uint8)}}
    K = {"wmr": g["wmr"], "wmh": g["wmh"], "iso": g["iso"]}
    assert D["IC"]["eval"].size == int(g["nF"]) * nIC
    return D, K, {k: g[k] for k in ("x", "y", "Ax", "Aty", "gwmr", "gwmh", "giso")}


def runs_dictionary(seed, nS=288, nIC=1, nEC=1, nISO=2):
    """Voxels whose entries form runs of one orientation of every length from
    1 to 200: runs that cross 64-entry chunks and groups of 8, voxels of a
    single run, and runs of one orientation that straddle slab boundaries
    (orientations 99/100/101 with 100-orientation slabs)."""
    rng = np.random.default_rng(seed)
    nVOX, ndirs = 40, 500
    vox, ori = [], []
    for v in range(nVOX):
        if v == 5:
            continue                                  # an empty voxel
        if v % 7 == 0:
            lens = [int(rng.integers(60, 200))]       # a single long run
        else:
            lens = list(rng.integers(1, 12, size=int(rng.integers(3, 25))))
            lens[int(rng.integers(0, len(lens)))] = int(rng.integers(65, 140))
        dirs = rng.choice(ndirs, size=len(lens), replace=False)
        dirs[: min(3, len(dirs))] = [99, 100, 101][: min(3, len(dirs))]
        for L, o in zip(lens, dirs):
            vox += [v] * int(L)
            ori += [int(o)] * int(L)
    n = len(vox)
    nF = n // 6
    D, K = random_dictionary(seed, nVOX=nVOX, nF=nF, n=n, nS=nS, ndirs=ndirs, nIC=nIC, nEC=nEC, nISO=nISO,
                             empty_voxels=(5,))
    perm = rng.permutation(n)                         # entries of a run from many fibres, any order
    D["IC"]["vox"] = np.asarray(vox, np.uint32)
    D["IC"]["dir"] = np.asarray(ori, np.uint16)
    D["IC"]["str"] = D["IC"]["str"][perm]
    return D, K

