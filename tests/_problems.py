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
