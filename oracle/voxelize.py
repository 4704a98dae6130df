"""ORACLE — test infrastructure only.

Only ``tests/`` and the ``cpu_baseline`` leg of ``bench.py --mode voxelize``
may import this module.  It wraps ``voxelize_oracle.c``, the CPU restatement
of the reference's tractogram voxeliser (trk2dictionary_c.cpp:111-869 for
streamlines already in memory; see that file's header for the line map).

Pinning: tests/test_oracle_voxelize_ref.py runs the reference's own
trk2dictionary() (compiled from the reference tree into oracle/_ref by
oracle/Makefile; ``reference_trk2dictionary`` below) on the same streamlines
and checks every dictionary file bitwise equal; tests/test_oracle_voxelize.py
checks straight streamlines against their hand-computed voxel crossings,
orientations and norms, random smooth streamlines against a dense-sampling
estimate of each voxel's length, the centroid mode against the polyline
steps, blur replicas of a straight streamline against the offset lines they
must follow, the EC peak rules against hand-derived orientations, and the
ISO and TDI definitions.
"""
import ctypes
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

P = ctypes.c_void_p


class _Params(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int * 3), ("pixdim", ctypes.c_float * 3), ("shift", ctypes.c_float * 3),
                ("min_seg_len", ctypes.c_float), ("min_fiber_len", ctypes.c_float),
                ("max_fiber_len", ctypes.c_float), ("do_intersect", ctypes.c_int),
                ("mask", P), ("iso", P), ("peaks", P), ("Np", ctypes.c_int), ("vf_thr", ctypes.c_float),
                ("flip", ctypes.c_int * 3), ("peaks_affine", P), ("n_rep", ctypes.c_int),
                ("rho", P), ("angle", P), ("weight", P), ("apply_blur", P), ("hash", P),
                ("nthreads", ctypes.c_int)]


class _Result(ctypes.Structure):
    _fields_ = [("IC_f", P), ("IC_v", P), ("IC_o", P), ("IC_len", P), ("n_ic", ctypes.c_int64),
                ("TRK_norm", P), ("TRK_len", P), ("TRK_lenTot", P), ("TRK_kept", P),
                ("n_kept", ctypes.c_int64), ("EC_v", P), ("EC_o", P), ("n_ec", ctypes.c_int64),
                ("ISO_v", P), ("n_iso", ctypes.c_int64), ("tdi", P), ("status", ctypes.c_int)]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libvoxelize_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        lib.vo_run.argtypes = [ctypes.POINTER(_Params), P, P, ctypes.c_int, ctypes.POINTER(_Result)]
        lib.vo_run.restype = ctypes.c_int
        lib.vo_free.argtypes = [ctypes.POINTER(_Result)]
        lib.vo_free.restype = None
        _LIB = lib
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def _take(addr, dtype, count):
    count = int(count)
    if count == 0 or not addr:
        return np.zeros(0, dtype=dtype)
    buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype).copy()


def trk2dictionary(fibers, fib_ptr, dim, pixdim, hash_table, fiber_shift=(0.0, 0.0, 0.0),
                   min_seg_len=1e-3, min_fiber_len=0.0, max_fiber_len=250.0, do_intersect=True,
                   mask=None, iso_mask=None, peaks=None, vf_THR=0.1, flip_peaks=(False, False, False),
                   peaks_affine=None, blur=None, blur_apply_to=None, n_threads=None):
    """Voxelise streamlines on the host cores.

    fibers   float32 [n_points, 3] coordinates in mm (voxmm space)
    fib_ptr  int64 [n_count+1], first point of each streamline
    mask / iso_mask  float32 [Nx, Ny, Nz] (C order) or None
    peaks    float32 [Nx, Ny, Nz, 3*Np] or None
    blur     (rho, angle, weights) float64 arrays, or None for one replica
    Returns a dict of the dictionary arrays (the dictionary_*.dict contents),
    'tdi' (float32 [Nx*Ny*Nz], x fastest) and 'status' (1: a crossing loop
    hit its cap).
    """
    fibers = np.ascontiguousarray(fibers, dtype=np.float32).reshape(-1, 3)
    fib_ptr = np.ascontiguousarray(fib_ptr, dtype=np.int64)
    n_count = fib_ptr.size - 1
    hash_table = np.ascontiguousarray(hash_table, dtype=np.int16)
    if blur is None:
        rho, ang, wts = (np.zeros(1), np.zeros(1), np.ones(1))
    else:
        rho, ang, wts = (np.ascontiguousarray(b, dtype=np.float64) for b in blur)
    mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
    iso_mask = None if iso_mask is None else np.ascontiguousarray(iso_mask, dtype=np.float32)
    Np = 0
    if peaks is not None:
        peaks = np.ascontiguousarray(peaks, dtype=np.float32)
        Np = peaks.shape[3] // 3
        peaks_affine = np.eye(3) if peaks_affine is None else peaks_affine
        peaks_affine = np.ascontiguousarray(peaks_affine, dtype=np.float64)
    apply = None if blur_apply_to is None else np.ascontiguousarray(blur_apply_to, dtype=np.uint8)
    sh = np.broadcast_to(np.asarray(fiber_shift, dtype=np.float32), (3,))
    p = _Params()
    p.dim[:] = [int(d) for d in dim]
    p.pixdim[:] = [float(x) for x in pixdim]
    p.shift[:] = [float(x) for x in sh]
    p.min_seg_len, p.min_fiber_len, p.max_fiber_len = min_seg_len, min_fiber_len, max_fiber_len
    p.do_intersect = 1 if do_intersect else 0
    p.mask, p.iso, p.peaks = _ptr(mask), _ptr(iso_mask), _ptr(peaks)
    p.Np = Np
    p.vf_thr = vf_THR
    p.flip[:] = [-1 if f else 1 for f in flip_peaks]
    p.peaks_affine = _ptr(peaks_affine) if peaks is not None else None
    p.n_rep = rho.size
    p.rho, p.angle, p.weight = _ptr(rho), _ptr(ang), _ptr(wts)
    p.apply_blur = _ptr(apply)
    p.hash = _ptr(hash_table)
    if n_threads is None:
        try:
            n_threads = len(os.sched_getaffinity(0))
        except AttributeError:
            n_threads = os.cpu_count() or 1
    p.nthreads = max(1, min(255, int(n_threads)))
    r = _Result()
    rc = _lib().vo_run(ctypes.byref(p), fibers.ctypes.data_as(P), fib_ptr.ctypes.data_as(P), n_count,
                       ctypes.byref(r))
    if rc < 0:
        raise MemoryError("voxelize oracle: allocation failed")
    nvox = int(np.prod([int(d) for d in dim]))
    out = {
        "IC_f": _take(r.IC_f, np.uint32, r.n_ic), "IC_v": _take(r.IC_v, np.uint32, r.n_ic),
        "IC_o": _take(r.IC_o, np.uint16, r.n_ic), "IC_len": _take(r.IC_len, np.float32, r.n_ic),
        "TRK_norm": _take(r.TRK_norm, np.float32, r.n_kept), "TRK_len": _take(r.TRK_len, np.float32, r.n_kept),
        "TRK_lenTot": _take(r.TRK_lenTot, np.float32, r.n_kept),
        "TRK_kept": _take(r.TRK_kept, np.uint8, n_count),
        "EC_v": _take(r.EC_v, np.uint32, r.n_ec), "EC_o": _take(r.EC_o, np.uint16, r.n_ec),
        "ISO_v": _take(r.ISO_v, np.uint32, r.n_iso), "tdi": _take(r.tdi, np.float32, nvox),
        "status": int(r.status),
    }
    _lib().vo_free(ctypes.byref(r))
    return out


# ------------------------------------------------ the reference's own code --
_REF = None


def reference_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libtrk2dictionary_ref.so"))


def _ref_lib():
    """trk2dictionary() of the reference (trk2dictionary_c.cpp:111-286),
    compiled from the reference tree by oracle/Makefile into _ref/."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libtrk2dictionary_ref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle, with the reference tree present)")
        with open(os.path.join(_HERE, "_ref", "trk2dictionary.sym")) as f:
            sym = f.read().split()[0]
        lib = ctypes.CDLL(path)
        fn = getattr(lib, sym)
        I, F, C = ctypes.c_int, ctypes.c_float, ctypes.c_char_p
        fn.argtypes = [C, I, I, I, I, F, F, F, I, I, I, F, F, F, F, F, F, P, I, F, I, I, I, P, P,
                       ctypes.POINTER(ctypes.c_void_p), C, I, P, I, P, P, P, P, P, P, I, I]
        fn.restype = I
        _REF = fn
    return _REF


def _write_trk(path, fibers, fib_ptr, dim, pixdim):
    hdr = bytearray(1000)
    hdr[0:6] = b"TRACK\0"
    np.frombuffer(hdr, "<i2", 3, 6)[:] = np.asarray(dim, np.int16)
    np.frombuffer(hdr, "<f4", 3, 12)[:] = np.asarray(pixdim, np.float32)
    np.frombuffer(hdr, "<i4", 1, 988)[0] = len(fib_ptr) - 1
    np.frombuffer(hdr, "<i4", 1, 992)[0] = 2
    np.frombuffer(hdr, "<i4", 1, 996)[0] = 1000
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        for i in range(len(fib_ptr) - 1):
            pts = np.ascontiguousarray(fibers[fib_ptr[i]:fib_ptr[i + 1]], np.float32)
            f.write(np.int32(pts.shape[0]).tobytes())
            f.write(pts.astype("<f4").tobytes())


def reference_trk2dictionary(workdir, fibers, fib_ptr, dim, pixdim, hash_table, fiber_shift=(0.0, 0.0, 0.0),
                             min_seg_len=1e-3, min_fiber_len=0.0, max_fiber_len=250.0, do_intersect=True,
                             mask=None, iso_mask=None, peaks=None, vf_THR=0.1, flip_peaks=(False, False, False),
                             peaks_affine=None, blur=None, blur_apply_to=None, n_threads=1):
    """Run the reference's own trk2dictionary() on these streamlines (written
    to a .trk file in workdir) and read back what it writes, concatenated and
    renumbered as trk2dictionary.pyx:551-612 does.  Returns the same keys as
    trk2dictionary() above, except 'tdi': 'has_ic' marks the voxels its
    per-thread TDI buffers saw (the presence test of :291-302 and :420-428)."""
    os.makedirs(workdir, exist_ok=True)
    fibers = np.ascontiguousarray(fibers, np.float32).reshape(-1, 3)
    fib_ptr = np.ascontiguousarray(fib_ptr, np.int64)
    n = fib_ptr.size - 1
    trk = os.path.join(workdir, "t.trk")
    _write_trk(trk, fibers, fib_ptr, dim, pixdim)
    temp = os.path.join(workdir, "temp")
    os.makedirs(temp, exist_ok=True)
    nvox = int(np.prod([int(d) for d in dim]))
    nt = max(1, min(int(n_threads), n if n > 0 else 1))
    tdis = [np.zeros(nvox, np.float64) for _ in range(nt)]
    tptr = (ctypes.c_void_p * nt)(*[t.ctypes.data for t in tdis])
    if blur is None:
        rho, ang, wts = np.zeros(1), np.zeros(1), np.ones(1)
    else:
        rho, ang, wts = (np.ascontiguousarray(b, np.float64) for b in blur)
    apply = np.ones(max(n, 1), np.uint8) if blur_apply_to is None else \
        np.ascontiguousarray(blur_apply_to, np.uint8)
    mask = None if mask is None else np.ascontiguousarray(mask, np.float32)
    iso_mask = None if iso_mask is None else np.ascontiguousarray(iso_mask, np.float32)
    Np = 0
    aff = None
    if peaks is not None:
        peaks = np.ascontiguousarray(peaks, np.float32)
        Np = peaks.shape[3] // 3
        aff = np.ascontiguousarray(np.eye(3) if peaks_affine is None else peaks_affine, np.float64)
    sh = np.broadcast_to(np.asarray(fiber_shift, np.float32), (3,))
    hash_table = np.ascontiguousarray(hash_table, np.int16)
    # the reference prints its progress on the C stdout: send it to stderr, so
    # that a caller's own stdout (a bench's one JSON line) stays clean
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        t0 = time.perf_counter()
        ret = _ref_lib()(trk.encode(), 1000, int(dim[0]), int(dim[1]), int(dim[2]), float(pixdim[0]),
                         float(pixdim[1]), float(pixdim[2]), n, 0, 0, float(sh[0]), float(sh[1]), float(sh[2]),
                         float(min_seg_len), float(min_fiber_len), float(max_fiber_len), _ptr(peaks), Np,
                         float(vf_THR), -1 if flip_peaks[0] else 1, -1 if flip_peaks[1] else 1,
                         -1 if flip_peaks[2] else 1, _ptr(mask), _ptr(iso_mask), tptr, temp.encode(),
                         1 if do_intersect else 0, _ptr(aff), int(rho.size), _ptr(rho), _ptr(ang), _ptr(wts),
                         _ptr(apply), None, _ptr(hash_table), nt, 0)
        seconds = time.perf_counter() - t0
    finally:
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
    if ret != 1:
        raise RuntimeError("the reference's trk2dictionary() generated no dictionary")
    rd = lambda name, dt: np.fromfile(name, dtype=dt)
    out = {}
    parts = {k: [] for k in ("IC_f", "IC_v", "IC_o", "IC_len", "TRK_norm", "TRK_len", "TRK_lenTot", "TRK_kept")}
    dts = {"IC_f": "<u4", "IC_v": "<u4", "IC_o": "<u2", "IC_len": "<f4", "TRK_norm": "<f4", "TRK_len": "<f4",
           "TRK_lenTot": "<f4", "TRK_kept": "u1"}
    discarded = 0
    for j in range(nt):
        for k in parts:
            a = rd(os.path.join(temp, f"dictionary_{k}_{j}.dict"), dts[k])
            if k == "IC_f" and j > 0:
                a = (a.astype(np.int64) - discarded).astype(np.uint32)
            parts[k].append(a)
        discarded += int(np.count_nonzero(parts["TRK_kept"][j] == 0))
    for k in parts:
        out[k] = np.concatenate(parts[k]).astype(np.dtype(dts[k]).newbyteorder("="))
    out["EC_v"] = rd(os.path.join(workdir, "dictionary_EC_v.dict"), "<u4").astype(np.uint32)
    out["EC_o"] = rd(os.path.join(workdir, "dictionary_EC_o.dict"), "<u2").astype(np.uint16)
    out["ISO_v"] = rd(os.path.join(workdir, "dictionary_ISO_v.dict"), "<u4").astype(np.uint32)
    X, Y, Z = (int(d) for d in dim)
    has = np.zeros((X, Y, Z), bool)
    for t in tdis:
        has |= t.reshape(X, Y, Z) != 0          # C order [x][y][z] (:528)
    out["has_ic"] = has.ravel(order="F")          # x fastest, as the linear voxel ids
    out["seconds"] = seconds                      # the trk2dictionary() call alone
    return out
