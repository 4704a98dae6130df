"""ORACLE — test infrastructure only.

Only ``tests/`` and the ``cpu_baseline`` leg of ``bench.py --mode voxelize``
may import this module.  It wraps ``voxelize_oracle.c``, the CPU restatement
of the reference's tractogram voxeliser (trk2dictionary_c.cpp:101-869 for
streamlines already in memory; see that file's header for the line map).

Pinning (tests/test_oracle_voxelize.py): straight streamlines through the
grid against their hand-computed voxel crossings, orientations and norms;
random smooth streamlines against a dense-sampling estimate of each (voxel,
orientation) length; the centroid mode against the polyline steps; blur
replicas of a straight streamline against the offset lines they must follow;
the EC peak selection, hemisphere flip, axis flips and affine against
hand-derived orientations; the ISO and TDI definitions.
"""
import ctypes
import os

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
