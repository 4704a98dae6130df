"""Tractogram -> COMMIT dictionary, voxelised on the GPU.

Drop-in for ``commit.trk2dictionary.run`` (trk2dictionary.pyx:57-654): reads a
.trk or .tck tractogram and the optional mask / ISO mask / peaks NIfTI files,
voxelises every streamline on the device (``commit_hip_trk2dictionary``,
include/commit_hip.h) and writes the files ``Evaluation.load_dictionary``
reads: dictionary_{IC_f, IC_v, IC_o, IC_len, TRK_norm, TRK_len, TRK_lenTot,
TRK_kept, EC_v, EC_o, ISO_v}.dict, dictionary_tdi.nii.gz,
dictionary_mask.nii.gz and dictionary_info.pickle.

``voxelize`` is the in-memory entry point (streamline arrays in, dictionary
arrays out).  There is no CPU path: without libcommit_hip.so this raises.

Differences from the reference, all outside the voxeliser's arithmetic:
* the orientation hash table is built from this package's own direction set
  (``synth.hemisphere_directions`` / ``synth.hash_table``), since amico's
  precomputed tables are not available; pass ``hash_table`` to use another;
* streamline clustering (``blur_clust_thr > 0``) needs dicelib and raises;
* TRK, TCK and NIfTI files are read by this package's own readers (no nibabel).
"""
import ctypes
import os
import pickle
import shutil
import time
from os.path import dirname, exists, isdir, join, splitext

import numpy as np

from . import _lib, nifti

MAX_FIB_LEN = 10000          # trk2dictionary_c.cpp:16


# ------------------------------------------------------------------ blur --
def blur_replicas(blur_core_extent=0.0, blur_gauss_extent=0.0, blur_spacing=0.25, blur_gauss_min=0.1):
    """Replica offsets (rho, angle) and weights of the streamline blur, as
    trk2dictionary.pyx:210-239 builds them.  Returns (rho, angle, weights, sigma)."""
    if blur_core_extent < 0:
        raise ValueError("The extent of the core must be non-negative")
    if blur_gauss_extent < 0:
        raise ValueError("The extent of the blur must be non-negative")
    if (blur_gauss_extent > 0 or blur_core_extent > 0) and blur_spacing <= 0:
        raise ValueError("The grid spacing of the blur must be positive")
    sigma = 0.0
    if (blur_gauss_extent == 0 and blur_core_extent == 0) or blur_spacing == 0:
        return np.array([0.0]), np.array([0.0]), np.array([1.0]), sigma
    tmp = np.arange(0, blur_core_extent + blur_gauss_extent + 1e-6, blur_spacing)
    tmp = np.concatenate((tmp, -tmp[1:][::-1]))
    x, y = np.meshgrid(tmp, tmp)
    r = np.sqrt(x * x + y * y)
    idx = r <= blur_core_extent + blur_gauss_extent
    rho = r[idx]
    angle = np.arctan2(y, x)[idx]
    weights = np.empty(rho.size, np.double)
    if blur_gauss_extent == 0:
        weights[:] = 1.0
    else:
        sigma = float(np.float32(blur_gauss_extent / np.sqrt(-2.0 * np.log(blur_gauss_min))))
        for i in range(rho.size):
            if rho[i] <= blur_core_extent:
                weights[i] = 1.0
            else:
                weights[i] = np.exp(-(rho[i] - blur_core_extent) ** 2 / (2.0 * sigma ** 2))
    return rho, angle, weights, sigma


def default_hash_table(ndirs=500):
    """181 x 181 (colatitude, longitude) degree table -> nearest of this
    package's ``ndirs`` hemisphere directions."""
    from . import synth
    return synth.hash_table(synth.hemisphere_directions(ndirs))


# -------------------------------------------------------------- readers --
def _trk_header(raw):
    if raw[:5] != b"TRACK":
        raise ValueError("not a TrackVis .trk file")
    endian = "<" if np.frombuffer(raw, "<i4", 1, 996)[0] == 1000 else ">"
    if np.frombuffer(raw, endian + "i4", 1, 996)[0] != 1000:
        raise ValueError("TrackVis header size is not 1000")
    h = {
        "dimensions": np.frombuffer(raw, endian + "i2", 3, 6).astype(int),
        "voxel_sizes": np.frombuffer(raw, endian + "f4", 3, 12).astype(np.float32),
        "nb_scalars_per_point": int(np.frombuffer(raw, endian + "i2", 1, 36)[0]),
        "nb_properties_per_streamline": int(np.frombuffer(raw, endian + "i2", 1, 238)[0]),
        "vox_to_ras": np.frombuffer(raw, endian + "f4", 16, 440).reshape(4, 4).astype(np.float64),
        "nb_streamlines": int(np.frombuffer(raw, endian + "i4", 1, 988)[0]),
        "endian": endian,
    }
    return h


def read_trk(filename):
    """Streamlines of a .trk file as (fibers float32 [n_points, 3] in voxmm,
    fib_ptr int64 [n+1], header), the coordinates read_fiberTRK returns
    (trk2dictionary_c.cpp:873-893).  A streamline of MAX_FIB_LEN or more
    points makes the reference generate no dictionary (:195): ValueError."""
    with open(filename, "rb") as f:
        raw = f.read()
    h = _trk_header(raw[:1000])
    e = h["endian"]
    ns, npr, n = h["nb_scalars_per_point"], h["nb_properties_per_streamline"], h["nb_streamlines"]
    stride = 3 + ns
    counts = np.empty(n, np.int64)
    starts = np.empty(n, np.int64)
    pos = 1000
    for i in range(n):
        N = int(np.frombuffer(raw, e + "i4", 1, pos)[0])
        if N >= MAX_FIB_LEN or N <= 0:
            raise ValueError(f"streamline {i} has {N} points (supported: 1..{MAX_FIB_LEN - 1})")
        counts[i] = N
        starts[i] = pos + 4
        pos += 4 + 4 * (N * stride + npr)
    fib_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=fib_ptr[1:])
    fibers = np.empty((int(fib_ptr[-1]), 3), np.float32)
    for i in range(n):
        pts = np.frombuffer(raw, e + "f4", int(counts[i]) * stride, int(starts[i])).reshape(-1, stride)
        fibers[fib_ptr[i]:fib_ptr[i + 1]] = pts[:, :3]
    return fibers, fib_ptr, h


def write_trk(filename, fibers, fib_ptr, dim, voxel_size, vox_to_ras=None):
    """Minimal TrackVis v2 writer (no scalars or properties), coordinates in voxmm."""
    hdr = bytearray(1000)
    hdr[0:6] = b"TRACK\0"
    np.frombuffer(hdr, "<i2", 3, 6)[:] = np.asarray(dim, np.int16)
    np.frombuffer(hdr, "<f4", 3, 12)[:] = np.asarray(voxel_size, np.float32)
    if vox_to_ras is None:
        vox_to_ras = np.diag([float(v) for v in voxel_size] + [1.0])
    np.frombuffer(hdr, "<f4", 16, 440)[:] = np.asarray(vox_to_ras, np.float32).ravel()
    hdr[948:952] = b"RAS\0"
    n = len(fib_ptr) - 1
    np.frombuffer(hdr, "<i4", 1, 988)[0] = n
    np.frombuffer(hdr, "<i4", 1, 992)[0] = 2
    np.frombuffer(hdr, "<i4", 1, 996)[0] = 1000
    fibers = np.ascontiguousarray(fibers, np.float32).reshape(-1, 3)
    with open(filename, "wb") as f:
        f.write(bytes(hdr))
        for i in range(n):
            pts = fibers[fib_ptr[i]:fib_ptr[i + 1]]
            f.write(np.int32(pts.shape[0]).astype("<i4").tobytes())
            f.write(pts.astype("<f4").tobytes())


def _tck_header(filename):
    with open(filename, "rb") as f:
        first = f.readline()
        if first.strip() != b"mrtrix tracks":
            raise ValueError(f"{filename}: not an MRtrix .tck file")
        h = {}
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{filename}: header without END")
            line = line.strip()
            if line == b"END":
                break
            if b":" in line:
                k, v = line.split(b":", 1)
                h[k.strip().decode()] = v.strip().decode()
    if "file" not in h:
        raise ValueError(f"{filename}: no 'file' entry")
    fparts = h["file"].split()
    h["_offset_data"] = int(fparts[1])
    dt = h.get("datatype", "Float32LE")
    if dt not in ("Float32LE", "Float32BE"):
        raise ValueError(f"{filename}: unsupported datatype {dt}")
    h["_dtype"] = "<f4" if dt.endswith("LE") else ">f4"
    return h


def read_tck(filename, toVOXMM):
    """Streamlines of a .tck file mapped to voxmm with the float32 3x4 matrix
    toVOXMM, evaluated as read_fiberTCK does (trk2dictionary_c.cpp:896-911):
    ((p0*m0 + p1*m1) + p2*m2) + m3 in float arithmetic."""
    h = _tck_header(filename)
    data = np.fromfile(filename, dtype=h["_dtype"], offset=h["_offset_data"]).astype(np.float32).reshape(-1, 3)
    n_count = int(h.get("count", "0"))
    end = np.nonzero(np.isinf(data[:, 0]))[0]
    if end.size:
        data = data[:end[0]]
    sep = np.nonzero(np.isnan(data[:, 0]) | np.isnan(data[:, 1]) | np.isnan(data[:, 2]))[0]
    if sep.size < n_count:
        raise ValueError(f"{filename}: {sep.size} streamlines found, header count {n_count}")
    sep = sep[:n_count]
    starts = np.concatenate(([0], sep[:-1] + 1)) if n_count else np.zeros(0, np.int64)
    counts = sep - starts
    if np.any(counts >= MAX_FIB_LEN):
        raise ValueError(f"a streamline has {MAX_FIB_LEN} or more points")
    keep = np.ones(data.shape[0], bool)
    keep[sep] = False
    if n_count:
        keep[sep[-1] + 1:] = False
    else:
        keep[:] = False
    P = data[keep]
    M = np.asarray(toVOXMM, np.float32).reshape(-1)
    f32 = np.float32
    out = np.empty_like(P)
    for r in range(3):
        m0, m1, m2, m3 = (f32(M[4 * r + j]) for j in range(4))
        out[:, r] = ((P[:, 0] * m0 + P[:, 1] * m1) + P[:, 2] * m2) + m3
    fib_ptr = np.zeros(n_count + 1, np.int64)
    np.cumsum(counts, out=fib_ptr[1:])
    return out, fib_ptr, h


def write_tck(filename, fibers_ras, fib_ptr):
    """Minimal MRtrix .tck writer (Float32LE, world coordinates)."""
    n = len(fib_ptr) - 1
    head = f"mrtrix tracks\ncount: {n}\ndatatype: Float32LE\nfile: . {{off}}\nEND\n"
    off = len(head.format(off=0)) + 8
    head = head.format(off=off)
    head = head + " " * (off - len(head))
    parts = []
    fibers_ras = np.asarray(fibers_ras, np.float32).reshape(-1, 3)
    nan = np.full((1, 3), np.nan, np.float32)
    for i in range(n):
        parts.append(fibers_ras[fib_ptr[i]:fib_ptr[i + 1]])
        parts.append(nan)
    parts.append(np.full((1, 3), np.inf, np.float32))
    with open(filename, "wb") as f:
        f.write(head.encode())
        f.write(np.concatenate(parts).astype("<f4").tobytes())


def tck_to_voxmm(affine, pixdim):
    """toVOXMM of trk2dictionary.pyx:406-411: the inverse of the reference
    image's affine with the voxel scaling removed, as float32 [12]."""
    M = np.asarray(affine, np.float64).copy()
    M[:3, :3] = M[:3, :3].dot(np.diag([np.float64(1) / pixdim[0], np.float64(1) / pixdim[1],
                                       np.float64(1) / pixdim[2]]))
    return np.ravel(np.linalg.inv(M)).astype("<f4")


# ------------------------------------------------------------- device ---
class Voxeliser:
    """A device voxeliser (commit_hip_vox); keeps its buffers between runs."""

    def __init__(self, device=0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.commit_hip_vox_create(ctypes.byref(h), int(device)), "commit_hip_vox_create")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.commit_hip_vox_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, fibers, fib_ptr, dim, pixdim, hash_table, fiber_shift=(0.0, 0.0, 0.0), min_seg_len=1e-3,
            min_fiber_len=0.0, max_fiber_len=250.0, do_intersect=True, mask=None, iso_mask=None, peaks=None,
            vf_THR=0.1, flip_peaks=(False, False, False), peaks_affine=None, blur=None, blur_apply_to=None,
            stream=None, n_points=None):
        """Voxelise.  fibers/fib_ptr are numpy arrays, or torch CUDA tensors
        (float32 [n_points, 3] and int64 [n+1]) already in device memory; then
        mask, iso_mask and peaks are used (or copied) as device tensors too."""
        on_dev = hasattr(fibers, "data_ptr")
        if on_dev:
            n_count = int(fib_ptr.numel()) - 1
            n_points = int(fibers.shape[0]) if n_points is None else int(n_points)
            pf, pp = ctypes.c_void_p(fibers.data_ptr()), ctypes.c_void_p(fib_ptr.data_ptr())
            keep = ()
        else:
            fibers = np.ascontiguousarray(fibers, np.float32).reshape(-1, 3)
            fib_ptr = np.ascontiguousarray(fib_ptr, np.int64)
            n_count = fib_ptr.size - 1
            n_points = fibers.shape[0]
            pf, pp = _lib.ptr(fibers), _lib.ptr(fib_ptr)
            keep = (fibers, fib_ptr)
        hash_table = np.ascontiguousarray(hash_table, np.int16)
        if hash_table.size != 181 * 181:
            raise ValueError("the orientation hash table must have 181*181 entries")
        if blur is None:
            rho, ang, wts = np.zeros(1), np.zeros(1), np.ones(1)
        else:
            rho, ang, wts = (np.ascontiguousarray(b, np.float64) for b in blur[:3])
        as_f32 = (lambda a: a.float().contiguous()) if on_dev else (lambda a: np.ascontiguousarray(a, np.float32))
        istensor = lambda a: hasattr(a, "data_ptr")
        mask = None if mask is None else (as_f32(mask) if istensor(mask) else np.ascontiguousarray(mask, np.float32))
        iso_mask = None if iso_mask is None else (as_f32(iso_mask) if istensor(iso_mask) else
                                                  np.ascontiguousarray(iso_mask, np.float32))
        Np = 0
        aff = None
        if peaks is not None:
            peaks = as_f32(peaks) if istensor(peaks) else np.ascontiguousarray(peaks, np.float32)
            if peaks.ndim != 4 or peaks.shape[3] % 3:
                raise ValueError("PEAKS dataset must have 3*k volumes")
            Np = peaks.shape[3] // 3
            aff = np.ascontiguousarray(np.eye(3) if peaks_affine is None else peaks_affine, np.float64)
        apply = None if blur_apply_to is None else np.ascontiguousarray(blur_apply_to, np.uint8)
        if apply is not None and apply.size != n_count:
            raise ValueError('"blur_apply_to" must have one value per streamline')
        sh = np.broadcast_to(np.asarray(fiber_shift, np.float32), (3,))
        dim = [int(d) for d in dim]
        for name, a in (("mask", mask), ("iso_mask", iso_mask)):
            if a is not None and tuple(a.shape) != tuple(dim):
                raise ValueError(f"{name} shape {a.shape} does not match the grid {tuple(dim)}")
        if peaks is not None and tuple(peaks.shape[:3]) != tuple(dim):
            raise ValueError(f"peaks shape {peaks.shape[:3]} does not match the grid {tuple(dim)}")
        dptr = _lib.ptr
        if on_dev:
            import torch
            held = []

            def dptr(a):
                if a is None:
                    return None
                t = a if hasattr(a, "data_ptr") else torch.from_numpy(np.ascontiguousarray(a)).to(fibers.device)
                held.append(t)
                return ctypes.c_void_p(t.data_ptr())
        f = ctypes.c_float
        rc = self._lib.commit_hip_trk2dictionary(
            self._h, pf, pp, n_count, n_points, 1 if on_dev else 0, dim[0], dim[1], dim[2],
            f(pixdim[0]), f(pixdim[1]), f(pixdim[2]), f(sh[0]), f(sh[1]), f(sh[2]),
            f(min_seg_len), f(min_fiber_len), f(max_fiber_len), dptr(peaks), Np, f(vf_THR),
            -1 if flip_peaks[0] else 1, -1 if flip_peaks[1] else 1, -1 if flip_peaks[2] else 1,
            dptr(mask), dptr(iso_mask), 1 if do_intersect else 0, _lib.ptr(aff), int(rho.size),
            _lib.ptr(rho), _lib.ptr(ang), _lib.ptr(wts), _lib.ptr(apply), _lib.ptr(hash_table),
            None if stream is None else ctypes.c_void_p(stream))
        del keep
        _lib.check(rc, "commit_hip_trk2dictionary")
        self._dim = dim
        self._n_count = n_count

    def sizes(self):
        v = [ctypes.c_int64() for _ in range(5)]
        _lib.check(self._lib.commit_hip_vox_sizes(self._h, *[ctypes.byref(x) for x in v]), "commit_hip_vox_sizes")
        return dict(zip(("n_ic", "n_kept", "n_ec", "n_iso", "n_pieces"), (x.value for x in v)))

    def fetch(self):
        s = self.sizes()
        nvox = int(np.prod(self._dim))
        out = {
            "IC_f": np.empty(s["n_ic"], np.uint32), "IC_v": np.empty(s["n_ic"], np.uint32),
            "IC_o": np.empty(s["n_ic"], np.uint16), "IC_len": np.empty(s["n_ic"], np.float32),
            "TRK_norm": np.empty(s["n_kept"], np.float32), "TRK_len": np.empty(s["n_kept"], np.float32),
            "TRK_lenTot": np.empty(s["n_kept"], np.float32), "TRK_kept": np.empty(self._n_count, np.uint8),
            "EC_v": np.empty(s["n_ec"], np.uint32), "EC_o": np.empty(s["n_ec"], np.uint16),
            "ISO_v": np.empty(s["n_iso"], np.uint32), "tdi": np.empty(nvox, np.float32),
        }
        keys = ("IC_f", "IC_v", "IC_o", "IC_len", "TRK_norm", "TRK_len", "TRK_lenTot", "TRK_kept", "EC_v", "EC_o",
                "ISO_v", "tdi")
        _lib.check(self._lib.commit_hip_vox_fetch(self._h, *[_lib.ptr(out[k]) for k in keys]), "commit_hip_vox_fetch")
        return out


def voxelize(fibers, fib_ptr, dim, pixdim, hash_table=None, ndirs=500, device=0, **kw):
    """Voxelise streamlines on the GPU; returns the dictionary arrays (the
    dictionary_*.dict contents) and 'tdi' (float32, x fastest)."""
    if hash_table is None:
        hash_table = default_hash_table(ndirs)
    v = Voxeliser(device)
    try:
        v.run(fibers, fib_ptr, dim, pixdim, hash_table, **kw)
        return v.fetch()
    finally:
        v.close()


# ------------------------------------------------------------ the flow ---
def _is_scalar(x):
    return np.isscalar(x)


def run(filename_tractogram=None, path_out=None, filename_peaks=None, filename_mask=None, filename_iso_mask=None,
        do_intersect=True, fiber_shift=0, min_seg_len=1e-3, min_fiber_len=0.0, max_fiber_len=250.0,
        vf_THR=0.1, peaks_use_affine=False, flip_peaks=(False, False, False), blur_clust_groupby=None,
        blur_clust_thr=0, blur_spacing=0.25, blur_core_extent=0.0, blur_gauss_extent=0.0,
        blur_gauss_min=0.1, blur_apply_to=None, TCK_ref_image=None, ndirs=500, n_threads=-1,
        keep_temp=False, verbose=3, hash_table=None, device=0):
    """trk2dictionary.run (trk2dictionary.pyx:57-654) with the voxelisation on
    the GPU.  Same arguments and files written; ``n_threads`` is recorded in
    dictionary_info.pickle only (the device partitions itself), ``keep_temp``
    is accepted (no temporary files are made)."""
    t0 = time.time()
    if _is_scalar(fiber_shift):
        shift = (fiber_shift, fiber_shift, fiber_shift)
    elif len(fiber_shift) == 3:
        shift = tuple(fiber_shift)
    else:
        raise ValueError('"fiber_shift" must be a scalar or a vector with 3 elements')
    rho, ang, wts, blur_sigma = blur_replicas(blur_core_extent, blur_gauss_extent, blur_spacing, blur_gauss_min)
    if min_seg_len < 0:
        raise ValueError('"min_seg_len" must be >= 0')
    if min_fiber_len < 0:
        raise ValueError('"min_fiber_len" must be >= 0')
    if max_fiber_len < min_fiber_len:
        raise ValueError('"max_fiber_len" must be >= "min_fiber_len"')
    if filename_tractogram is None:
        raise ValueError('"filename_tractogram" not defined')
    if np.isscalar(blur_clust_thr):
        blur_clust_thr = np.array([blur_clust_thr])
    if blur_clust_thr[0] > 0:
        raise NotImplementedError("streamline clustering (blur_clust_thr > 0) needs dicelib, which this build "
                                  "does not include")
    if path_out is None:
        path_out = dirname(filename_tractogram) or "."
        if not isdir(path_out):
            raise ValueError('"path_out" cannot be inferred from "filename_tractogram"')
        path_out = join(path_out, "COMMIT")
    os.makedirs(path_out, exist_ok=True)
    if n_threads == 0 or n_threads > 255:
        raise ValueError("Number of n_threads must be between 1 and 255")
    if n_threads == -1:
        n_threads = os.cpu_count() or 1
    if not exists(filename_tractogram):
        raise FileNotFoundError(f"Tractogram file not found: {filename_tractogram}")
    ext = splitext(filename_tractogram)[1]
    if ext not in (".trk", ".tck"):
        raise ValueError("Invalid input file: only .trk and .tck are supported")

    niiREF = niiMASK = niiPEAKS = None
    if ext == ".trk":
        fibers, fib_ptr, hdr = read_trk(filename_tractogram)
        Nx, Ny, Nz = (int(d) for d in hdr["dimensions"])
        Px, Py, Pz = (float(p) for p in hdr["voxel_sizes"])
        n_count = hdr["nb_streamlines"]
    else:
        if TCK_ref_image is None:
            if filename_peaks is not None:
                TCK_ref_image = filename_peaks
            elif filename_mask is not None:
                TCK_ref_image = filename_mask
            else:
                raise ValueError('TCK files do not contain information about the geometry. Use "TCK_ref_image" '
                                 'for that')
        niiREF = nifti.load(TCK_ref_image)
        Nx, Ny, Nz = niiREF.shape[:3]
        Px, Py, Pz = (float(np.float32(p)) for p in niiREF.pixdim[:3])
        toVOXMM = tck_to_voxmm(niiREF.affine, (Px, Py, Pz))
        fibers, fib_ptr, hdr = read_tck(filename_tractogram, toVOXMM)
        n_count = fib_ptr.size - 1
    if Nx >= 2 ** 16 or Ny >= 2 ** 16 or Nz >= 2 ** 16:
        raise ValueError("The max dim size is 2^16 voxels")
    if blur_apply_to is None:
        blur_apply_to = np.repeat([True], n_count)
    elif np.asarray(blur_apply_to).size != n_count:
        raise ValueError('"blur_apply_to" must have one value per streamline')

    mask = None
    if filename_mask is not None:
        niiMASK = nifti.load(filename_mask)
        mask = np.ascontiguousarray(np.asarray(niiMASK.data).astype(np.float32))
    peaks = None
    peaks_aff = None
    if filename_peaks is not None:
        niiPEAKS = nifti.load(filename_peaks)
        if niiPEAKS.shape[3] % 3:
            raise ValueError("PEAKS dataset must have 3*k volumes")
        if vf_THR < 0 or vf_THR > 1:
            raise ValueError('"vf_THR" must be between 0 and 1')
        peaks = np.ascontiguousarray(np.asarray(niiPEAKS.data).astype(np.float32))
        peaks_aff = np.ascontiguousarray(niiPEAKS.affine[:3, :3].T) if peaks_use_affine else np.eye(3)
    iso = None
    if filename_iso_mask is not None:
        iso = np.ascontiguousarray(np.asarray(nifti.load(filename_iso_mask).data).astype(np.float32))

    if hash_table is None:
        hash_table = default_hash_table(ndirs)
    d = voxelize(fibers, fib_ptr, (Nx, Ny, Nz), (Px, Py, Pz), hash_table=hash_table, device=device,
                 fiber_shift=shift, min_seg_len=min_seg_len, min_fiber_len=min_fiber_len,
                 max_fiber_len=max_fiber_len, do_intersect=do_intersect, mask=mask, iso_mask=iso, peaks=peaks,
                 vf_THR=vf_THR, flip_peaks=flip_peaks, peaks_affine=peaks_aff, blur=(rho, ang, wts),
                 blur_apply_to=np.asarray(blur_apply_to, bool))

    info = {
        "filename_tractogram": filename_tractogram, "iso_mask": filename_iso_mask is not None,
        "TCK_ref_image": TCK_ref_image, "path_out": path_out, "filename_peaks": filename_peaks,
        "filename_mask": filename_mask, "do_intersect": do_intersect, "fiber_shift": fiber_shift,
        "min_seg_len": min_seg_len, "min_fiber_len": min_fiber_len, "max_fiber_len": max_fiber_len,
        "vf_THR": vf_THR, "peaks_use_affine": peaks_use_affine, "flip_peaks": list(flip_peaks),
        "blur_core_extent": blur_core_extent, "blur_gauss_extent": blur_gauss_extent,
        "blur_gauss_min": blur_gauss_min, "blur_spacing": blur_spacing, "blur_sigma": blur_sigma,
        "blur_apply_to": np.asarray(blur_apply_to), "ndirs": ndirs, "n_threads": n_threads,
    }
    with open(join(path_out, "dictionary_info.pickle"), "wb+") as f:
        pickle.dump(info, f, protocol=2)
    for key, dt in (("IC_f", "<u4"), ("IC_v", "<u4"), ("IC_o", "<u2"), ("IC_len", "<f4"), ("TRK_norm", "<f4"),
                    ("TRK_len", "<f4"), ("TRK_lenTot", "<f4"), ("TRK_kept", "u1"), ("EC_v", "<u4"),
                    ("EC_o", "<u2"), ("ISO_v", "<u4")):
        d[key].astype(dt).tofile(join(path_out, f"dictionary_{key}.dict"))
    # TDI and mask maps (trk2dictionary.pyx:603-650)
    if niiREF is not None:
        aff = niiREF.affine
    elif niiMASK is not None:
        aff = niiMASK.affine
    elif niiPEAKS is not None:
        aff = niiPEAKS.affine
    else:
        aff = np.diag([Px, Py, Pz, 1])
    tdi = np.reshape(d["tdi"], (Nx, Ny, Nz), order="F")
    nifti.save(nifti.Nifti1Image(tdi, aff, (Px, Py, Pz)), join(path_out, "dictionary_tdi.nii.gz"))
    mimg = mask if mask is not None else (tdi > 0).astype(np.float32)
    nifti.save(nifti.Nifti1Image(mimg, aff, (Px, Py, Pz)), join(path_out, "dictionary_mask.nii.gz"))
    tmp = join(path_out, "temp")
    if exists(tmp) and not keep_temp:
        shutil.rmtree(tmp, ignore_errors=True)
    d["time"] = time.time() - t0
    return d
