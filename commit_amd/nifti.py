"""Minimal NIfTI-1 single-file (.nii / .nii.gz) reader and writer.

Covers what the COMMIT flow exchanges with disk: 3D/4D volumes of
uint8/int16/int32/float32/float64, the voxel sizes (pixdim) and the sform
affine.  Data are stored in Fortran (x fastest) order as the format requires.
"""
import gzip
import struct

import numpy as np

_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8,
           512: np.uint16, 768: np.uint32}
_CODES = {np.dtype(v): k for k, v in _DTYPES.items()}


class Nifti1Image:
    def __init__(self, data, affine=None, pixdim=None, descrip=""):
        self.data = np.asarray(data)
        self.affine = np.eye(4) if affine is None else np.asarray(affine, dtype=np.float64)
        if pixdim is None:
            pixdim = tuple(float(np.linalg.norm(self.affine[:3, i])) for i in range(3))
        self.pixdim = tuple(pixdim)
        self.descrip = descrip

    @property
    def shape(self):
        return self.data.shape


def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def load(path):
    with _open(path, "rb") as f:
        raw = f.read()
    endian = "<"
    if struct.unpack("<i", raw[:4])[0] != 348:
        endian = ">"
        if struct.unpack(">i", raw[:4])[0] != 348:
            raise ValueError(f"{path}: not a NIfTI-1 file")
    dim = struct.unpack(endian + "8h", raw[40:56])
    datatype = struct.unpack(endian + "h", raw[70:72])[0]
    pixdim = struct.unpack(endian + "8f", raw[76:108])
    vox_offset = int(struct.unpack(endian + "f", raw[108:112])[0])
    scl_slope, scl_inter = struct.unpack(endian + "2f", raw[112:120])
    sform_code = struct.unpack(endian + "h", raw[254:256])[0]
    srow = np.array(struct.unpack(endian + "12f", raw[280:328])).reshape(3, 4)
    descrip = raw[148:228].split(b"\0")[0].decode(errors="replace")
    if datatype not in _DTYPES:
        raise ValueError(f"{path}: unsupported NIfTI datatype {datatype}")
    ndim = dim[0]
    shape = tuple(int(d) for d in dim[1:1 + ndim])
    dt = np.dtype(_DTYPES[datatype]).newbyteorder(endian)
    count = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dt, count=count, offset=vox_offset).reshape(shape, order="F")
    data = data.astype(dt.newbyteorder("="))
    if scl_slope not in (0.0, 1.0) or scl_inter != 0.0:
        data = data * scl_slope + scl_inter
    affine = np.eye(4)
    if sform_code > 0:
        affine[:3] = srow
    else:
        affine[:3, :3] = np.diag(pixdim[1:4])
    return Nifti1Image(data, affine, tuple(float(p) for p in pixdim[1:4]), descrip)


def save(img, path):
    data = np.asarray(img.data)
    if data.dtype == np.bool_:
        data = data.astype(np.uint8)
    if np.dtype(data.dtype) not in _CODES:
        data = data.astype(np.float32)
    code = _CODES[np.dtype(data.dtype)]
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    dims = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<h", hdr, 70, code)
    struct.pack_into("<h", hdr, 72, data.dtype.itemsize * 8)
    pix = [1.0] + list(img.pixdim[:3]) + [1.0] * 4
    struct.pack_into("<8f", hdr, 76, *pix)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<f", hdr, 112, 1.0)
    hdr[148:148 + min(79, len(img.descrip))] = img.descrip.encode()[:79]
    struct.pack_into("<h", hdr, 252, 1)      # qform_code
    struct.pack_into("<h", hdr, 254, 1)      # sform_code
    struct.pack_into("<12f", hdr, 280, *img.affine[:3].ravel())
    hdr[344:348] = b"n+1\0"
    with _open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(np.asfortranarray(data).tobytes(order="F"))
