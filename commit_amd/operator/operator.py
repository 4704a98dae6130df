"""Drop-in LinearOperator backed by libcommit_hip.so.

Mirrors commit.operator.operator.LinearOperator (operator.pyx:10-184): the same
constructor arguments, ``.T``, ``.shape``, ``.dot``, ``nROWS``, ``nCOLS`` and
``adjoint``, the same size check and exception.  The dictionary is uploaded to
HBM once, when the operator is built; ``.T`` shares that device plan.

``dot`` accepts
  * a float64 numpy vector (the reference's interface): the vector is copied to
    the GPU, the product runs there, and a numpy vector comes back;
  * a torch tensor on the GPU of the plan's precision: the product runs on the
    current torch stream and a device tensor is returned (used by the device
    FISTA and the autograd Function; no host round trip).
"""
import ctypes

import numpy as np
import torch

from .. import _lib


class _Plan:
    """Owns one device-resident dictionary (commit_hip_plan)."""

    def __init__(self, DICTIONARY, KERNELS, nolut, precision, device):
        lib = _lib.load()
        IC, EC, ISO = DICTIONARY["IC"], DICTIONARY["EC"], DICTIONARY["ISO"]
        self.nIC = int(KERNELS["wmr"].shape[0])
        self.nEC = int(KERNELS["wmh"].shape[0])
        self.nISO = int(KERNELS["iso"].shape[0])
        # number of samples, as operator.pyx:55-61
        if KERNELS["wmr"].size > 0:
            self.nS = int(KERNELS["wmr"].shape[2])
        elif KERNELS["wmh"].size > 0:
            self.nS = int(KERNELS["wmh"].shape[2])
        else:
            self.nS = int(KERNELS["wmr"].shape[1])
        self.ndirs = int(KERNELS["wmr"].shape[1])
        self.nF = int(IC["nSTR"])
        self.n = int(IC["n"])
        self.nE = int(EC["n"])
        self.nV = int(ISO["n"])
        self.nVOX = int(IC["nVOX"])
        self.precision = precision
        self.device = device
        self.dtype = torch.float64 if precision == _lib.FP64 else torch.float32
        u32 = lambda a: np.ascontiguousarray(a, dtype=np.uint32)
        u16 = lambda a: np.ascontiguousarray(a, dtype=np.uint16)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        ev = IC.get("eval")
        if ev is None:
            ev = np.ones(self.nF * self.nIC, dtype=np.uint32)
        self._host = dict(ICf=u32(IC["str"]), ICv=u32(IC["vox"]), ICo=u16(IC["dir"]), ICl=f32(IC["len"]),
                          ICeval=u32(ev), ECv=u32(EC["vox"]), ECo=u16(EC["dir"]), ISOv=u32(ISO["vox"]),
                          wmr=f32(KERNELS["wmr"]), wmh=f32(KERNELS["wmh"]), iso=f32(KERNELS["iso"]))
        h = self._host
        if h["ICeval"].size != self.nF * self.nIC:
            raise RuntimeError("DICTIONARY['IC']['eval'] must hold nSTR*nIC entries")
        handle = ctypes.c_void_p()
        P = _lib.ptr
        rc = lib.commit_hip_plan_create(
            ctypes.byref(handle), self.nS, self.ndirs, self.nF, self.n, self.nE, self.nV, self.nVOX,
            P(h["ICf"]), P(h["ICeval"]), P(h["ICv"]), P(h["ICo"]), P(h["ICl"]), P(h["ECv"]), P(h["ECo"]),
            P(h["ISOv"]), P(h["wmr"]), P(h["wmh"]), P(h["iso"]), self.nIC, self.nEC, self.nISO,
            1 if nolut else 0, precision, device)
        _lib.check(rc, "commit_hip_plan_create")
        self.handle = handle
        nr, nc = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.commit_hip_plan_shape(handle, ctypes.byref(nr), ctypes.byref(nc)), "commit_hip_plan_shape")
        self.nROWS, self.nCOLS = int(nr.value), int(nc.value)
        self.device_bytes = int(lib.commit_hip_plan_device_bytes(handle))
        self._lib = lib

    def set_precision(self, precision):
        _lib.check(self._lib.commit_hip_plan_set_precision(self.handle, int(precision)), "commit_hip_plan_set_precision")
        self.precision = int(precision)
        self.dtype = torch.float64 if self.precision == _lib.FP64 else torch.float32

    def set_eval(self, ICeval):
        ev = np.ascontiguousarray(ICeval, dtype=np.uint32)
        if ev.size != self.nF * self.nIC:
            raise RuntimeError("eval mask must hold nSTR*nIC entries")
        self._host["ICeval"] = ev
        _lib.check(self._lib.commit_hip_plan_set_eval(self.handle, _lib.ptr(ev)), "commit_hip_plan_set_eval")

    def run(self, adjoint, v_in, v_out, stream):
        fn = self._lib.commit_hip_At if adjoint else self._lib.commit_hip_A
        rc = fn(self.handle, ctypes.c_void_p(v_in.data_ptr()), ctypes.c_void_p(v_out.data_ptr()),
                ctypes.c_void_p(stream))
        _lib.check(rc, "commit_hip_At" if adjoint else "commit_hip_A")

    def run_multi(self, adjoint, v_in, v_out, nrhs, sqrtw, stream):
        lib = self._lib
        if adjoint:
            rc = lib.commit_hip_At_multi(self.handle, ctypes.c_void_p(v_in.data_ptr()),
                                         ctypes.c_void_p(sqrtw.data_ptr() if sqrtw is not None else None),
                                         ctypes.c_void_p(v_out.data_ptr()), nrhs, _lib.FP32, ctypes.c_void_p(stream))
        else:
            rc = lib.commit_hip_A_multi(self.handle, ctypes.c_void_p(v_in.data_ptr()),
                                        ctypes.c_void_p(v_out.data_ptr()), nrhs, _lib.FP32, ctypes.c_void_p(stream))
        _lib.check(rc, "commit_hip_At_multi" if adjoint else "commit_hip_A_multi")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.commit_hip_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


class LinearOperator:
    """Wrapper of the MI355X implementation of the COMMIT linear operator A."""

    def __init__(self, DICTIONARY, KERNELS, THREADS, nolut=False, precision=_lib.FP64, device=None, _plan=None):
        self.DICTIONARY = DICTIONARY
        self.KERNELS = KERNELS
        self.THREADS = THREADS     # accepted for interface parity; the GPU plan partitions itself
        self.nolut = bool(nolut)
        if device is None:
            device = torch.cuda.current_device() if torch.cuda.is_available() else 0
        self._plan = _plan if _plan is not None else _Plan(DICTIONARY, KERNELS, self.nolut, precision, device)
        self.nSAMPLES = self._plan.nS
        self.ndirs = self._plan.ndirs
        self.adjoint = 0
        self.nROWS = self._plan.nROWS
        self.nCOLS = self._plan.nCOLS

    @property
    def T(self):
        """Transpose of the explicit matrix (shares the device plan)."""
        C = LinearOperator(self.DICTIONARY, self.KERNELS, self.THREADS, self.nolut, _plan=self._plan)
        C.adjoint = 1 - self.adjoint
        return C

    @property
    def shape(self):
        return (self.nROWS, self.nCOLS) if not self.adjoint else (self.nCOLS, self.nROWS)

    @property
    def dtype(self):
        return self._plan.dtype

    @property
    def device(self):
        return torch.device("cuda", self._plan.device)

    @property
    def precision(self):
        """Storage precision of the vectors (commit_amd._lib.FP64 or FP32)."""
        return self._plan.precision

    def set_precision(self, precision):
        """Switch the vector storage type of the shared device plan (FP64: the
        reference's float64 vectors; FP32: float32 vectors, FP64 accumulation)."""
        self._plan.set_precision(precision)

    def set_eval(self, ICeval):
        """Update DICTIONARY['IC']['eval'] on the device (debias, core.pyx:1482-1487)."""
        self._plan.set_eval(ICeval)
        self.DICTIONARY["IC"]["eval"] = np.ascontiguousarray(ICeval, dtype=np.uint32)

    def _check_device(self, t, what):
        if t.device.index != self._plan.device:
            raise RuntimeError(f"{what}: tensor on {t.device}, the plan lives on cuda:{self._plan.device}")

    def matvec(self, v_in, out=None):
        """Device product: v_in is a contiguous GPU tensor of the plan's dtype."""
        if v_in.numel() != self.shape[1]:
            raise RuntimeError("A.dot(): dimensions do not match")
        if v_in.dtype != self._plan.dtype or not v_in.is_cuda or not v_in.is_contiguous():
            raise RuntimeError(f"A.dot(): expected a contiguous CUDA tensor of {self._plan.dtype}")
        self._check_device(v_in, "A.dot()")
        if out is None:
            out = torch.empty(self.shape[0], dtype=self._plan.dtype, device=v_in.device)
        elif out.numel() != self.shape[0] or out.dtype != self._plan.dtype or not out.is_contiguous():
            raise RuntimeError(f"A.dot(): out must be a contiguous {self._plan.dtype} tensor of {self.shape[0]} entries")
        else:
            self._check_device(out, "A.dot()")
        stream = torch.cuda.current_stream(v_in.device).cuda_stream
        self._plan.run(self.adjoint, v_in, out, stream)
        return out

    def matmat(self, V, out=None, sqrtw=None):
        """Multi-RHS device product on the matrix cores: V is a contiguous
        float32 CUDA tensor of shape [shape[1], R] (R right-hand sides, the RHS
        index fastest); returns [shape[0], R] float32.  For the adjoint,
        ``sqrtw`` (float32 [nROWS]) weights the rows of V first.  No
        counterpart in the reference, which calls dot() once per vector."""
        if V.dim() != 2 or V.shape[0] != self.shape[1]:
            raise RuntimeError("A.matmat(): dimensions do not match")
        if V.dtype != torch.float32 or not V.is_cuda or not V.is_contiguous():
            raise RuntimeError("A.matmat(): expected a contiguous CUDA float32 tensor")
        self._check_device(V, "A.matmat()")
        R = int(V.shape[1])
        if out is None:
            out = torch.empty((self.shape[0], R), dtype=torch.float32, device=V.device)
        elif out.shape != (self.shape[0], R) or out.dtype != torch.float32 or not out.is_contiguous():
            raise RuntimeError("A.matmat(): bad output tensor")
        else:
            self._check_device(out, "A.matmat()")
        if sqrtw is not None:
            if not self.adjoint:
                raise RuntimeError("A.matmat(): row weights apply to the adjoint only")
            if sqrtw.numel() != self.nROWS or sqrtw.dtype != torch.float32 or not sqrtw.is_contiguous():
                raise RuntimeError("A.matmat(): sqrtw must be contiguous float32 [nROWS]")
            self._check_device(sqrtw, "A.matmat()")
        stream = torch.cuda.current_stream(V.device).cuda_stream
        self._plan.run_multi(self.adjoint, V, out, R, sqrtw, stream)
        return out

    def dot(self, v_in):
        """Matrix-vector product with A (or A^T after ``.T``).

        numpy float64 in -> numpy float64 out (operator.pyx:114-184);
        CUDA tensor in -> CUDA tensor out."""
        if isinstance(v_in, torch.Tensor) and v_in.is_cuda:
            return self.matvec(v_in)
        v = np.asarray(v_in)
        if v.ndim != 1 or v.size != self.shape[1]:
            raise RuntimeError("A.dot(): dimensions do not match")
        if v.dtype != np.float64:
            raise ValueError("Buffer dtype mismatch, expected 'double'")
        xd = torch.from_numpy(np.ascontiguousarray(v)).to(self.device, dtype=self._plan.dtype)
        out = self.matvec(xd)
        return out.to("cpu", dtype=torch.float64).numpy()
