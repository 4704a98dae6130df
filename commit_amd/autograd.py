"""Differentiable COMMIT operator: y = A(LUT) x as a torch.autograd.Function.

Backward with respect to x is the HIP adjoint (commit_hip_At); backward with
respect to the response-function tables is the HIP scatter-reduction
commit_hip_lut_grad.  With the tables built from the compartment
diffusivities by differentiable torch code (``szb_tables``), the
diffusivities can be fitted by gradient (``fit_diffusivities``).

No counterpart in the reference: its tables are fixed inputs computed once by
AMICO (core.pyx:355-356, 380-382) and the operator has no derivative.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _upload(op, wmr, wmh, iso):
    lib = _lib.load()
    _lib.check(lib.commit_hip_plan_set_lut(op._plan.handle, _ptr(wmr), _ptr(wmh), _ptr(iso), _stream(op.device)),
               "commit_hip_plan_set_lut")


class OperatorFunction(torch.autograd.Function):
    """y = A x with the plan's tables replaced by (wmr, wmh, iso).

    x: device vector of the operator's precision; wmr [nIC,ndirs,nS],
    wmh [nEC,ndirs,nS], iso [nISO,nS]: float32 device tensors (or None to keep
    the plan's table).  Gradients: x -> A^T gy; tables -> commit_hip_lut_grad.
    The tables are uploaded into the operator's device plan, which every
    operator sharing it uses: they stay there after the call (a later
    A.dot runs with them) until other tables are uploaded."""

    @staticmethod
    def forward(ctx, x, wmr, wmh, iso, op):
        tabs = [None if t is None else t.detach().to(torch.float32).contiguous() for t in (wmr, wmh, iso)]
        _upload(op, *tabs)
        xd = x.detach().contiguous()
        y = op.matvec(xd)
        ctx.op = op
        ctx.tabs = tabs
        ctx.save_for_backward(xd)
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        op = ctx.op
        gy = gy.contiguous().to(op.dtype)
        _upload(op, *ctx.tabs)             # the tables of this forward
        gx = op.T.matvec(gy) if ctx.needs_input_grad[0] else None
        want = [ctx.needs_input_grad[k] and ctx.tabs[k - 1] is not None for k in (1, 2, 3)]
        grads = [None, None, None]
        if any(want):
            out = [torch.empty(t.shape, dtype=torch.float64, device=x.device) if w else None
                   for t, w in zip(ctx.tabs, want)]
            lib = _lib.load()
            _lib.check(lib.commit_hip_lut_grad(op._plan.handle, _ptr(x), _ptr(gy), _ptr(out[0]), _ptr(out[1]),
                                               _ptr(out[2]), _stream(x.device)), "commit_hip_lut_grad")
            grads = [None if o is None else o.to(torch.float32) for o in out]
        return gx, grads[0], grads[1], grads[2], None


def apply(op, x, wmr=None, wmh=None, iso=None):
    """Differentiable y = A(wmr, wmh, iso) x on the device plan of ``op``."""
    return OperatorFunction.apply(x, wmr, wmh, iso, op)


def szb_tables(d_par, d_perp, bvals, bvecs, dirs, d_isos=(1.7e-3, 3.0e-3), normalize=True):
    """Stick-Zeppelin-Ball tables as differentiable torch functions of the
    diffusivities (the closed forms of the synthetic model, with the
    reference's kernel normalisation: each RF divided by the norm of its
    orientation-0 row, core.pyx:412-431).  d_par, d_perp: 0-d tensors."""
    dev, dt = d_par.device, d_par.dtype
    g = torch.as_tensor(np.asarray(bvecs), dtype=dt, device=dev)
    u = torch.as_tensor(np.asarray(dirs), dtype=dt, device=dev)
    b = torch.as_tensor(np.asarray(bvals), dtype=dt, device=dev)[None, :]
    c2 = (u @ g.T) ** 2
    wmr = torch.exp(-b * d_par * c2)[None]
    wmh = torch.exp(-b * (d_perp + (d_par - d_perp) * c2))[None]
    iso = torch.stack([torch.exp(-b[0] * d) for d in d_isos]) if len(d_isos) else torch.zeros((0, b.shape[1]))
    if normalize:
        wmr = wmr / torch.linalg.vector_norm(wmr[:, 0, :], dim=-1)[:, None, None]
        wmh = wmh / torch.linalg.vector_norm(wmh[:, 0, :], dim=-1)[:, None, None]
        iso = iso / torch.linalg.vector_norm(iso, dim=-1)[:, None]
    return wmr, wmh, iso


def fit_diffusivities(op, x, y, K, theta0, iters=30, d_isos=(1.7e-3, 3.0e-3), history=None):
    """Fit (d_par, d_perp) by L-BFGS on 0.5 ||A(theta) x - y||^2 with x fixed.

    op: LinearOperator (its precision sets the vector dtype); x, y: device
    vectors; K: kernels dict holding 'bvals', 'bvecs' and the directions under
    'dirs'; theta0: starting (d_par, d_perp).  Returns the fitted pair.
    On return the device plan holds the tables of the fitted pair (the last
    L-BFGS evaluation need not be at it), and op.KERNELS['wmr' / 'wmh' /
    'iso'] are replaced by those float32 tables."""
    dev = op.device
    scale = 1e-3
    p = torch.tensor([theta0[0] / scale, theta0[1] / scale], dtype=torch.float64, device=dev, requires_grad=True)
    opt = torch.optim.LBFGS([p], lr=1.0, max_iter=iters, tolerance_grad=1e-12, tolerance_change=1e-14,
                            line_search_fn="strong_wolfe")
    xd = x.to(op.dtype).contiguous()
    yd = y.to(op.dtype)

    def closure():
        opt.zero_grad()
        wmr, wmh, iso = szb_tables(p[0] * scale, p[1] * scale, K["bvals"], K["bvecs"], K["dirs"], d_isos)
        r = apply(op, xd, wmr.float(), wmh.float() if K["wmh"].shape[0] else None, iso.float()) - yd
        loss = 0.5 * torch.dot(r.double(), r.double())
        loss.backward()
        if history is not None:
            history.append((float(loss.item()), float(p[0].item()) * scale, float(p[1].item()) * scale))
        return loss

    opt.step(closure)
    dp, dq = float(p[0].item()) * scale, float(p[1].item()) * scale
    with torch.no_grad():
        tt = lambda v: torch.tensor(v, dtype=torch.float64, device=dev)
        wmr, wmh, iso = (t.float().contiguous() for t in
                         szb_tables(tt(dp), tt(dq), K["bvals"], K["bvecs"], K["dirs"], d_isos))
        _upload(op, wmr, wmh if K["wmh"].shape[0] else None, iso)
    if isinstance(op.KERNELS, dict):
        op.KERNELS["wmr"] = wmr.cpu().numpy()
        if K["wmh"].shape[0]:
            op.KERNELS["wmh"] = wmh.cpu().numpy()
        op.KERNELS["iso"] = iso.cpu().numpy()
    return dp, dq
