"""Regularised least squares on the GPU: FISTA with backtracking and the
COMMIT proximal operators, behind the interface of commit.solvers
(solvers.py:16-391).

``init_regularisation`` returns the same dict as the reference, with
``omega`` and ``prox`` callables (numpy arrays or CUDA tensors), plus a
device description of the penalty used by the GPU FISTA.  ``solve`` /
``fista`` keep the reference's signatures, control flow, step-size rule,
backtracking and stopping tests; with a device operator every vector stays
in HBM and each step runs as a handful of HIP kernels (fused residual,
adjoint, proximal step, one deterministic multi-sum), reading back only the
scalars the control flow needs.
"""
import ctypes
import sys
from math import sqrt

import numpy as np
import torch

from . import _lib

eps = np.finfo(float).eps
list_regularizers = [None, "lasso", "group_lasso", "sparse_group_lasso"]
_REG = {None: _lib.REG_NONE, "lasso": _lib.REG_LASSO, "group_lasso": _lib.REG_GROUP_LASSO,
        "sparse_group_lasso": _lib.REG_SPARSE_GROUP_LASSO}


def _slice_range(a, b, n):
    """[a, b) of the Python slice x[a:b] on a length-n vector (a, b >= 0)."""
    a, b = min(int(a), n), min(int(b), n)
    return (a, b) if b > a else (0, 0)


class DeviceRegularisation:
    """Device image of one regularisation: the commit_hip_prox_params struct,
    its weight and group arrays in HBM, and the recipe that turns the
    deterministic sums of commit_hip_fista_sums into Omega(x)."""

    def __init__(self, rp, n, device):
        self.n = int(n)
        self.device = device
        p = _lib.ProxParams()
        starts = (rp.get("startIC", 0), rp.get("startEC", 0), rp.get("startISO", 0))
        sizes = (rp.get("sizeIC", 0), rp.get("sizeEC", 0), rp.get("sizeISO", 0))
        regs = (rp["regIC"], rp["regEC"], rp["regISO"])
        nns = (rp.get("nnIC"), rp.get("nnEC"), rp.get("nnISO"))
        for c in range(3):
            if regs[c] not in list_regularizers:
                raise ValueError("Regularisation not in the list")
            if c > 0 and regs[c] in ("group_lasso", "sparse_group_lasso"):
                raise ValueError("group penalties are available only for the IC compartment")
            p.start[c], p.size[c] = int(starts[c]), int(sizes[c])
            p.reg[c] = _REG[regs[c]]
            p.nn[c] = 1 if nns[c] else 0
        # per-coefficient weights (solvers.py:30-42)
        w_all = np.ones(self.n, dtype=np.float64)
        dIC, dISO = rp.get("dictIC_params"), rp.get("dictISO_params")
        wic = dIC is not None and "coeff_weights_kept" in dIC
        wiso = dISO is not None and "coeff_weights" in dISO
        if wic:
            w_all[starts[0]:starts[0] + sizes[0]] = dIC["coeff_weights_kept"]
        if wiso:
            w_all[starts[2]:starts[2] + sizes[2]] = dISO["coeff_weights"]
        self.w = torch.from_numpy(w_all).to(device)
        p.w = self.w.data_ptr()
        lam = [0.0, 0.0, 0.0]
        self.omega_terms = []            # (kind, factor) composing Omega from the sums
        l1 = [(0, 0, 0)] * 3
        reg = regs[0]
        if reg == "lasso":
            lam[0] = float(rp["lambdaIC"])
            p.weighted[0] = 1 if wic else 0
            # solvers.py:57/63 slice x[startIC:sizeIC]
            a, b = _slice_range(starts[0], sizes[0], self.n)
            l1[0] = (a, b, 1 if wic else 0)
            self.omega_terms.append(("l1", 0, lam[0]))
        elif reg in ("group_lasso", "sparse_group_lasso"):
            groups = dIC["group_idx_kept"]
            gw = np.ascontiguousarray(dIC["group_weights"], dtype=np.float64)
            if len(groups) != gw.size:
                raise ValueError("Number of groups and weights do not match")
            sizes_g = np.array([np.asarray(g).size for g in groups], dtype=np.int64)
            gptr = np.zeros(len(groups) + 1, dtype=np.int32)
            gptr[1:] = np.cumsum(sizes_g)
            gidx = (np.concatenate([np.asarray(g, dtype=np.int32) for g in groups])
                    if len(groups) else np.zeros(0, np.int32))
            if gidx.size and (gidx.min() < 0 or gidx.max() >= self.n):
                raise ValueError("group index outside x")
            self.overlap = int(np.unique(gidx).size != gidx.size)
            self.gptr = torch.from_numpy(gptr).to(device)
            self.gidx = torch.from_numpy(np.ascontiguousarray(gidx)).to(device)
            self.gw = torch.from_numpy(gw).to(device)
            p.ngroups = len(groups)
            p.group_ptr, p.group_idx, p.group_w = self.gptr.data_ptr(), self.gidx.data_ptr(), self.gw.data_ptr()
            p.groups_overlap = self.overlap
            if reg == "group_lasso":
                p.lam_group = float(rp["lambdaIC"])
                self.omega_terms.append(("group", None, p.lam_group))
            else:
                lam[0], p.lam_group = float(rp["lambdaIC"][0]), float(rp["lambdaIC"][1])
                p.weighted[0] = 1 if wic else 0
                # the l1 part sums over the whole x (proximals.pyx:150, 175-176)
                l1[0] = (0, self.n, 1 if wic else 0)
                self.omega_terms.append(("sgl", lam[0], p.lam_group))
        if regs[1] == "lasso":
            lam[1] = float(rp["lambdaEC"])
            a, b = _slice_range(starts[1], starts[1] + sizes[1], self.n)
            l1[1] = (a, b, 0)
            self.omega_terms.append(("l1", 1, lam[1]))
        if regs[2] == "lasso":
            lam[2] = float(rp["lambdaISO"])
            p.weighted[2] = 1 if wiso else 0
            if wiso:   # solvers.py:183 slices x[startISO:sizeISO]
                a, b = _slice_range(starts[2], sizes[2], self.n)
            else:
                a, b = _slice_range(starts[2], starts[2] + sizes[2], self.n)
            l1[2] = (a, b, 1 if wiso else 0)
            self.omega_terms.append(("l1", 2, lam[2]))
        for c in range(3):
            p.lam[c] = lam[c]
        self.params = p
        self.l1_range = (ctypes.c_int64 * 6)(*[v for t in l1 for v in t[:2]])
        self.l1_weighted = (ctypes.c_int * 3)(*[t[2] for t in l1])
        self.ngroups = int(p.ngroups)
        self.work = torch.empty(_lib.sums_work(self.ngroups), dtype=torch.float64, device=device)
        self.out = torch.empty(8, dtype=torch.float64, device=device)

    def sums(self, x, xhat=None, grad=None, prev=None, out=None):
        """The eight deterministic sums of commit_hip_fista_sums (FP64
        accumulation for float64 and float32 vectors alike), on the host; with
        out (a float64 device tensor of 8), written there without a host sync."""
        lib = _lib.load()
        P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        p = self.params
        fn = lib.commit_hip_fista_sums_f32 if x.dtype == torch.float32 else lib.commit_hip_fista_sums
        rc = fn(self.n, P(x), P(xhat), P(grad), P(prev), self.l1_range, self.l1_weighted,
                                       ctypes.c_void_p(self.w.data_ptr()), self.ngroups,
                                       ctypes.c_void_p(p.group_ptr or 0), ctypes.c_void_p(p.group_idx or 0),
                                       ctypes.c_void_p(p.group_w or 0), P(self.out if out is None else out),
                                       P(self.work), _stream(x))
        _lib.check(rc, "commit_hip_fista_sums")
        return self.out.cpu().numpy() if out is None else None

    def omega_from(self, s):
        """Omega(x) from the sums (solvers.py:200 and the per-compartment omegas)."""
        om = 0.0
        for kind, a, b in self.omega_terms:
            if kind == "l1":
                om += b * s[4 + a]
            elif kind == "group":
                om += b * s[7] if b != 0 else 0.0
            elif kind == "sgl":
                if a != 0 or b != 0:
                    om += a * s[4] + b * s[7]
        return om

    def prox(self, x, scale, xin=None, grad=None, step=0.0):
        lib = _lib.load()
        P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        src = x if xin is None else xin
        fn = lib.commit_hip_prox_f32 if x.dtype == torch.float32 else lib.commit_hip_prox
        rc = fn(ctypes.byref(self.params), self.n, P(src), P(grad), float(step), float(scale),
                                 P(x), _stream(x))
        _lib.check(rc, "commit_hip_prox")
        return x


def _stream(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _device_reg(rp, n, device):
    return DeviceRegularisation(rp, n, device)


def init_regularisation(regularisation_params, n=None, device=None):
    """solvers.init_regularisation (solvers.py:16-206): adds 'omega' and 'prox'.

    The callables accept numpy vectors (as the reference's do: prox works in
    place and returns the vector) or CUDA tensors.  n defaults to
    sizeIC + sizeEC + sizeISO."""
    rp = regularisation_params
    if rp["regIC"] not in list_regularizers or rp["regEC"] not in list_regularizers or \
            rp["regISO"] not in list_regularizers:
        raise ValueError("Regularisation not in the list")
    if n is None:
        n = int(rp.get("sizeIC", 0)) + int(rp.get("sizeEC", 0)) + int(rp.get("sizeISO", 0))
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    dev = _device_reg(rp, n, device)

    def _as_dev(x):
        if isinstance(x, torch.Tensor) and x.is_cuda:
            return x, False
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device), True

    def omega(x):
        xd, _ = _as_dev(x)
        return dev.omega_from(dev.sums(xd))

    def prox(x, scaling):
        xd, host = _as_dev(x)
        dev.prox(xd, scaling)
        if host:
            x[...] = xd.cpu().numpy()
        return x

    rp["omega"] = omega
    rp["prox"] = prox
    rp["_device"] = dev
    return rp


def evaluate_model(y, A, x, regularisation=None):
    """0.5 ||Ax - y||^2 + Omega(x) (what solvers.py:209-215 intends; the
    reference unpacks the dict by mistake and cannot run)."""
    omega = (lambda v: 0.0) if regularisation is None else regularisation["omega"]
    r = np.asarray(A.dot(np.ascontiguousarray(x, dtype=np.float64))) - y
    return 0.5 * np.linalg.norm(r) ** 2 + omega(x)


def solve(y, A, At, tol_fun=1e-4, tol_x=1e-6, max_iter=1000, verbose=True, x0=None, regularisation=None,
          confidence_array=None, lipschitz_iters=0):
    """solvers.solve (solvers.py:218-241).

    The vectors are stored in the operator's precision (A.precision): float64
    reproduces the reference; float32 keeps x, y and the FISTA vectors in
    float32 while the stopping norms, the objective, the Lipschitz estimate
    and the final residual stay FP64.  lipschitz_iters > 0 replaces the
    reference's one-step estimate (solvers.py:282-286) by that many FP64
    power iterations on A^T W A."""
    if regularisation is None:
        omega = lambda x: 0.0
        prox = lambda x, _: x
    else:
        omega, prox = regularisation["omega"], regularisation["prox"]
    if x0 is None:
        x0 = np.zeros(A.shape[1])
    if confidence_array is not None:
        confidence_array = np.sqrt(confidence_array)
    return fista(y, A, At, omega, prox, confidence_array, tol_fun, tol_x, max_iter, verbose, x0,
                 _dev=None if regularisation is None else regularisation.get("_device"),
                 lipschitz_iters=lipschitz_iters)


def fista(y, A, At, omega, prox, sqrt_W=None, tol_fun=1e-4, tol_x=1e-6, max_iter=1000, verbose=False, x0=None,
          _dev=None, lipschitz_iters=0):
    """FISTA with backtracking (solvers.py:244-391; Beck & Teboulle).

    Returns (x, opt_details) with x a numpy float64 vector."""
    if not hasattr(A, "_plan"):
        raise TypeError("fista needs a commit_amd LinearOperator (device plan)")
    device = A.device
    n = A.shape[1]
    if _dev is None:
        rp = {"startIC": 0, "sizeIC": n, "startEC": n, "sizeEC": 0, "startISO": n, "sizeISO": 0,
              "regIC": None, "regEC": None, "regISO": None, "nnIC": False, "nnEC": False, "nnISO": False}
        _dev = DeviceRegularisation(rp, n, device)
    return _fista_device(y, A, _dev, sqrt_W, tol_fun, tol_x, max_iter, verbose, x0, lipschitz_iters)


class _Precision:
    """Temporarily switch the shared device plan to another vector precision."""

    def __init__(self, plan, precision):
        self.plan, self.precision = plan, precision

    def __enter__(self):
        self.saved = self.plan.precision
        self.plan.set_precision(self.precision)

    def __exit__(self, *exc):
        self.plan.set_precision(self.saved)


def _fista_device(y, A, reg, sqrt_W, tol_fun, tol_x, max_iter, verbose, x0, lipschitz_iters=0):
    lib = _lib.load()
    plan = A._plan.handle
    device = A.device
    m, n = A.nROWS, A.nCOLS
    dt = A._plan.dtype
    fp32 = dt == torch.float32
    vec = dict(dtype=dt, device=device)
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    axpby = lib.commit_hip_axpby_f32 if fp32 else lib.commit_hip_axpby

    def dev_vec(v, size, dtype=dt):
        if isinstance(v, torch.Tensor):
            t = v.to(device=device, dtype=dtype).contiguous()
        else:
            t = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(device=device, dtype=dtype)
        if t.numel() != size:
            raise RuntimeError("A.dot(): dimensions do not match")
        return t

    yd = dev_vec(y, m)
    sw = None if sqrt_W is None else dev_vec(sqrt_W, m)
    xhat = dev_vec(np.zeros(n) if x0 is None else x0, n).clone()
    res = torch.empty(m, **vec)
    grad = torch.empty(n, **vec)
    x = torch.empty(n, **vec)
    tmp = torch.empty(n, **vec)
    rsq = torch.empty(1, dtype=torch.float64, device=device)
    # the loop's scalars, read back once per trial step: the eight sums,
    # ||r(x)||^2 of the trial point, ||r(xhat)||^2 after a momentum update
    sc = torch.empty(10, dtype=torch.float64, device=device)
    st = _stream(yd)

    # residuals of the trial point and of the previous iterate: the momentum
    # point's residual is their affine combination (the residual is affine in x)
    rbuf = [torch.empty(m, **vec), torch.empty(m, **vec)]

    def residual(v, with_y=True, store=True, out=None, into=None):
        dst = (res if into is None else into) if store else None
        _lib.check(lib.commit_hip_A_residual(plan, P(v), P(yd) if with_y else None, P(sw), P(dst),
                                             P(rsq if out is None else out), st), "commit_hip_A_residual")
        return float(rsq.item()) if out is None else None

    def residual64(v, yv, swv, out):
        """||sqrtw (A v - y)||^2 with FP64 vectors on the same plan."""
        with _Precision(A._plan, _lib.FP64):
            _lib.check(lib.commit_hip_A_residual(plan, P(v), P(yv), P(swv), P(out), P(rsq), st),
                       "commit_hip_A_residual")
        return float(rsq.item())

    def gradient():
        _lib.check(lib.commit_hip_At_weighted(plan, P(res), P(sw), P(grad), st), "commit_hip_At_weighted")

    # initialisation (solvers.py:260-286)
    res_sq = residual(xhat)
    gradient()
    reg.prox(xhat, 1.0)
    reg_term = reg.omega_from(reg.sums(xhat))
    prev_obj = 0.5 * sqrt(res_sq) ** 2 + reg_term
    told = 1
    beta = 0.9
    prev_x = xhat.clone()
    qfval = prev_obj
    if lipschitz_iters > 0:
        L = _lipschitz_power(A, lib, plan, sw, grad, int(lipschitz_iters), st)
    else:
        if fp32:                       # the estimate is taken in FP64
            g64 = grad.double()
            sw64 = None if sw is None else sw.double()
            a_grad = residual64(g64, None, sw64, torch.empty(m, dtype=torch.float64, device=device))
            g_sq = float(torch.dot(g64, g64).item())
        else:
            a_grad = residual(grad, with_y=False)
            g_sq = reg.sums(grad)[3]
        with np.errstate(divide="ignore", invalid="ignore"):   # numpy semantics of solvers.py:283-286
            L = (np.sqrt(np.float64(a_grad)) / np.sqrt(np.float64(g_sq))) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        step_size = np.float64(1.9) / L
    if verbose == 4:
        print()
        print("      |  1/2||Ax-y||^2      Omega      |  Cost function    Abs error      Rel error    |      Abs x          Rel x    ")
        print("------|--------------------------------|-----------------------------------------------|------------------------------")
    it = 1

    qf_pending = False                                        # ||r(xhat)||^2 waits in sc[9]

    def trial(step):
        # x = prox(xhat - step*grad, step); q, objective   (solvers.py:299-312)
        nonlocal qfval, qf_pending
        reg.prox(x, step, xin=xhat, grad=grad, step=step)
        reg.sums(x, xhat, grad, prev_x, out=sc[:8])
        residual(x, out=sc[8:9], into=rbuf[0])                # r(x), kept for the momentum point
        v = sc.cpu().numpy()                                  # one host sync per trial step
        if qf_pending:
            qfval = 0.5 * sqrt(v[9]) ** 2
            qf_pending = False
        s = v[:8]
        reg_x = reg.omega_from(s)
        q = qfval + s[0] + 0.5 / step * sqrt(s[1]) ** 2 + reg_x
        rn = sqrt(v[8])
        return reg_x, q, rn, 0.5 * rn ** 2 + reg_x, s

    while True:
        if verbose == 4:
            print("%4d  |" % it, end="")
            sys.stdout.flush()
        reg_term_x, q, res_norm, curr_obj, s = trial(step_size)
        while curr_obj > q:                                   # backtracking (solvers.py:315-332)
            step_size = beta * step_size
            reg_term_x, q, res_norm, curr_obj, s = trial(step_size)
        abs_obj = abs(curr_obj - prev_obj)
        rel_obj = abs_obj / curr_obj
        abs_x = sqrt(s[2])
        rel_x = abs_x / (sqrt(s[3]) + eps)
        if verbose == 4:
            print("  %13.7e  %13.7e  |  %13.7e  %13.7e  %13.7e  |  %13.7e  %13.7e" % (
                0.5 * res_norm ** 2, reg_term_x, curr_obj, abs_obj, rel_obj, abs_x, rel_x), flush=True)
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
        # FISTA update (solvers.py:359-375): xhat = x + (told-1)/t * (x - prev_x)
        t = 0.5 * (1 + sqrt(1 + 4 * told ** 2))
        mom = (told - 1) / t
        _lib.check(axpby(n, 1.0, P(x), -1.0, P(prev_x), P(tmp), st), "commit_hip_axpby")
        _lib.check(axpby(n, 1.0, P(x), mom, P(tmp), P(xhat), st), "commit_hip_axpby")
        # r(xhat) = (1 + mom) r(x) - mom r(prev_x), with its norm (read with the
        # next trial's scalars), formed while the gradient's A^T loads it; the
        # first update has no r(prev_x) and runs A xhat
        if it > 1:
            _lib.check(lib.commit_hip_At_combined(plan, 1.0 + mom, P(rbuf[0]), -mom, P(rbuf[1]), P(sw), P(res),
                                                  P(grad), P(sc[9:10]), st), "commit_hip_At_combined")
        else:
            residual(xhat, out=sc[9:10])
            gradient()
        it += 1
        prev_obj = curr_obj
        prev_x.copy_(x)
        rbuf.reverse()                                        # r(prev_x) := r(x)
        told = t
        qf_pending = True
    if verbose == 4:
        print("< Stopping criterion: %s >" % criterion, flush=True)
    details = {"residual": 0.5 * res_norm ** 2, "regterm": reg_term_x, "cost_function": curr_obj,
               "abs_cost": abs_obj, "rel_cost": rel_obj, "abs_x": abs_x, "rel_x": rel_x,
               "iterations": it, "stopping_criterion": criterion}
    if fp32:
        # final residual reported in FP64 (x as stored, y and weights in float64)
        x64 = x.double()
        y64 = dev_vec(y, m, torch.float64)
        sw64 = None if sqrt_W is None else dev_vec(sqrt_W, m, torch.float64)
        r64 = torch.empty(m, dtype=torch.float64, device=device)
        rn64 = sqrt(residual64(x64, y64, sw64, r64))
        details["residual"] = 0.5 * rn64 ** 2
        details["cost_function"] = details["residual"] + reg_term_x
        details["precision"] = "fp32"
    return x.double().cpu().numpy(), details


def _lipschitz_power(A, lib, plan, sw, g, iters, st):
    """Largest eigenvalue of A^T W A by `iters` FP64 power iterations started
    from g (the first gradient), with the operator in FP64."""
    device = A.device
    m, n = A.nROWS, A.nCOLS
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    v = g.double().clone()
    nv = float(torch.linalg.vector_norm(v).item())
    if not nv > 0:
        return np.float64(np.nan)
    v /= nv
    sw64 = None if sw is None else sw.double()
    r = torch.empty(m, dtype=torch.float64, device=device)
    rsq = torch.empty(1, dtype=torch.float64, device=device)
    w = torch.empty(n, dtype=torch.float64, device=device)
    lam = 0.0
    with _Precision(A._plan, _lib.FP64):
        for _ in range(iters):
            _lib.check(lib.commit_hip_A_residual(plan, P(v), None, P(sw64), P(r), P(rsq), st), "commit_hip_A_residual")
            _lib.check(lib.commit_hip_At_weighted(plan, P(r), P(sw64), P(w), st), "commit_hip_At_weighted")
            lam = float(torch.linalg.vector_norm(w).item())
            if not lam > 0:
                break
            v = w / lam
    return np.float64(lam)
