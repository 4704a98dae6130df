"""Batched fits: many tractograms of uneven size and many right-hand sides
(lambda sweeps, bootstrap signals) solved together on one device.

* ``combine`` stacks tractograms into one block-diagonal dictionary, so that a
  single plan (and a single launch of every product) covers all of them.
* ``BatchedFista`` runs S x R independent FISTA fits (S tractograms, R columns)
  with the reference's control flow (solvers.py:244-391) evaluated per fit on
  the device; one trial step is a fixed kernel sequence, captured once in a
  HIP graph and replayed.

No counterpart in the reference, which builds one operator per tractogram and
calls ``solve`` once per right-hand side.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .operator.operator import LinearOperator
from .solvers import _REG, _slice_range

_CRITERIA = {1: "Absolute tolerance on the objective", 2: "Relative tolerance on the objective",
             3: "Absolute tolerance on the unknowns", 4: "Relative tolerance on the unknowns",
             5: "Maximum number of iterations"}


class FistaCol(ctypes.Structure):
    """commit_hip_fista_col (include/commit_hip.h)."""
    _fields_ = [("step", ctypes.c_double), ("told", ctypes.c_double), ("prev_obj", ctypes.c_double),
                ("qfval", ctypes.c_double), ("coef", ctypes.c_double), ("cost", ctypes.c_double),
                ("regterm", ctypes.c_double), ("residual", ctypes.c_double), ("abs_cost", ctypes.c_double),
                ("rel_cost", ctypes.c_double), ("abs_x", ctypes.c_double), ("rel_x", ctypes.c_double),
                ("lam", ctypes.c_double * 3), ("lam_group", ctypes.c_double),
                ("iterations", ctypes.c_int32), ("done", ctypes.c_int32), ("criterion", ctypes.c_int32),
                ("advance", ctypes.c_int32), ("trials", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class MultiLayout(ctypes.Structure):
    """commit_hip_multi_layout (include/commit_hip.h)."""
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("nrhs", ctypes.c_int), ("nprob", ctypes.c_int),
                ("segx", ctypes.c_void_p), ("xpieces", ctypes.c_void_p), ("xpseg", ctypes.c_void_p),
                ("nxpieces", ctypes.c_int64), ("ypieces", ctypes.c_void_p), ("ypseg", ctypes.c_void_p),
                ("nypieces", ctypes.c_int64), ("gseg", ctypes.c_void_p), ("vseg", ctypes.c_void_p)]


class Combined:
    """Block-diagonal stack of S dictionaries sharing one set of kernels.

    Unknowns keep the reference's compartment-major layout over the stack:
    x = [IC: RF i, tractogram s, fibre f | EC: RF i, s, peak e | ISO: RF i, s, voxel k],
    rows are the tractograms' voxels one after another (voxel-major, nS each)."""

    def __init__(self, dictionaries, kernels):
        self.K = kernels
        self.S = len(dictionaries)
        nIC, nEC, nISO = (kernels["wmr"].shape[0], kernels["wmh"].shape[0], kernels["iso"].shape[0])
        nS = kernels["wmr"].shape[2] if kernels["wmr"].size else kernels["iso"].shape[1]
        self.nIC, self.nEC, self.nISO, self.nS = nIC, nEC, nISO, nS
        nF = [int(D["IC"]["nSTR"]) for D in dictionaries]
        nE = [int(D["EC"]["n"]) for D in dictionaries]
        nI = [int(D["ISO"]["n"]) for D in dictionaries]
        nV = [int(D["IC"]["nVOX"]) for D in dictionaries]
        self.fo, self.eo, self.io, self.vo = (np.concatenate([[0], np.cumsum(a)[:-1]]).astype(np.int64)
                                              for a in (nF, nE, nI, nV))
        self.nF, self.nE, self.nI, self.nV = nF, nE, nI, nV
        F, E, I, V = sum(nF), sum(nE), sum(nI), sum(nV)
        self.F, self.E, self.I, self.V = F, E, I, V
        cat = lambda key, sub, off: np.concatenate(
            [np.asarray(D[key][sub]).astype(np.int64) + o for D, o in zip(dictionaries, off)]) \
            if dictionaries else np.zeros(0)
        IC = {"str": cat("IC", "str", self.fo).astype(np.uint32), "vox": cat("IC", "vox", self.vo).astype(np.uint32),
              "dir": np.concatenate([D["IC"]["dir"] for D in dictionaries]).astype(np.uint16),
              "len": np.concatenate([D["IC"]["len"] for D in dictionaries]).astype(np.float32),
              "nSTR": F, "nVOX": V}
        IC["n"] = int(IC["str"].size)
        order = np.argsort(IC["vox"], kind="mergesort")          # voxel-sorted (core.pyx:509-513)
        for k in ("str", "vox", "dir", "len"):
            IC[k] = IC[k][order]
        IC["eval"] = np.ones(F * nIC, dtype=np.uint32)
        EC = {"vox": cat("EC", "vox", self.vo).astype(np.uint32),
              "dir": np.concatenate([D["EC"]["dir"] for D in dictionaries]).astype(np.uint16)}
        EC["n"] = int(EC["vox"].size)
        ISO = {"vox": cat("ISO", "vox", self.vo).astype(np.uint32)}
        ISO["n"] = int(ISO["vox"].size)
        self.D = {"IC": IC, "EC": EC, "ISO": ISO}
        self.n = F * nIC + E * nEC + I * nISO
        self.m = V * nS

    def columns(self, s):
        """Indices of tractogram s's unknowns in the stacked x, in its own order."""
        parts = []
        for i in range(self.nIC):
            parts.append(i * self.F + self.fo[s] + np.arange(self.nF[s]))
        b = self.nIC * self.F
        for i in range(self.nEC):
            parts.append(b + i * self.E + self.eo[s] + np.arange(self.nE[s]))
        b += self.nEC * self.E
        for i in range(self.nISO):
            parts.append(b + i * self.I + self.io[s] + np.arange(self.nI[s]))
        return np.concatenate(parts).astype(np.int64) if parts else np.zeros(0, np.int64)

    def rows(self, s):
        return np.arange(self.vo[s] * self.nS, (self.vo[s] + self.nV[s]) * self.nS, dtype=np.int64)

    def x_ranges(self, s):
        """[begin, end) ranges of tractogram s's unknowns (one per compartment RF)."""
        out = []
        for i in range(self.nIC):
            b = i * self.F + self.fo[s]
            out.append((b, b + self.nF[s]))
        base = self.nIC * self.F
        for i in range(self.nEC):
            b = base + i * self.E + self.eo[s]
            out.append((b, b + self.nE[s]))
        base += self.nEC * self.E
        for i in range(self.nISO):
            b = base + i * self.I + self.io[s]
            out.append((b, b + self.nI[s]))
        return [(int(a), int(b)) for a, b in out if b > a]

    def operator(self, precision=_lib.FP32, device=None):
        return LinearOperator(self.D, self.K, None, precision=precision, device=device)


def _pieces(ranges, chunk):
    """Split (begin, end, problem) ranges into pieces of at most `chunk` rows."""
    pcs, seg = [], []
    for b, e, s in ranges:
        for a in range(b, e, chunk):
            pcs.append((a, min(e, a + chunk)))
            seg.append(s)
    return np.asarray(pcs, dtype=np.int64).reshape(-1, 2), np.asarray(seg, dtype=np.int32)


class BatchedFista:
    """S x R independent FISTA fits over one (block-diagonal) operator.

    A     : LinearOperator of the stacked dictionary (any precision; the
            products run through the float32 multi-RHS kernels)
    layout: Combined describing the problems (None: one problem, all of A)
    y     : float32-convertible [m] (shared by the columns) or [m, R]
    rp    : regularisation dict as for solvers.init_regularisation, over the
            stacked x (group indices in stacked x; groups must not overlap)
    lambdas: per fit penalty, array [S, R] (lasso or group lasso of IC) or
            [S, R, 2] (sparse group lasso: l1, group); EC/ISO lambdas from rp
    combine: form the residual at the momentum point from the trial point's
            and the previous iterate's, r(xhat) = (1 + coef) r(x) - coef
            r(prev_x), inside the A^T product (one product A X per trial step);
            False recomputes A xhat as the reference does (solvers.py:362-368)
    """

    def __init__(self, A, y, rp, lambdas, R, layout=None, sqrt_W=None, tol_fun=1e-4, tol_x=1e-6, max_iter=100,
                 chunk=1 << 15, combine=True):
        self.lib = _lib.load()
        self.A, self.At = A, A.T
        self.dev = A.device
        n, m = A.nCOLS, A.nROWS
        self.n, self.m, self.R = n, m, int(R)
        if not 1 <= self.R <= 16:
            raise ValueError("batched fits take 1..16 columns")
        S = 1 if layout is None else layout.S
        self.S = S
        self.tol_fun, self.tol_x, self.max_iter = float(tol_fun), float(tol_x), int(max_iter)
        dev = self.dev
        f32 = dict(dtype=torch.float32, device=dev)
        # problems: unknown ranges and row ranges
        if layout is None:
            xr = [(0, n, 0)]
            yr = [(0, m, 0)]
            segx = np.zeros(n, np.int32)
        else:
            xr, yr = [], []
            segx = np.zeros(n, np.int32)
            for s in range(S):
                for b, e in layout.x_ranges(s):
                    xr.append((b, e, s))
                    segx[b:e] = s
                rr = layout.rows(s)
                if rr.size:
                    yr.append((int(rr[0]), int(rr[-1]) + 1, s))
        xp, xs = _pieces(xr, chunk)
        yp, ys = _pieces(yr, chunk * 8)
        nS = int(A.nSAMPLES)
        vseg = np.zeros(m // nS if nS else 0, np.int32)          # problem of each voxel
        for b, e, s in yr:
            vseg[b // nS:e // nS] = s
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self._keep = [t(segx), t(xp), t(xs), t(yp), t(ys), t(vseg)]
        L = MultiLayout()
        L.n, L.m, L.nrhs, L.nprob = n, m, self.R, S
        L.segx, L.xpieces, L.xpseg = (self._keep[0].data_ptr(), self._keep[1].data_ptr(), self._keep[2].data_ptr())
        L.nxpieces = int(xp.shape[0])
        L.ypieces, L.ypseg = self._keep[3].data_ptr(), self._keep[4].data_ptr()
        L.nypieces = int(yp.shape[0])
        L.vseg = self._keep[5].data_ptr()
        # penalty
        from .solvers import DeviceRegularisation
        reg = DeviceRegularisation(dict(rp), n, dev)
        self.reg = reg
        self.params = reg.params
        ng = int(reg.params.ngroups) if rp["regIC"] in ("group_lasso", "sparse_group_lasso") else 0
        if ng:
            if reg.overlap:
                raise ValueError("batched fits need non-overlapping groups")
            gidx = reg.gidx.cpu().numpy()
            gptr = reg.gptr.cpu().numpy()
            gseg = np.array([segx[gidx[gptr[g]]] if gptr[g + 1] > gptr[g] else 0 for g in range(ng)], np.int32)
            for g in range(ng):
                if gptr[g + 1] > gptr[g] and np.any(segx[gidx[gptr[g]:gptr[g + 1]]] != gseg[g]):
                    raise ValueError("a group spans two problems")
            self._keep.append(t(gseg))
            L.gseg = self._keep[-1].data_ptr()
        self.L = L
        self.work = torch.empty(int(self.lib.commit_hip_fista_multi_work(ctypes.byref(L), ng)) + 1,
                                dtype=torch.float64, device=dev)
        nrw = int(self.lib.commit_hip_A_multi_residual_work(A._plan.handle, ctypes.byref(L)))
        if nrw < 0:
            raise RuntimeError("commit_hip_A_multi_residual_work failed")
        self.rwork = torch.empty(nrw + 1, dtype=torch.float64, device=dev)
        # vectors
        yv = torch.as_tensor(np.asarray(y, dtype=np.float32)) if not isinstance(y, torch.Tensor) else y
        yv = yv.to(**f32).contiguous()
        if yv.dim() == 1:
            if yv.numel() != m:
                raise RuntimeError("y: dimensions do not match")
            self.ypc = 0
        else:
            if tuple(yv.shape) != (m, self.R):
                raise RuntimeError("y: expected [m] or [m, R]")
            self.ypc = 1
        self.y = yv
        self.sw = None if sqrt_W is None else torch.as_tensor(np.asarray(sqrt_W, np.float32)).to(**f32).contiguous()
        self.xhat = torch.zeros((n, self.R), **f32)
        self.x = torch.zeros((n, self.R), **f32)
        self.prev = torch.zeros((n, self.R), **f32)
        self.grad = torch.zeros((n, self.R), **f32)
        self.combine = bool(combine)
        # residual rows: with combine, r(prev_x) in r[sel] and the trial point's in r[1 - sel]
        self.r = [torch.zeros((m, self.R), **f32)]
        if self.combine:
            self.r.append(torch.zeros((m, self.R), **f32))
        self.sel = torch.zeros(1, dtype=torch.int32, device=dev)
        self.res = self.r[0]
        nq = S * self.R
        self.sums = torch.zeros(nq * 8, dtype=torch.float64, device=dev)
        self.rsq_x = torch.zeros(nq, dtype=torch.float64, device=dev)
        self.rsq_h = torch.zeros(nq, dtype=torch.float64, device=dev)
        # per-fit penalties
        lam = np.asarray(lambdas, dtype=np.float64)
        kind = rp["regIC"]
        if kind == "sparse_group_lasso":
            if lam.shape != (S, self.R, 2):
                raise ValueError("sparse group lasso lambdas must have shape [S, R, 2]")
        elif kind is not None and lam.shape != (S, self.R):
            raise ValueError("lambdas must have shape [S, R]")
        cols = (FistaCol * nq)()
        p = self.params
        for s in range(S):
            for r in range(self.R):
                c = cols[s * self.R + r]
                c.lam[0], c.lam[1], c.lam[2] = p.lam[0], p.lam[1], p.lam[2]
                c.lam_group = p.lam_group
                if kind == "lasso":
                    c.lam[0] = lam[s, r]
                elif kind == "group_lasso":
                    c.lam_group = lam[s, r]
                elif kind == "sparse_group_lasso":
                    c.lam[0], c.lam_group = lam[s, r, 0], lam[s, r, 1]
                c.step = 1.0
        self.kind = kind
        self._cols_host = cols
        self.col = torch.empty(ctypes.sizeof(cols), dtype=torch.uint8, device=dev)
        self._upload_cols()
        self.graph = None

    # ------------------------------------------------------------- pieces --
    def _st(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def _upload_cols(self):
        host = torch.frombuffer(bytearray(bytes(self._cols_host)), dtype=torch.uint8)
        self.col.copy_(host)

    def _download_cols(self):
        raw = self.col.cpu().numpy().tobytes()
        cols = (FistaCol * (self.S * self.R)).from_buffer_copy(raw)
        self._cols_host = cols
        return cols

    def _P(self, t):
        return None if t is None else ctypes.c_void_p(t.data_ptr())

    def _check(self, rc, what):
        _lib.check(rc, what)

    def _prox(self, xin, grad, out):
        self._check(self.lib.commit_hip_prox_multi(ctypes.byref(self.params), ctypes.byref(self.L), self._P(xin),
                                                   self._P(grad), self._P(out), self._P(self.col), self._st()),
                    "commit_hip_prox_multi")

    def _sums(self, x, xhat=None, grad=None, prev=None):
        reg = self.reg
        self._check(self.lib.commit_hip_fista_multi_sums(ctypes.byref(self.params), ctypes.byref(self.L), self._P(x),
                                                         self._P(xhat), self._P(grad), self._P(prev), reg.l1_range,
                                                         reg.l1_weighted, self._P(self.sums), self._P(self.work),
                                                         self._st()), "commit_hip_fista_multi_sums")

    def _resid(self, X, y, ypc, res, rsq):
        """res = sqrtw .* (A X - y) (res None: norms only; y None: zero) and the
        squared norm of every fit, in one pass (commit_hip_A_multi_residual)."""
        self._check(self.lib.commit_hip_A_multi_residual(self.A._plan.handle, ctypes.byref(self.L), self._P(X),
                                                         self._P(y), ypc, self._P(self.sw), self._P(res),
                                                         self._P(rsq), self._P(self.rwork), self._st()),
                    "commit_hip_A_multi_residual")

    def _omega_host(self, s8, c):
        reg = self.params.reg
        om = 0.0
        if reg[0] == _REG["lasso"]:
            om += c.lam[0] * s8[4]
        elif reg[0] == _REG["group_lasso"]:
            om += c.lam_group * s8[7] if c.lam_group != 0 else 0.0
        elif reg[0] == _REG["sparse_group_lasso"]:
            if c.lam[0] != 0 or c.lam_group != 0:
                om += c.lam[0] * s8[4] + c.lam_group * s8[7]
        if reg[1] == _REG["lasso"]:
            om += c.lam[1] * s8[5]
        if reg[2] == _REG["lasso"]:
            om += c.lam[2] * s8[6]
        return om

    # --------------------------------------------------------------- init --
    def initialise(self, x0=None):
        """FISTA initialisation (solvers.py:260-286) for every fit."""
        if x0 is not None:
            self.xhat.copy_(torch.as_tensor(x0, dtype=torch.float32).reshape(self.n, self.R))
        # residual and gradient at x0 (r[0]: r(prev_x) of the first momentum step, whose coef is 0)
        self.sel.zero_()
        self._resid(self.xhat, self.y, self.ypc, self.r[0], self.rsq_h)
        self.At.matmat(self.r[0], out=self.grad, sqrtw=self.sw)
        # in-place prox(xhat, 1) with step 1 (col.step = 1 from the constructor)
        zero = torch.zeros_like(self.grad)
        self._prox(self.xhat, zero, self.x)
        self.xhat.copy_(self.x)
        self._sums(self.xhat)
        s_x = self.sums.cpu().numpy().reshape(-1, 8)
        rsq0 = self.rsq_h.cpu().numpy()
        # Lipschitz quotient (||A grad|| / ||grad||)^2 per fit (solvers.py:282-286)
        self._resid(self.grad, None, 0, None, self.rsq_x)
        self._sums(self.grad)
        s_g = self.sums.cpu().numpy().reshape(-1, 8)
        a_g = self.rsq_x.cpu().numpy()
        cols = self._download_cols()
        for q in range(self.S * self.R):
            c = cols[q]
            reg_term = self._omega_host(s_x[q], c)
            c.prev_obj = 0.5 * np.sqrt(rsq0[q]) ** 2 + reg_term
            c.qfval = c.prev_obj
            c.told = 1.0
            with np.errstate(divide="ignore", invalid="ignore"):
                Lq = (np.sqrt(np.float64(a_g[q])) / np.sqrt(np.float64(s_g[q][3]))) ** 2
                c.step = float(np.float64(1.9) / Lq)
            c.iterations = 1
            c.done = c.criterion = c.advance = c.trials = 0
        self._cols_host = cols
        self._upload_cols()
        self.prev.copy_(self.xhat)

    # --------------------------------------------------------------- step --
    def trial(self):
        """One trial step of every unfinished fit (the captured sequence)."""
        self._prox(self.xhat, self.grad, self.x)
        self._sums(self.x, self.xhat, self.grad, self.prev)
        if self.combine:
            self._check(self.lib.commit_hip_A_multi_trial(self.A._plan.handle, ctypes.byref(self.L), self._P(self.x),
                                                          self._P(self.y), self.ypc, self._P(self.sw),
                                                          self._P(self.r[0]), self._P(self.r[1]), self._P(self.sel),
                                                          self._P(self.rsq_x), self._P(self.rwork), self._st()),
                        "commit_hip_A_multi_trial")
        else:
            self._resid(self.x, self.y, self.ypc, None, self.rsq_x)
        self._check(self.lib.commit_hip_fista_multi_control(ctypes.byref(self.params), ctypes.byref(self.L),
                                                            self.tol_fun, self.tol_x, self.max_iter,
                                                            self._P(self.sums), self._P(self.rsq_x),
                                                            self._P(self.col), self._st()),
                    "commit_hip_fista_multi_control")
        self._check(self.lib.commit_hip_fista_multi_update(ctypes.byref(self.L), self._P(self.x), self._P(self.xhat),
                                                           self._P(self.prev), self._P(self.col), self._st()),
                    "commit_hip_fista_multi_update")
        if self.combine:
            # gradient and ||r(xhat)||^2 of the fits that advanced, then r(prev_x) back into r[sel]
            self._check(self.lib.commit_hip_At_multi_momentum(self.A._plan.handle, ctypes.byref(self.L),
                                                              self._P(self.r[0]), self._P(self.r[1]),
                                                              self._P(self.sel), self._P(self.sw), self._P(self.col),
                                                              self._P(self.grad), self._P(self.rsq_h),
                                                              self._P(self.rwork), self._st()),
                        "commit_hip_At_multi_momentum")
            self._check(self.lib.commit_hip_fista_multi_qf(ctypes.byref(self.L), self._P(self.rsq_h),
                                                           self._P(self.col), self._st()), "commit_hip_fista_multi_qf")
            self._check(self.lib.commit_hip_fista_multi_roles(self.A._plan.handle, ctypes.byref(self.L),
                                                              self._P(self.r[0]), self._P(self.r[1]),
                                                              self._P(self.sel), self._P(self.col), self._st()),
                        "commit_hip_fista_multi_roles")
            return
        self._resid(self.xhat, self.y, self.ypc, self.res, self.rsq_h)
        self._check(self.lib.commit_hip_fista_multi_qf(ctypes.byref(self.L), self._P(self.rsq_h), self._P(self.col),
                                                       self._st()), "commit_hip_fista_multi_qf")
        self.At.matmat(self.res, out=self.grad, sqrtw=self.sw)

    def capture(self):
        """Capture one trial step in a HIP graph (after initialise)."""
        side = torch.cuda.Stream(self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            self.trial()
        torch.cuda.current_stream(self.dev).wait_stream(side)
        self.graph = g
        return g

    def all_done(self):
        cols = self._download_cols()
        return all(cols[q].done for q in range(self.S * self.R))

    def run(self, use_graph=True, check_every=8, max_trials=None):
        """Trial steps until every fit has stopped; returns the number of steps."""
        if max_trials is None:
            max_trials = 20 * self.max_iter + 100
        if use_graph and self.graph is None:
            self.capture()
        k = 0
        while k < max_trials:
            for _ in range(check_every):
                if use_graph:
                    self.graph.replay()
                else:
                    self.trial()
                k += 1
            torch.cuda.synchronize(self.dev)
            if self.all_done():
                break
        return k

    # ------------------------------------------------------------ results --
    def details(self):
        cols = self._download_cols()
        out = []
        for s in range(self.S):
            row = []
            for r in range(self.R):
                c = cols[s * self.R + r]
                row.append({"residual": c.residual, "regterm": c.regterm, "cost_function": c.cost,
                            "abs_cost": c.abs_cost, "rel_cost": c.rel_cost, "abs_x": c.abs_x, "rel_x": c.rel_x,
                            "iterations": int(c.iterations),
                            "stopping_criterion": _CRITERIA.get(int(c.criterion), "not stopped"),
                            "trials": int(c.trials)})
            out.append(row)
        return out

    def solution(self):
        """x as float64 numpy [n, R] (stacked layout)."""
        return self.x.double().cpu().numpy()
