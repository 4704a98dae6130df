"""Hand-derived FISTA trajectories (solvers.py:244-391) on tiny diagonal
problems, shared by the oracle pins (test_oracle_fista.py) and the device
solver tests (test_gpu_fista_trajectory.py).

The expected values below come from the mathematics of the method, not from
running any implementation.  With f(x) = 0.5 ||A x - y||^2, the first gradient
g = A^T (A x0 - y), the step constant L = (||A g|| / ||g||)^2
(solvers.py:282-287) and a trial point x = xhat - s g where the
non-negativity prox does not act:

    f(x) = f(xhat) - s ||g||^2 + s^2/2 ||A g||^2
    q    = f(xhat) - s ||g||^2 + s/2   ||g||^2            (solvers.py:305)

so the backtracking test f(x) > q (solvers.py:315) holds exactly when
s > ||g||^2 / ||A g||^2.  At the first iteration that threshold is 1/L, and
the trial steps are 1.9/L, 0.9 * 1.9/L, ... (solvers.py:287, 318): the first
step backtracks while 1.9 * 0.9^k > 1, that is exactly 7 times
(1.9 * 0.9^6 = 1.0097, 1.9 * 0.9^7 = 0.9088).  Coordinates where the prox
clips to 0 add the same constant to f(x) and q and do not change this.

For a diagonal A = diag(a) started at 0 and a fixed step s, plain gradient
steps give x_k[i] = (y_i / a_i) (1 - (1 - s a_i^2)^k); the momentum
coefficient of iteration k+1 is (t_k - 1) / t_{k+1} with t_1 = 1,
t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 (Beck & Teboulle), so t_2 is the golden
ratio phi = (1 + sqrt 5)/2, t_3 = (1 + sqrt(7 + 2 sqrt 5))/2 (since
1 + 4 phi^2 = 7 + 2 sqrt 5), and the first non-zero coefficient is
(phi - 1) / t_3, at the third iteration.

The stopping cases are built so that a given test of the ordered list
(solvers.py:342-356) fires together with every later one; the expected
criterion is the first, so any reordering of the five tests changes at least
one expected answer.
"""
from math import sqrt

import numpy as np

C7 = 1.9 * 0.9 ** 7                  # the accepted first step times L
EPS = np.finfo(float).eps
PHI = (1.0 + sqrt(5.0)) / 2.0
T3 = (1.0 + sqrt(7.0 + 2.0 * sqrt(5.0))) / 2.0
MOM3 = (PHI - 1.0) / T3              # momentum coefficient of the third iteration

CRIT = {1: "Absolute tolerance on the objective", 2: "Relative tolerance on the objective",
        3: "Absolute tolerance on the unknowns", 4: "Relative tolerance on the unknowns",
        5: "Maximum number of iterations"}


def _case(name, a, y, x0=None, conf=None, max_iter=1, tol_fun=1e-4, tol_x=1e-6, **exp):
    a = np.asarray(a, dtype=np.float64)
    return dict(name=name, a=a, y=np.asarray(y, dtype=np.float64),
                x0=None if x0 is None else np.asarray(x0, dtype=np.float64),
                conf=None if conf is None else np.asarray(conf, dtype=np.float64),
                max_iter=max_iter, tol_fun=tol_fun, tol_x=tol_x, **exp)


def cases(a_huge=1e20):
    """The trajectory cases; a_huge is the large diagonal entry of the
    unknowns-tolerance cases (the device tests pass its float32 value)."""
    out = []

    # A = [2], y = [3]: one iteration, 7 backtracks, x1 = 1.9*0.9^7 * y/a;
    # f(x0) = y^2/2, f(x1) = y^2/2 (1 - c)^2.
    a, y = 2.0, 3.0
    x1 = C7 * y / a
    f0, f1 = 0.5 * y * y, 0.5 * y * y * (1 - C7) ** 2
    out.append(_case("scalar_first_step", [a], [y], max_iter=1, x=[x1], iterations=1, criterion=CRIT[5],
                     backtracks=[7], n_A=2 + 8, n_At=1,
                     details=dict(residual=f1, regterm=0.0, cost_function=f1, abs_cost=f0 - f1,
                                  rel_cost=(f0 - f1) / f1, abs_x=x1, rel_x=x1 / (x1 + EPS))))

    # A = diag(1, 2), y = (1, 1): g = -(1, 2), L = ||(1, 4)||^2 / ||(1, 2)||^2 = 17/5.
    # Iteration 1: s = (1.9 / 3.4) 0.9^7 (7 backtracks); iteration 2 keeps s
    # (s < ||g||^2/||Ag||^2 for the new gradient), no momentum yet (t_1 = 1):
    # x2[i] = (1/a_i)(1 - (1 - s a_i^2)^2).
    s = 1.9 / 3.4 * 0.9 ** 7
    d = np.array([1.0, 2.0])
    x2 = (1.0 / d) * (1 - (1 - s * d * d) ** 2)
    out.append(_case("diag2_second_iterate", d, [1.0, 1.0], max_iter=2, x=x2, iterations=2, criterion=CRIT[5],
                     backtracks=[7, 0], n_A=2 + 8 + 1 + 1, n_At=2))

    # A = [1], y = [1], three iterations: the error e = 1 - x follows
    # e1 = 1 - c, e2 = e1^2 (step c, no momentum), then the momentum point
    # xhat2 = x2 + m (x2 - x1) has error e2 + m (e2 - e1) and one step with
    # s = c multiplies the error by (1 - c).  tol 0: only the iteration cap stops.
    e1 = 1 - C7
    e2 = e1 * e1
    e3 = e1 * (e2 + MOM3 * (e2 - e1))
    out.append(_case("scalar_momentum", [1.0], [1.0], max_iter=3, tol_fun=0.0, tol_x=0.0, x=[1 - e3],
                     iterations=3, criterion=CRIT[5], backtracks=[7, 0, 0], n_A=2 + 8 + 1 + 1 + 1 + 1, n_At=3))

    # Confidence weights enter as sqrt(w) (solvers.py:238-239): diag(1, 2)
    # with w = (4, 1) is the unweighted problem diag(2, 2), y' = (2, 1).
    # L = 4, x1 = c/4 * A'^T y' = c (1, 1/2).
    out.append(_case("weighted_first_step", [1.0, 2.0], [1.0, 1.0], conf=[4.0, 1.0], max_iter=1,
                     x=[C7, 0.5 * C7], iterations=1, criterion=CRIT[5], backtracks=[7], n_A=10, n_At=1))

    # Infeasible start, A = [1], y = [1], x0 = [-1]: the residual, gradient and
    # prev_obj are taken at x0 (res = -2, g = -2, prev_obj = 2), then xhat is
    # projected to 0 (solvers.py:262-273).  With step s the trial point is
    # x = 2s, q = 2 - 4s + (0.5/s)(2s)^2 = 2 - 2s and f(x) = (2s - 1)^2 / 2, so
    # f > q exactly when s > sqrt(3)/2: 1.9*0.9^7 = 0.909 backtracks again,
    # 1.9*0.9^8 = 0.818 is accepted (8 backtracks), x1 = 2 * 1.9 * 0.9^8.
    out.append(_case("infeasible_start", [1.0], [1.0], x0=[-1.0], max_iter=1, x=[2 * 1.9 * 0.9 ** 8],
                     iterations=1, criterion=CRIT[5], backtracks=[8], n_A=2 + 9, n_At=1))

    # Stop 1 fires with all the others: A = [1], y = [-1].  g = 1, x = prox(-1.9) = 0
    # = xhat, so q = f = 1/2, no backtrack, abs_obj = 0 = abs_x = rel_obj = rel_x.
    out.append(_case("stop_abs_objective", [1.0], [-1.0], max_iter=1, x=[0.0], iterations=1, criterion=CRIT[1],
                     backtracks=[0], n_A=3, n_At=1))

    # Stop 2 fires with 3, 4 and 5: a clipped coordinate (y = -1000) holds the
    # objective at ~5e5, the x-change c*1e-20 of the huge coordinate is below
    # eps, and the unchanged coordinate x = 1 keeps ||x|| = 1.
    # abs_obj = 0.5 (1 - (1-c)^2) ~ 0.4958, rel_obj ~ 1e-6 < tol_fun.
    ah = a_huge
    out.append(_case("stop_rel_objective", [ah, 1.0, 1.0], [1.0, 1.0, -1000.0], x0=[0.0, 1.0, 0.0], max_iter=1,
                     x=[C7 / ah, 1.0, 0.0], iterations=1, criterion=CRIT[2], backtracks=[7], n_A=10, n_At=1))

    # Stop 3 fires with 4 and 5: as above without the clipped coordinate, so
    # rel_obj = (1 - (1-c)^2) / (1-c)^2 ~ 119; abs_x = c/a_huge < eps; rel_x tiny.
    out.append(_case("stop_abs_unknowns", [ah, 1.0], [1.0, 1.0], x0=[0.0, 1.0], max_iter=1,
                     x=[C7 / ah, 1.0], iterations=1, criterion=CRIT[3], backtracks=[7], n_A=10, n_At=1))

    # Stop 4 fires with 5: a = 1e6, abs_x = c * 1e-6 = 9.1e-7 > eps and
    # rel_x = 9.1e-7 / ||(9.1e-7, 1)|| < tol_x = 1e-6.
    out.append(_case("stop_rel_unknowns", [1e6, 1.0], [1.0, 1.0], x0=[0.0, 1.0], max_iter=1,
                     x=[C7 * 1e-6, 1.0], iterations=1, criterion=CRIT[4], backtracks=[7], n_A=10, n_At=1))
    return out
