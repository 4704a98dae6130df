"""Pins the oracle's FISTA trajectory (oracle.fista, restating
solvers.py:244-391) against hand-derived iterates (tests/_fista_cases.py):
the step constant 1.9/L, the backtracking factor 0.9, the momentum sequence,
the square root of the confidence weights and the order of the five stopping
tests.  The operator counts its products, so the number of backtracking steps
is checked as well as the iterates they lead to."""
import numpy as np
import pytest

from _fista_cases import cases
from oracle import oracle


class _Counting:
    """Dense diagonal operator that counts A.dot and A.T.dot calls."""

    def __init__(self, a, counts, adjoint=False):
        self.a, self.counts, self.adjoint = np.asarray(a, np.float64), counts, adjoint
        self.shape = (self.a.size, self.a.size)

    def dot(self, v):
        self.counts["At" if self.adjoint else "A"] += 1
        return self.a * v

    @property
    def T(self):
        return _Counting(self.a, self.counts, not self.adjoint)


def _nn(n):
    rp = {"startIC": 0, "sizeIC": n, "startEC": n, "sizeEC": 0, "startISO": n, "sizeISO": 0,
          "regIC": None, "regEC": None, "regISO": None, "nnIC": True, "nnEC": True, "nnISO": True}
    return oracle.init_regularisation(rp)


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_fista_hand_derived_trajectory(case):
    counts = {"A": 0, "At": 0}
    A = _Counting(case["a"], counts)
    x, det = oracle.solve(case["y"], A, A.T, tol_fun=case["tol_fun"], tol_x=case["tol_x"],
                          max_iter=case["max_iter"], x0=None if case["x0"] is None else case["x0"].copy(),
                          regularisation=_nn(case["a"].size), confidence_array=case["conf"])
    assert det["iterations"] == case["iterations"]
    assert det["stopping_criterion"] == case["criterion"]
    # 2 products before the loop, one per trial step, one per momentum point
    assert counts["A"] == case["n_A"], counts
    assert counts["At"] == case["n_At"], counts
    exp = np.asarray(case["x"])
    assert np.allclose(x, exp, rtol=1e-13, atol=0.0), (x, exp)
    for k, v in case.get("details", {}).items():
        assert det[k] == pytest.approx(v, rel=1e-12, abs=0.0), k


def test_fista_cases_cover_every_stopping_test():
    got = {c["criterion"] for c in cases()}
    assert len(got) == 5
