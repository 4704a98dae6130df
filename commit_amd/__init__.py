"""commit_amd — MI355X-native implementation of COMMIT's implicit linear
operator (A x, A^T y) and FISTA fit, behind the reference's drop-in surface
(commit.Evaluation, commit.operator.operator.LinearOperator)."""
__version__ = "0.1.0"


def __getattr__(name):
    # heavy modules (torch) load on first use
    if name in ("Evaluation", "setup"):
        from . import core
        return getattr(core, name)
    if name == "LinearOperator":
        from .operator.operator import LinearOperator
        return LinearOperator
    raise AttributeError(name)


__all__ = ["Evaluation", "setup", "LinearOperator"]
