"""MI355X-native batched Rodas3 chemistry step: a drop-in for the reference's CudaSolver call.

Importing the package does not import torch; CudaSolver and its result types load on first use.
"""

from ._constants import ATOL, COUNTS, DRIFT, REC_STEPS, T_MAX, T_MIN, X_SUM_TOL
from .kinvulcan import VulcanNetwork

__all__ = ["ATOL", "COUNTS", "DRIFT", "REC_STEPS", "T_MAX", "T_MIN", "X_SUM_TOL", "CallResult", "CudaSolver",
           "LaunchResult", "VulcanNetwork"]  # fmt: skip


def __getattr__(name: str):
    if name in ("CudaSolver", "CallResult", "LaunchResult"):
        from . import kincuda

        return getattr(kincuda, name)
    raise AttributeError(name)
