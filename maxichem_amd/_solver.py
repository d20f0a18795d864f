"""What the solver's API needs besides the kernels: input checks, the counts, refusal messages and the lifetime.

The checks, messages and count semantics are the reference's drop-in contract (maxichem/_solver.py:34-234): the same
ValueError texts before any launch and after a call that refused cells, the same count slots and check().
"""

from __future__ import annotations

import math
import operator
from typing import Any

import numpy as np
import torch

from ._constants import COUNT_SLOTS, COUNTS, DRIFT, REC_STEPS, T_MAX, T_MIN, X_SUM_TOL


def device_of(device: torch.device | str) -> torch.device:
    """A HIP device with its index ("cuda" alone means the current one)."""
    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError(f"a GPU device expected, got {d}")
    return torch.device("cuda", torch.cuda.current_device() if d.index is None else d.index)


def check_tensor(name: str, a: torch.Tensor, shape: tuple[int, ...], device: torch.device) -> None:
    """Refuse a tensor that is not float64 of this shape on this device (maxichem/_solver.py:34-39)."""
    if a.dtype != torch.float64 or a.device != device or tuple(a.shape) != shape:
        raise ValueError(
            f"{name}: float64 {list(shape)} on {device} expected, got {a.dtype} {list(a.shape)} on {a.device}"
        )


def check_inputs(dev, ns, x, T, P, dt, h, rtol, atol, max_steps) -> int:
    """Refuse inputs of the wrong form or range before any launch; the number of cells (_solver.py:42-70)."""
    if x.dim() != 2 or x.shape[1] != ns:
        raise ValueError(f"x: [B, {ns}] expected, got {list(x.shape)}")
    B = x.shape[0]
    check_tensor("x", x, (B, ns), dev)
    for name, a in (("T", T), ("P", P), ("dt", dt), ("h", h)):
        check_tensor(name, a, (B,), dev)
    if not (math.isfinite(rtol) and rtol > 0.0 and math.isfinite(atol) and atol > 0.0):
        raise ValueError(f"rtol and atol finite and positive expected, got {rtol}, {atol}")
    try:
        ms = operator.index(max_steps)
    except TypeError:
        raise ValueError(f"max_steps: an integer expected, got {max_steps!r}") from None
    most = int(np.iinfo(np.int32).max)  # tries are counted in an int, up to REC_STEPS * max_steps
    if ms < 1 or REC_STEPS * ms > most:
        raise ValueError(f"max_steps from 1 to {most // REC_STEPS} expected, got {ms}")
    return int(B)


def check_photolysis(J, cells: int, branches: int, device: torch.device):
    """The call's J: None without photolysis; refuse a J that breaks the rule (_solver.py:73-92)."""
    if not branches:
        if J is None:
            return None
        if J.dim() != 2 or J.shape[1] != 0:
            raise ValueError(
                f"J: none or an empty [{cells}, 0] one expected, since this solver has no photolysis; got "
                f"{list(J.shape)}"
            )
        check_tensor("J", J, (cells, 0), device)
        return None
    if J is None:
        raise ValueError(f"J [B, {branches}] is required with photolysis: pass zeros for dark cells")
    check_tensor("J", J, (cells, branches), device)
    return J.contiguous()


def _first(mask: torch.Tensor) -> int:
    return int(torch.nonzero(mask)[0, 0])


class Solver:
    """The lifetime, the counts on the device and the refusal logic (maxichem/_solver.py:117-234)."""

    nj: int

    def __init__(self, net: Any, device: torch.device | str, drift: float | None, photo: bool) -> None:
        if photo and (not net.photo or net.photo_missing):
            raise ValueError(
                "photo=True needs a network with photolysis branches whose species are all in its species "
                f"list (branches {len(net.photo)}, missing species {net.photo_missing})"
            )
        value = float(DRIFT if drift is None else drift)
        if not (math.isfinite(value) and value >= 0.0):
            raise ValueError(f"drift: a finite allowance of at least 0 expected, got {drift}")
        self.drift = value
        self.dev = device_of(device)
        self._module = None
        self._tables: dict[str, torch.Tensor] = {}
        self.count = torch.zeros(len(COUNT_SLOTS), dtype=torch.int32, device=self.dev)
        self.total = torch.zeros(len(COUNT_SLOTS), dtype=torch.int64, device=self.dev)
        self.flag: torch.Tensor | None = None

    def close(self) -> None:
        """Unload the kernels and drop the tables; the counts stay readable."""
        if self._module is not None:
            with torch.cuda.device(self.dev):
                torch.cuda.synchronize(self.dev)
                self._module.unload()
            self._module = None
        self._tables = {}

    def __enter__(self):
        return self

    def __exit__(self, *exception: object) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            if getattr(self, "_module", None) is not None:
                self._module.unload()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass

    def _require_open(self) -> None:
        if self._module is None:
            raise RuntimeError("this solver was closed; build a new one")

    @property
    def fallback(self) -> dict[str, int]:
        """The last call's counts (one small copy from the device)."""
        return dict(zip(COUNTS, self.count[1:-1].tolist()))

    def _refuse(self, x, T, P, dt, J, status) -> None:
        """Raise ValueError if the call just launched refused a cell, naming the first of each kind."""
        n = self.fallback
        if not (n["bad_input"] or n["cold"] or n["hot"]):
            return
        refused = status != 0
        cold, hot = refused & (T < T_MIN), refused & (T > T_MAX)
        problems = []
        for kind, cells, side in (("cold", cold, f"below {T_MIN:g} K"), ("hot", hot, f"above {T_MAX:g} K")):
            if n[kind]:
                i = _first(cells)
                problems.append(
                    f"{n[kind]} cells {side}, outside the solver's range {T_MIN:g}-{T_MAX:g} K (first: cell {i}, "
                    f"T = {float(T[i]):g} K); leave them out of the call"
                )
        if n["bad_input"]:
            i = _first(refused & torch.logical_not(cold | hot))
            total = 0.0
            for v in x[i].tolist():  # in index order, as the C ABI's solver adds them
                total += v
            where = (
                f"cell {i}: sum(x) - 1 = {total - 1.0:g} (added in another order than the kernel's, so "
                f"it can differ in the last bits), min(x) = {float(x[i].min()):g}, "
                f"T = {float(T[i]):g} K, P = {float(P[i]):g} dyn cm^-2, dt = {float(dt[i]):g} s"
                + ("" if J is None else f", min(J) = {float(J[i].min()):g} s^-1")
            )
            problems.append(
                f"{n['bad_input']} cells of bad input: x negative or non-finite, |sum(x) - 1| above {X_SUM_TOL:g} "
                "or no x above 0; T non-finite, P non-finite, not above 0 or so large that P / (k_B T) overflows, dt "
                f"negative or non-finite, or J negative or non-finite (first: {where}); pass x normalized to sum 1 in "
                "float64, clamped at 0"
            )
        raise ValueError("the call refused cells: " + "; ".join(problems))

    def check(self) -> dict[str, int]:
        """The counts since construction; raises if any cell was outside [T_MIN, T_MAX]."""
        n = dict(zip(COUNTS, self.total[1:-1].tolist()))
        if n["cold"] or n["hot"]:
            raise ValueError(
                f"{n['cold']} cells below {T_MIN:g} K and {n['hot']} above {T_MAX:g} K: outside the chemistry "
                "solver's range"
            )
        return n
