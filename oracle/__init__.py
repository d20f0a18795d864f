"""The CPU FP64 oracle: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, as the checker or the CPU baseline.
The product package maxichem_amd never imports it. It wraps oracle/liboracle.so (oracle.c, built by oracle/Makefile),
a scalar C restatement of the reference solver's arithmetic (maxichem/kernels/common.cuh, call.cuh, guard.cuh,
recover.cuh).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
OC_SLOTS = 12

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


class _Net(ctypes.Structure):
    _fields_ = [
        ("ns", ctypes.c_int), ("nr", ctypes.c_int), ("nrt", ctypes.c_int), ("nj", ctypes.c_int),
        ("nel", ctypes.c_int), ("perm", _i), ("rx", _i), ("rc", _d), ("rf", _i), ("nasa", _d), ("atoms", _d),
        ("drift", ctypes.c_double),
    ]  # fmt: skip


def build() -> Path:
    """Compile liboracle.so with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.oracle_call.restype = ctypes.c_int
        _lib.oracle_call_lm.restype = ctypes.c_int
        _lib.oracle_constant.restype = ctypes.c_double
        _lib.oracle_constant.argtypes = [ctypes.c_char_p]
        _lib.oracle_step_factor.restype = ctypes.c_double
        _lib.oracle_step_factor.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double]  # fmt: skip
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Oracle:
    """The oracle over one network's tables (maxichem_amd._tables.Tables) and a drift allowance."""

    def __init__(self, tab, drift: float = 1e-6) -> None:
        self.tab = tab
        self._keep = [
            np.ascontiguousarray(tab.perm, np.int32),
            np.ascontiguousarray(tab.rx, np.int32),
            np.ascontiguousarray(tab.rc, np.float64),
            np.ascontiguousarray(tab.rf, np.int32),
            np.ascontiguousarray(tab.nasa, np.float64),
            np.ascontiguousarray(tab.atoms, np.float64),
        ]
        perm, rx, rc, rf, nasa, atoms = self._keep
        self.net = _Net(
            tab.ns, tab.nr, tab.nrt, tab.nj, tab.nel, _p(perm, _i), _p(rx, _i), _p(rc, _d), _p(rf, _i), _p(nasa, _d),
            _p(atoms, _d), float(drift),
        )  # fmt: skip

    def call(self, x, T, P, dt, h, *, rtol=1e-3, atol=1e-25, max_steps=4000, J=None, threads=0, lm=None):
        """One call of the solver on the CPU: dict of x, h, done, steps, status, flag, tries and count [12]. lm: the
        cells' log rate multipliers [B, nr] (the tangent kernels' primal), or None."""
        x = np.ascontiguousarray(x, np.float64)
        B = x.shape[0]
        T, P, dt, h = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (B,))) for a in (T, P, dt, h))
        Jp = None
        if self.tab.nj:
            J = np.zeros((B, self.tab.nj)) if J is None else J
            J = np.ascontiguousarray(J, np.float64)
            Jp = _p(J, _d)
        out = {
            "x": np.empty_like(x), "h": np.empty(B), "done": np.empty(B, np.int32), "steps": np.empty(B),
            "status": np.empty(B, np.int32), "flag": np.empty(B, np.int32), "tries": np.empty(B),
            "count": np.zeros(OC_SLOTS, np.int64),
        }  # fmt: skip
        lmp = None
        if lm is not None:
            lm = np.ascontiguousarray(lm, np.float64)
            assert lm.shape == (B, self.tab.nr)
            lmp = _p(lm, _d)
        lib().oracle_call_lm(
            ctypes.byref(self.net), ctypes.c_int(B), _p(x, _d), _p(T, _d), _p(P, _d), _p(dt, _d), _p(h, _d), Jp, lmp,
            ctypes.c_double(rtol), ctypes.c_double(atol), ctypes.c_int(max_steps), _p(out["x"], _d),
            _p(out["h"], _d), _p(out["done"], _i), _p(out["steps"], _d), _p(out["status"], _i), _p(out["flag"], _i),
            _p(out["tries"], _d), out["count"].ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
            ctypes.c_int(threads),
        )  # fmt: skip
        out["done"] = out["done"].astype(bool)
        return out

    def rates(self, T: float, P: float, J=None) -> tuple[np.ndarray, np.ndarray]:
        kf, kr = np.empty(self.tab.nr), np.empty(self.tab.nr)
        Jp = None
        if self.tab.nj:
            J = np.zeros(self.tab.nj) if J is None else np.ascontiguousarray(J, np.float64)
            Jp = _p(J, _d)
        lib().oracle_rates(ctypes.byref(self.net), ctypes.c_double(T), ctypes.c_double(P), Jp, _p(kf, _d), _p(kr, _d))
        return kf, kr

    def rhs_jacobian(self, T: float, P: float, N, J=None) -> tuple[np.ndarray, np.ndarray]:
        ns = self.tab.ns
        N = np.ascontiguousarray(N, np.float64)
        F, JN = np.empty(ns), np.empty((ns, ns))
        Jp = None
        if self.tab.nj:
            J = np.zeros(self.tab.nj) if J is None else np.ascontiguousarray(J, np.float64)
            Jp = _p(J, _d)
        lib().oracle_rhs_jacobian(
            ctypes.byref(self.net), ctypes.c_double(T), ctypes.c_double(P), Jp, _p(N, _d), _p(F, _d), _p(JN, _d)
        )
        return F, JN

    def try_loop(self, x0, T: float, P: float, dt: float, h0: float, *, rtol=1e-3, atol=1e-25, max_steps=4000,
                 J=None, dense: bool = False) -> dict:  # fmt: skip
        """One cell's tries without the guard and the recovery, as the call kernel takes them or (dense) as one
        recovery pass does: the next step size h, the time reached t, the accepted steps and the tries."""
        x0 = np.ascontiguousarray(x0, np.float64)
        Jp = None
        if self.tab.nj:
            J = np.zeros(self.tab.nj) if J is None else np.ascontiguousarray(J, np.float64)
            Jp = _p(J, _d)
        h, t, nok, tries = (ctypes.c_double() for _ in range(4))
        lib().oracle_try_loop(
            ctypes.byref(self.net), _p(x0, _d), ctypes.c_double(T), ctypes.c_double(P), Jp, ctypes.c_double(dt),
            ctypes.c_double(h0), ctypes.c_double(rtol), ctypes.c_double(atol), ctypes.c_int(max_steps),
            ctypes.c_int(int(dense)), ctypes.byref(h), ctypes.byref(t), ctypes.byref(nok), ctypes.byref(tries),
        )  # fmt: skip
        return {"h": h.value, "t": t.value, "steps": nok.value, "tries": tries.value}

    def accepted_steps(self, x0, T: float, P: float, dt: float, h0: float, *, rtol=1e-3, atol=1e-25,
                       max_steps=4000, J=None) -> np.ndarray:  # fmt: skip
        """The sizes of one cell's accepted tries in a call (the call kernel's tries, without the guard)."""
        x0 = np.ascontiguousarray(x0, np.float64)
        Jp = None
        if self.tab.nj:
            J = np.zeros(self.tab.nj) if J is None else np.ascontiguousarray(J, np.float64)
            Jp = _p(J, _d)
        steps = np.empty(max_steps)
        n = lib().oracle_accepted_steps(
            ctypes.byref(self.net), _p(x0, _d), ctypes.c_double(T), ctypes.c_double(P), Jp, ctypes.c_double(dt),
            ctypes.c_double(h0), ctypes.c_double(rtol), ctypes.c_double(atol), ctypes.c_int(max_steps),
            _p(steps, _d), ctypes.c_int(max_steps),
        )  # fmt: skip
        return steps[:n].copy()

    def replay(self, x0, T: float, P: float, steps, *, atol=1e-25, J=None) -> tuple[int, np.ndarray]:
        """The frozen primal: from x0, the given tries each accepted, then the guard (bits, x out)."""
        x0 = np.ascontiguousarray(x0, np.float64)
        steps = np.ascontiguousarray(steps, np.float64)
        Jp = None
        if self.tab.nj:
            J = np.zeros(self.tab.nj) if J is None else np.ascontiguousarray(J, np.float64)
            Jp = _p(J, _d)
        out = np.empty(self.tab.ns)
        code = lib().oracle_replay(
            ctypes.byref(self.net), _p(x0, _d), ctypes.c_double(T), ctypes.c_double(P), Jp, _p(steps, _d),
            ctypes.c_int(len(steps)), ctypes.c_double(atol), _p(out, _d),
        )  # fmt: skip
        return code, out

    def check_cell(self, x0, xout, done: bool, nok: float, atol: float = 1e-25) -> tuple[int, np.ndarray]:
        x0 = np.ascontiguousarray(x0, np.float64)
        xo = np.array(xout, np.float64)
        code = lib().oracle_check_cell(
            ctypes.byref(self.net), _p(x0, _d), _p(xo, _d), ctypes.c_int(int(done)), ctypes.c_double(nok),
            ctypes.c_double(atol),
        )  # fmt: skip
        return code, xo


def threshold_lu(a: np.ndarray, y: np.ndarray | None = None):
    """The recovery's threshold-pivoted LU (row order, factors, 1 / U_ii) and the solve of y's rows."""
    a = np.array(a, np.float64)
    n = a.shape[0]
    pr, dinv = np.empty(n, np.int32), np.empty(n)
    yy = np.zeros((0, n)) if y is None else np.array(np.atleast_2d(y), np.float64)
    lib().oracle_threshold_lu(
        ctypes.c_int(n), _p(a, _d), _p(pr, _i), _p(dinv, _d), ctypes.c_int(yy.shape[0]), _p(yy, _d)
    )
    return pr, a, dinv, yy


def step_factor(en: float, ok: bool, prej: bool, pen: float, pha: float, ha: float) -> float:
    """The oracle's step-size law: the factor on ha for the next try."""
    return float(lib().oracle_step_factor(en, int(ok), int(prej), pen, pha, ha))


def constant(name: str) -> float:
    return float(lib().oracle_constant(name.encode()))


def cores() -> int:
    return len(os.sched_getaffinity(0))
