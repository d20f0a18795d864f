"""The tangent (JVP) and adjoint (VJP) of one solver call on the GPU: ADSolver, with the reference's call forms
(maxichem/kincuda_ad.py: ADSolver.jvp, ADSolver.vjp, JvpResult, VjpResult, Cotangents; docs/method.md,
"Derivatives").

Built: both of the reference's contracts. adapt=True (its default) differentiates the step control: every try's
tangent, rejected ones included, the error's and the step factor's, so the tangents of dt and h enter and the
result's h has one. adapt=False (the frozen contract) holds the accepted step sizes fixed: rejected tries drop out, no
tangent through dt or h, and none of h out. The tangents of x, T, P and J enter under both. The call's result is
CudaSolver.launch's, bit for bit (the same kernels' arithmetic in the same order). A cell the guard fails is
recomputed by the recovery with its tangent (mxc_jvp_recover, on the recovery's dense factors), its flag 1 or 2 as in
launch; a cell that still fails (flag 2) has NaN tangents.

Beside them, both contracts take per-cell log rate multipliers lm [B, R] (a reaction's two coefficients times exp(lm)
after the collision limit, a photolysis branch's kf = J exp(lm)) and another rate table rc [R, W] for the call, with
their tangents dlm and drc (the nine Arrhenius columns of rc).

The adjoint (vjp) of the same map is assembled from the tangent's columns, one tangent call per input coordinate; the
rate table's cotangent follows from the multipliers' (a multiplier and the table's law enter a reaction's
coefficients alike) through one reduction kernel.
"""

from __future__ import annotations

import ctypes
from typing import NamedTuple

import torch

from . import _hip, _kernels
from ._constants import ATOL, DEFAULT_MAX_STEPS, DEFAULT_RTOL, REC_STEPS, WAVE
from ._solver import check_inputs, check_photolysis, check_tensor
from .kincuda import CudaSolver

__all__ = ["ADSolver", "Cotangents", "JvpResult", "VjpResult"]
DMAX = 64  # directions of one jvp call (the reference's bound on the leading axis)
ARRHENIUS_COLUMNS = 9  # ln A, B, C of k0, of the second k0 and of k_inf: the differentiated columns of rc
COLUMN_BYTES = 1 << 31  # device memory for one launch's unit tangents and outputs in vjp


class JvpResult(NamedTuple):
    """What ADSolver.jvp returns, per cell (the reference's JvpResult): x, h, done, steps as CudaSolver's call;
    dx [B, S] or [D, B, S] the tangent of x; dh [B] or [D, B] the tangent of h (zero under the frozen contract)."""

    x: torch.Tensor
    h: torch.Tensor
    done: torch.Tensor
    steps: torch.Tensor
    dx: torch.Tensor
    dh: torch.Tensor


class Cotangents(NamedTuple):
    """The cotangents of ADSolver.vjp's inputs (the reference's Cotangents), float64, each per unit of its input: x
    [B, S]; T, P, dt, h [B]; lm [B, R] and rc [R, 9] (None when vjp is asked not to build them); J [B, branches]
    (None without photolysis)."""

    x: torch.Tensor
    T: torch.Tensor
    P: torch.Tensor
    dt: torch.Tensor
    h: torch.Tensor
    lm: torch.Tensor | None
    rc: torch.Tensor | None
    J: torch.Tensor | None


class VjpResult(NamedTuple):
    """What ADSolver.vjp returns (the reference's VjpResult): x, h, done, steps as CudaSolver's call, and the inputs'
    cotangents."""

    x: torch.Tensor
    h: torch.Tensor
    done: torch.Tensor
    steps: torch.Tensor
    cotangents: Cotangents


class ADSolver(CudaSolver):
    """One solver call with its tangent on one GPU (the reference's ADSolver(net, device, *, cap, drift, photo)).

    cap, the reference's stored attempts of its adjoint, is accepted and checked (an integer from 1) for the same
    construction; the tangent does not store attempts."""

    _EXTRA_KERNELS = ("jvp.hip",)

    def __init__(self, net, device, *, cap: int = 64, drift: float | None = None, photo: bool = False) -> None:
        if not isinstance(cap, int) or isinstance(cap, bool):
            raise ValueError(f"cap: an integer expected, got {cap!r}")
        if cap < 1:
            raise ValueError(f"cap: at least 1 stored attempt expected, got {cap}")
        # the rates in global scratch: a tangent cell then needs 40 KB of LDS instead of 54 (NCHO), four cells per
        # compute unit instead of three (every SIMD busy); the call kernel of this module gives the same bits
        super().__init__(net, device, drift=drift, photo=photo, variant=(1, _kernels.DEFAULT_BATCH, 1, 1))
        self.cap = cap
        with torch.cuda.device(self.dev):
            self._jvp = self._module.function("mxc_jvp")
            self.lds_jvp = _kernels.jvp_lds_bytes(self.tab, 1, True)
            info = _hip.device_info(self.dev.index)
            if self.lds_jvp > info["lds_per_block"]:
                raise ValueError(f"one cell's tangent kernel needs {self.lds_jvp} B of LDS; the device has "
                                 f"{info['lds_per_block']}")  # fmt: skip
            self.max_blocks_jvp = info["compute_units"] * max(1, _hip.occupancy(self._jvp, WAVE, self.lds_jvp))
            self._rc_cotangent = self._module.function("mxc_rc_cotangent")
            self._jvp_recover = self._module.function("mxc_jvp_recover")
            self.lds_jvp_rec = _kernels.jvp_recover_lds_bytes(self.tab, True)
            if self.lds_jvp_rec > info["lds_per_block"]:
                raise ValueError(f"one cell's tangent recovery needs {self.lds_jvp_rec} B of LDS; the device has "
                                 f"{info['lds_per_block']}")  # fmt: skip
            self.rec_blocks_jvp = info["compute_units"] * max(
                1, _hip.occupancy(self._jvp_recover, WAVE * _kernels.REC_WAVES, self.lds_jvp_rec))
            # each tangent wavefront's and tangent recovery workgroup's slot of rates and rate tangents
            self._jrsc = torch.empty(max(self.max_blocks_jvp, self.rec_blocks_jvp) * _kernels.jvp_slot(self.tab),
                                     dtype=torch.float64, device=self.dev)  # fmt: skip
        # the rate table's width as the reference gives it: the nine Arrhenius columns, and the Troe parameters and
        # the efficiency when a reaction has them
        self.rcw = self.tab.rc.shape[1] if self.tab.has_troe or self.tab.has_eff else ARRHENIUS_COLUMNS

    def close(self) -> None:
        """As CudaSolver.close, and drop the tangent kernels' slots of rates."""
        super().close()
        self._jrsc = None

    @property
    def rc0(self) -> torch.Tensor:
        """The network's rate table [R, W] (the reference's ADSolver.rc0), which a call uses unless given another: per
        reaction ln A, B, C of k0, of the second k0 and of k_inf, then (W = 14) Troe a, T3, T1, T2 and the efficiency;
        photolysis rows are 0."""
        self._require_open()
        return self._tables["rc"][:, : self.rcw]

    def _rates(self, B: int, lm, rc):
        """Check lm [B, R] and rc [R, W]; return lm contiguous (or None), the Net struct of the call (the solver's, or a
        copy that points at rc widened to the kernels' 14 columns) and the widened table, which the caller keeps until
        the call's kernels have been launched."""
        if lm is not None:
            check_tensor("lm", lm, (B, self.nr), self.dev)
            lm = lm.contiguous()
        if rc is None:
            return lm, self._net, None
        check_tensor("rc", rc, (self.nr, self.rcw), self.dev)
        full = rc if self.rcw == self.tab.rc.shape[1] else torch.cat([rc, self._tables["rc"][:, self.rcw:]], 1)
        full = full.contiguous()
        net = self._net.clone()
        net[_kernels.NET_FIELDS.index("rc")] = full.data_ptr()
        return lm, net, full

    def jvp(self, x, T, P, dt, h, dx, dT, dP, ddt, dh, *, lm=None, dlm=None, rc=None, drc=None, adapt: bool = True,
            rtol: float = DEFAULT_RTOL, atol: float = ATOL, max_steps: int = DEFAULT_MAX_STEPS, J=None,
            dJ=None) -> JvpResult:  # fmt: skip
        """The call and its tangent. Everything after dh is keyword only.

        dx, dT, dP, ddt, dh: the inputs' tangents in their shapes, or with a leading axis of D <= 64 directions on
        every one, which then share one launch (each direction repeats the call with the same bits); dJ [B, branches]
        (or [D, B, branches]) with photolysis. adapt=True (the reference's default)
        differentiates the step control too: every try's tangent, the error's and the step factor's, so ddt and dh
        enter and the result's h has a tangent; adapt=False holds the accepted steps fixed (ddt and dh do not enter,
        the result's dh is zero). The result's x, h, done and steps are CudaSolver.launch's; dx, dh the tangents of x
        and h (those of a recomputed cell through the recovery's tries; NaN for one that still fails, flag 2).
        Refusals and counts as launch's.

        lm [B, R]: per-cell log rate multipliers (a reaction's two coefficients times exp(lm) after the collision
        limit, a photolysis branch's kf = J exp(lm)); a non-finite one makes its cell bad input. rc [R, W]: the rate
        table the call uses instead of rc0. dlm [B, R] and drc [R, 9] (the nine Arrhenius columns), or with the
        leading axis: their tangents."""
        self.flag = None
        self._require_open()
        B = check_inputs(self.dev, self.ns, x, T, P, dt, h, rtol, atol, max_steps)
        if dx.dim() not in (2, 3):
            raise ValueError(f"dx: [B, S] or [D, B, S] expected, got {list(dx.shape)}")
        one = dx.dim() == 2
        D = 1 if one else dx.shape[0]
        if not 1 <= D <= DMAX:
            raise ValueError(f"dx: from 1 to {DMAX} directions on the leading axis expected, got {list(dx.shape)}")
        lead = () if one else (D,)
        if dJ is not None and not self.nj and (dJ.dim() != len(lead) + 2 or dJ.shape[-1] != 0):
            raise ValueError(f"dJ: none or an empty {[*lead, B, 0]} one expected, since this solver has no photolysis; "
                             f"got {list(dJ.shape)}")  # fmt: skip
        shapes = {"dx": (*lead, B, self.ns), "dT": (*lead, B), "dP": (*lead, B), "ddt": (*lead, B), "dh": (*lead, B),
                  "dJ": (*lead, B, self.nj), "dlm": (*lead, B, self.nr), "drc": (*lead, self.nr, ARRHENIUS_COLUMNS)}
        tangents = (("dx", dx), ("dT", dT), ("dP", dP), ("ddt", ddt), ("dh", dh), ("dJ", dJ), ("dlm", dlm),
                    ("drc", drc))  # fmt: skip
        for name, a in tangents:
            if a is not None:
                check_tensor(name, a, shapes[name], self.dev)
        with torch.cuda.device(self.dev):
            x, T, P, dt, h = (a.contiguous() for a in (x, T, P, dt, h))
            lm, net, table = self._rates(B, lm, rc)
            jx = check_photolysis(J, B, self.nj, self.dev)
            dJ = dJ if self.nj else None
            tx = (dx if not one else dx[None]).contiguous()
            tT, tP, tdt, th, tJ, tlm, trc = (None if a is None else (a if not one else a[None]).contiguous()
                                             for a in (dT, dP, ddt, dh, dJ, dlm, drc))  # fmt: skip
            xd = torch.empty(D, B, self.ns, dtype=torch.float64, device=self.dev)
            hd = torch.empty(D, B, dtype=torch.float64, device=self.dev)
            xo, ho, done, nok, redo = self._jvp_launch(x, T, P, dt, h, jx, rtol, atol, max_steps, B, D, tx, tT, tP, tJ,
                                                       xd, adapt, tdt, th, hd, lm, tlm, trc, net)  # fmt: skip
            del table  # kept to here: the launches above read it
            self.total += self.count
            self._refuse(x, T, P, dt, jx, self._status)
        self.flag = redo
        return JvpResult(xo, ho, done.bool(), nok, xd[0] if one else xd, hd[0] if one else hd)

    def _jvp_launch(self, x, T, P, dt, h, jx, rtol, atol, max_steps, B, D, dx, dT, dP, dJ, xd, adapt=False, ddt=None,
                    dh=None, hd=None, lm=None, dlm=None, drc=None, net=None):  # fmt: skip
        """One launch of the tangent kernel over the D directions' (direction, cell) pairs (each tangent [D, ...]
        contiguous, or None), then the recovery of the listed cells with each direction's tangent. Returns the call's
        x, h, done, steps and flags; the tangents go to xd [D, B, S] and hd [D, B]."""
        xo, ho = torch.empty_like(x), torch.empty_like(h)
        done, redo, lst, status = (torch.empty(B, dtype=torch.int32, device=self.dev) for _ in range(4))
        nok = torch.empty(B, dtype=torch.float64, device=self.dev)
        self.tries = torch.zeros(B, dtype=torch.float64, device=self.dev)
        self.count.zero_()
        p = _hip.ptr
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        net = self._net if net is None else net
        args = [p(net), p(x), p(T), p(P), p(dt), p(h), p(ho), p(xo), p(done), p(nok), p(status), p(redo), p(lst),
                p(self.count), p(self.tries), p(jx), ctypes.c_int(B), ctypes.c_double(rtol), ctypes.c_double(atol),
                ctypes.c_int(max_steps), p(dx), p(dT), p(dP), p(dJ), p(xd), ctypes.c_int(int(bool(adapt))), p(ddt),
                p(dh), p(hd), p(lm), p(dlm), p(drc), ctypes.c_int(D), p(self._jrsc)]  # fmt: skip
        if B:
            _hip.launch(self._jvp, max(1, min(B * D, self.max_blocks_jvp)), WAVE, self.lds_jvp, stream, args)
        # the cells the guard failed: the recovery's result with each direction's tangent. Direction 0's launch keeps
        # the counts and the tries; the others run on a copy of the counts as the tangent kernel left them (the list's
        # length), write the same results again and leave the tries alone
        listed = self.count.clone()

        def row(a, d):
            return None if a is None else a[d]

        for d in range(D):
            count = self.count if d == 0 else listed.clone()
            rec_args = [p(net), p(x), p(T), p(P), p(dt), p(h), p(xo), p(ho), p(done), p(nok), p(redo), p(lst), p(count),
                        p(self.tries if d == 0 else None), p(jx), ctypes.c_double(rtol), ctypes.c_double(atol),
                        ctypes.c_int(REC_STEPS * max_steps), p(lm), p(dx[d]), p(row(dT, d)), p(row(dP, d)),
                        p(row(dJ, d)), p(xd[d]), ctypes.c_int(int(bool(adapt))), p(row(ddt, d)), p(row(dh, d)),
                        p(hd[d]), p(row(dlm, d)), p(row(drc, d)), p(self._jrsc)]  # fmt: skip
            _hip.launch(self._jvp_recover, max(1, min(B, self.rec_blocks_jvp)), WAVE * _kernels.REC_WAVES,
                        self.lds_jvp_rec, stream, rec_args)  # fmt: skip
        self._status = status
        return xo, ho, done, nok, redo

    def _column_chunks(self, columns, B: int):
        """The vjp's columns in launches of at most DMAX directions, each launch's unit tangents within
        COLUMN_BYTES."""
        per = 8 * B * (2 * self.ns + 4 + self.nj + self.nr)  # one direction's tangents and outputs, at most
        size = max(1, min(DMAX, COLUMN_BYTES // max(per, 1)))
        for i in range(0, len(columns), size):
            yield columns[i : i + size]

    def vjp(self, x, T, P, dt, h, gx, *, gh=None, lm=None, rc=None, adapt: bool = True, rtol: float = DEFAULT_RTOL,
            atol: float = ATOL, max_steps: int = DEFAULT_MAX_STEPS, J=None, rates: bool = True) -> VjpResult:  # fmt: skip
        """The call and its inputs' cotangents for the output cotangents gx [B, S] and gh [B] (optional), under
        either contract as jvp (the reference's ADSolver.vjp, maxichem/kincuda_ad.py:457-556). Everything after gx is
        keyword only; lm [B, R] and rc [R, W] as jvp.

        Built from the tangent by columns, not by a reverse pass: one tangent direction per input coordinate (each
        of x's S entries, T, P, each photolysis branch, under adapt=True dt and h, and with rates each thermal
        reaction's multiplier), every cell along the same unit direction, up to 64 directions in one launch, and each
        cell's cotangent gx . dx + gh dh of its column. A photolysis branch's multiplier cotangent is J times J's
        (kf = J exp(lm)); the rate table's is reduced from the multipliers' by mxc_rc_cotangent (a multiplier and the
        law enter a reaction's coefficients alike, the law's share 1 - wA of the collision limit). Its work is S + 2 +
        branches (+ 2) tangent calls, and the thermal reactions' count more with rates (rates=False leaves the
        cotangents of lm and rc None); a small batch runs its columns side by side. Under the
        frozen contract the cotangents of dt and h are zero. A cell that still fails the guard (flag 2) has NaN
        cotangents and is left out of the rate table's sum (fallback["still_failed"] counts them)."""
        self.flag = None
        self._require_open()
        B = check_inputs(self.dev, self.ns, x, T, P, dt, h, rtol, atol, max_steps)
        check_tensor("gx", gx, (B, self.ns), self.dev)
        if gh is not None:
            check_tensor("gh", gh, (B,), self.dev)
        with torch.cuda.device(self.dev):
            x, T, P, dt, h, gx = (a.contiguous() for a in (x, T, P, dt, h, gx))
            gh = None if gh is None else gh.contiguous()
            lm, net, table = self._rates(B, lm, rc)
            jx = check_photolysis(J, B, self.nj, self.dev)
            zero = torch.zeros(B, dtype=torch.float64, device=self.dev)
            g = Cotangents(x=torch.empty_like(x), T=torch.empty_like(T), P=torch.empty_like(T), dt=zero.clone(),
                           h=zero.clone(),
                           lm=torch.empty(B, self.nr, dtype=torch.float64, device=self.dev) if rates else None,
                           rc=torch.empty(self.nr, ARRHENIUS_COLUMNS, dtype=torch.float64, device=self.dev)
                           if rates else None,
                           J=torch.empty(B, self.nj, dtype=torch.float64, device=self.dev) if self.nj else None)  # fmt: skip
            # the input coordinates, each a unit direction for every cell at once: (what, index, where its
            # cotangent goes)
            columns = [("x", s, g.x[:, s]) for s in range(self.ns)] + [("T", 0, g.T), ("P", 0, g.P)]
            columns += [("J", b, g.J[:, b]) for b in range(self.nj)]
            if adapt:  # dt and h enter under the adaptive contract
                columns += [("dt", 0, g.dt), ("h", 0, g.h)]
            if rates:  # the thermal reactions' multipliers
                columns += [("lm", r, g.lm[:, r]) for r in range(self.tab.nrt)]
            result = None
            for chunk in self._column_chunks(columns, B):
                D = len(chunk)
                kinds = {k for k, _, _ in chunk}
                f64 = {"dtype": torch.float64, "device": self.dev}
                dx = torch.zeros(D, B, self.ns, **f64)
                dT, dP, ddt, dh = (torch.zeros(D, B, **f64) if k in kinds else None for k in ("T", "P", "dt", "h"))
                dJ = torch.zeros(D, B, self.nj, **f64) if "J" in kinds else None
                dlm = torch.zeros(D, B, self.nr, **f64) if "lm" in kinds else None
                unit = {"x": dx, "T": dT, "P": dP, "dt": ddt, "h": dh, "J": dJ, "lm": dlm}
                for d, (k, i, _) in enumerate(chunk):
                    if k in ("x", "J", "lm"):
                        unit[k][d, :, i] = 1.0
                    else:
                        unit[k][d] = 1.0
                xd = torch.empty(D, B, self.ns, **f64)
                hd = torch.empty(D, B, **f64)
                result = self._jvp_launch(x, T, P, dt, h, jx, rtol, atol, max_steps, B, D, dx, dT, dP, dJ, xd, adapt,
                                          ddt, dh, hd, lm, dlm, None, net)  # fmt: skip
                out = torch.einsum("dbs,bs->db", xd, gx)
                if gh is not None and adapt:
                    out += hd * gh
                for d, (_, _, dest) in enumerate(chunk):
                    dest.copy_(out[d])
            if rates and self.nj:  # a branch's kf = J exp(lm): d/dlm = J d/dJ
                g.lm[:, self.tab.nrt:] = jx * g.J
            xo, ho, done, nok, redo = result
            if rates:
                skip = (redo == 2).to(torch.int32)  # a cell that still fails has no tangent: left out of the sum
                stream = torch.cuda.current_stream(self.dev).cuda_stream
                _hip.launch(self._rc_cotangent, self.nr, WAVE, 0, stream,
                            [_hip.ptr(net), _hip.ptr(T), _hip.ptr(P), _hip.ptr(g.lm), _hip.ptr(skip), ctypes.c_int(B),
                             _hip.ptr(g.rc)])  # fmt: skip
            del table
            self.total += self.count
            self._refuse(x, T, P, dt, jx, self._status)
        self.flag = redo
        return VjpResult(xo, ho, done.bool(), nok, g)
