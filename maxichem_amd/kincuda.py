"""CudaSolver: the batched Rodas3 chemistry step on one MI355X, a drop-in for the reference's CudaSolver.

The call, its arguments, results, counts and refusals are the reference's (maxichem/kincuda.py:72-361); the kernels
are this package's HIP kernels (kernels/*.hip), compiled at run time by hiprtc for gfx950 and specialised to the
network, launched through the C ABI of libmaxichem_amd.so. The name keeps the reference's module and class names so a
caller's imports change only in the package name.
"""

from __future__ import annotations

import ctypes
from typing import NamedTuple

import torch

from . import _hip, _kernels, _tables
from ._constants import (
    ATOL,
    COUNTS,
    DEFAULT_MAX_STEPS,
    DEFAULT_RTOL,
    DRIFT,
    REC_STEPS,
    T_MAX,
    T_MIN,
    WAVE,
    X_SUM_TOL,
)
from ._solver import Solver, check_inputs, check_photolysis

__all__ = [
    "ATOL", "COUNTS", "DRIFT", "REC_STEPS", "T_MAX", "T_MIN", "X_SUM_TOL", "CallGraph", "CallResult", "CudaSolver",
    "LaunchResult",
]  # fmt: skip


class CallResult(NamedTuple):
    """Per cell: x [B, S] mole fractions, h [B] next step size (s), done [B] bool, steps [B] accepted steps."""

    x: torch.Tensor
    h: torch.Tensor
    done: torch.Tensor
    steps: torch.Tensor


class LaunchResult(NamedTuple):
    """CallResult's fields and flag [B] int32: 0 passed the guard, 1 recomputed and passes, 2 still fails."""

    x: torch.Tensor
    h: torch.Tensor
    done: torch.Tensor
    steps: torch.Tensor
    flag: torch.Tensor


class CudaSolver(Solver):
    """Batched chemistry solver on one GPU; calls run one at a time on torch's current stream.

    Parameters
    ----------
    net : VulcanNetwork
    device : torch.device or str
    drift : float, optional
        The guard's element-drift allowance (DRIFT if None). Keyword only, as photo.
    photo : bool
        The network's photolysis branches enter the call, which then takes J [B, branches].
    variant : (int, int[, int[, int]]), optional
        The call kernel's cells (wavefronts) per workgroup, stream batch, register budget in wavefronts per SIMD
        (1, or 2 for at most 256 registers with the excess spilled) and the rate coefficients' place (0: LDS, 1: a
        global scratch slot per wavefront); by default the autotuner's stored choice for this network and device,
        else (1, 16, 1, 0).
    tune : bool
        Run the autotuner now (maxichem_amd.tune) and store its choice.
    precision : str
        "fp64" (the method as the reference computes it) or "mixed": the trailing block of M0 factored and solved in
        FP32, each solve refined with FP64 residuals (at most 3 sweeps); rates, right-hand sides, error estimate,
        guard and recovery stay FP64.
    """

    _EXTRA_KERNELS: tuple[str, ...] = ()  # kernel files compiled into the module after the call's (a subclass's)

    def __init__(self, net, device, *, drift: float | None = None, photo: bool = False,
                 variant: tuple[int, ...] | None = None, tune: bool = False, precision: str = "fp64") -> None:  # fmt: skip
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision: 'fp64' or 'mixed' expected, got {precision!r}")
        super().__init__(net, device, drift, photo)
        self.precision = precision
        with torch.cuda.device(self.dev):
            if variant is None:
                from . import tune as _tune

                variant = _tune.choose(net, self.dev, photo, self.drift, run=tune, mixed=precision == "mixed")
            self._setup(net, photo, variant)

    def _setup(self, net, photo: bool, variant: tuple[int, ...]) -> None:
        tab = _tables.build(net, photo)
        self.tab = tab
        self._names = (list(net.species), list(net.elements))
        self.ns, self.nr, self.nj = tab.ns, tab.nr, tab.nj
        v = tuple(int(q) for q in variant)
        # cells (wavefronts) per workgroup, stream batch, register budget in wavefronts per SIMD, rates in global
        # scratch (0: in LDS)
        self.cpw, self.batch, self.wps, self.rates_global = (v + (1, 0)[len(v) - 2 :])[:4] if 2 <= len(v) <= 4 else (0,) * 4
        if (self.cpw not in (1, 2, 4, 8) or self.batch not in (2, 4, 8, 16) or self.wps not in (1, 2)
                or self.rates_global not in (0, 1)):  # fmt: skip
            raise ValueError(f"variant: cells per workgroup 1, 2, 4 or 8, a batch of 2-16, 1 or 2 wavefronts per SIMD "
                             f"and rates in global scratch 0 or 1 expected, got {variant}")  # fmt: skip
        dev = self.dev.index
        info = _hip.device_info(dev)
        mixed = self.precision == "mixed"
        self.lds_call, self.lds_rec = _kernels.lds_bytes(tab, self.cpw, mixed, bool(self.rates_global), self.wps,
                                                         self._EXTRA_KERNELS)  # fmt: skip
        for what, need in (("call", self.lds_call), ("recovery", self.lds_rec)):
            if need > info["lds_per_block"]:
                raise ValueError(
                    f"one cell's {what} kernel needs {need} B of LDS per workgroup; the device has "
                    f"{info['lds_per_block']}"
                )
        self._source = _kernels.source(tab, self.cpw, self.drift, self.batch, mixed, self.wps, bool(self.rates_global),
                                       self._EXTRA_KERNELS)  # fmt: skip
        self._code = code = _hip.compile_source(self._source)
        self._module = _hip.Module(code, dev)
        self._call = self._module.function("mxc_call")
        self._recover = self._module.function("mxc_recover")
        cus = info["compute_units"]
        self.max_blocks = cus * max(1, _hip.occupancy(self._call, WAVE * self.cpw, self.lds_call))
        self.rec_blocks = cus * max(1, _hip.occupancy(self._recover, WAVE * _kernels.REC_WAVES, self.lds_rec))
        # each call wavefront's and each recovery workgroup's slot of rate coefficients, when they are not in LDS
        slots = max(self.max_blocks * self.cpw, self.rec_blocks)
        self._rscratch = (torch.empty(slots * 2 * tab.nr, dtype=torch.float64, device=self.dev)
                          if self.rates_global else None)  # fmt: skip
        self._tables = {k: torch.from_numpy(v).to(self.dev) for k, v in _kernels.host_tables(tab).items()}
        # the kernels' Net struct: the tables' device pointers, in its field order
        self._net = torch.tensor([self._tables[k].data_ptr() for k in _kernels.NET_FIELDS], dtype=torch.int64,
                                 device=self.dev)  # fmt: skip
        self.tries: torch.Tensor | None = None
        self._fp64: torch.Tensor | None = None  # the mixed form's cells that fell back to FP64 (_fp64_cells)
        # a list to collect (start, end) events of each call kernel launch on the call's stream; None: not recorded
        # (when recorded, timing_recover collects the recovery kernel's (start, end) events the same way)
        self.timing: list | None = None
        self.timing_recover: list = []
        self.info = {
            "ns": tab.ns, "nr": tab.nr, "nj": tab.nj, "k0": tab.k0, "nt": tab.nt, "ndv": tab.ndv,
            "jac_terms": tab.jac_terms, "sv_terms": tab.sv_terms, "lds_call": self.lds_call, "lds_rec": self.lds_rec,
            "regs": _hip.attribute(self._call, _hip.ATTR_REGISTERS),
            "words_in_lds": not self._EXTRA_KERNELS and _kernels.words_in_lds(
                tab, self.cpw, self.precision == "mixed", self.wps, bool(self.rates_global)),
            "private_bytes": _hip.attribute(self._call, _hip.ATTR_PRIVATE_BYTES),
            "cells_per_cu": self.max_blocks // cus * self.cpw, "max_blocks": self.max_blocks,
            "rec_blocks": self.rec_blocks, "compute_units": cus, "drift": self.drift, "tables": tab.digest(),
            "variant": (self.cpw, self.batch, self.wps, self.rates_global), "precision": self.precision,
        }  # fmt: skip

    def launch(
        self,
        x: torch.Tensor,
        T: torch.Tensor,
        P: torch.Tensor,
        dt: torch.Tensor,
        h: torch.Tensor,
        *,
        rtol: float = DEFAULT_RTOL,
        atol: float = ATOL,
        max_steps: int = DEFAULT_MAX_STEPS,
        J: torch.Tensor | None = None,
    ) -> LaunchResult:
        """The call, the guard and the recovery; reads the counts back (the one host sync) and refuses bad input
        and T out of range (maxichem/kincuda.py:230-320)."""
        self.flag = None
        self._require_open()
        B = check_inputs(self.dev, self.ns, x, T, P, dt, h, rtol, atol, max_steps)
        with torch.cuda.device(self.dev):
            x, T, P, dt, h = (a.contiguous() for a in (x, T, P, dt, h))
            jx = check_photolysis(J, B, self.nj, self.dev)
            xo, ho = torch.empty_like(x), h.clone()
            done, redo, lst, status = (torch.empty(B, dtype=torch.int32, device=self.dev) for _ in range(4))
            nok = torch.empty(B, dtype=torch.float64, device=self.dev)
            order, xprop, planned = self._plan(x, T, P, dt, ho, jx, B)
            self.tries = torch.zeros(B, dtype=torch.float64, device=self.dev)
            self.count.zero_()
            if planned is not None:  # counts the plan took from the inputs before changing them
                self.count += planned
            stream = torch.cuda.current_stream(self.dev).cuda_stream
            self._launch_all(x, T, P, dt, h, xo, ho, done, nok, status, redo, lst, jx, B, rtol, atol, max_steps,
                             stream, order, xprop)  # fmt: skip
            self.total += self.count
            self._refuse(x, T, P, dt, jx, status)
        self.flag = redo
        return LaunchResult(xo, ho, done.bool(), nok, redo)

    def _plan(self, x, T, P, dt, h_start, jx, B):
        """The queue's order and proposals of a call; h_start (the first step sizes) may be changed in place.

        The plain form queues the previous call's costliest cells first when it had as many cells (a call usually
        follows one over the same cells) and proposes nothing. Results do not depend on the order. The third value
        is None, or counts [COUNT_SLOTS] (int32, on the device) the plan took from the inputs it changed, which the
        call adds to its own."""
        prev = self.tries
        order = None
        if prev is not None and prev.shape[0] == B:
            order = torch.argsort(prev, descending=True).to(torch.int32)
        return order, None, None

    def _launch_all(self, x, T, P, dt, h, xo, ho, done, nok, status, redo, lst, jx, B, rtol, atol, max_steps,
                    stream, order, xprop=None, fp64=None) -> None:  # fmt: skip
        """The call kernel (which guards each cell's result) and the recovery kernel on stream. ho holds each cell's
        first step size on entry (the recovery starts from h); xprop, when given, the proposed end states [B, NS]
        (NaN first entry: none); fp64, when given, the mixed form's per-cell FP64 flags to use instead of the
        solver's own (a captured call owns its flags)."""
        p = _hip.ptr
        net = p(self._net)
        call_args = [net, p(x), p(T), p(P), p(dt), p(ho), p(xo), p(done), p(nok), p(status), p(redo), p(lst),
                     p(self.count), p(self.tries), p(order), p(jx), ctypes.c_int(B), ctypes.c_double(rtol),
                     ctypes.c_double(atol), ctypes.c_int(max_steps), p(xprop), p(self._rscratch),
                     p(self._fp64_cells(B) if fp64 is None else fp64)]  # fmt: skip
        grid = max(1, min(-(-B // self.cpw), self.max_blocks))
        if self.timing is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        _hip.launch(self._call, grid, WAVE * self.cpw, self.lds_call, stream, call_args)
        if self.timing is not None:
            ev[1].record()
            self.timing.append(ev)
        rec_args = [net, p(x), p(T), p(P), p(dt), p(h), p(xo), p(ho), p(done), p(nok), p(redo), p(lst), p(self.count),
                    p(self.tries), p(jx), ctypes.c_double(rtol), ctypes.c_double(atol),
                    ctypes.c_int(REC_STEPS * max_steps), p(self._rscratch)]  # fmt: skip
        _hip.launch(self._recover, max(1, min(B, self.rec_blocks)), WAVE * _kernels.REC_WAVES, self.lds_rec, stream,
                    rec_args)
        if self.timing is not None:  # the recovery kernel's events: from the call kernel's end to its own
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            self.timing_recover.append((ev[1], end))

    def _fp64_cells(self, B: int) -> torch.Tensor | None:
        """The mixed form's per-cell flags [B] (int32): the cells whose FP32 refinement failed in an earlier call with
        as many cells, which then start in FP64 (a call usually repeats the previous call's cells, as for the queue's
        order); None in the FP64 form."""
        if self.precision != "mixed":
            return None
        if self._fp64 is None or self._fp64.shape[0] != B:
            self._fp64 = torch.zeros(B, dtype=torch.int32, device=self.dev)
        return self._fp64

    def close(self) -> None:
        """Unload the kernels and drop the tables and the scratch slots; the counts stay readable."""
        super().close()
        self._rscratch = None
        self._fp64 = None

    def export(self, folder) -> None:
        """Write what a caller without Python needs to run this solver's kernels (maxichem/kincuda.py:363-373): the
        code object, its source, the tables and meta.txt, which mxc_solver_open and maxichem::Solver
        (include/maxichem.hpp) load. The code object is for gfx950."""
        self._require_open()
        from . import _export

        _export.write(folder, tab=self.tab, source=self._source, code=self._code, species=self._names[0],
                      elements=self._names[1], cpw=self.cpw, batch=self.batch, wps=self.wps,
                      rates_global=bool(self.rates_global), precision=self.precision, drift=self.drift)  # fmt: skip

    def capture(self, x, T, P, dt, h, *, rtol: float = DEFAULT_RTOL, atol: float = ATOL,
                max_steps: int = DEFAULT_MAX_STEPS, J=None) -> "CallGraph":  # fmt: skip
        """This call on these tensors captured in a HIP graph (CallGraph): replay() repeats it on their contents."""
        return CallGraph(self, x, T, P, dt, h, rtol=rtol, atol=atol, max_steps=max_steps, J=J)

    def __call__(
        self,
        x: torch.Tensor,
        T: torch.Tensor,
        P: torch.Tensor,
        dt: torch.Tensor,
        h: torch.Tensor,
        *,
        rtol: float = DEFAULT_RTOL,
        atol: float = ATOL,
        max_steps: int = DEFAULT_MAX_STEPS,
        J: torch.Tensor | None = None,
    ) -> CallResult:
        """Advance each cell by dt: launch without the flag in its result."""
        r = self.launch(x, T, P, dt, h, rtol=rtol, atol=atol, max_steps=max_steps, J=J)
        return CallResult(r.x, r.h, r.done, r.steps)


class CallGraph:
    """One call of a solver captured in a HIP graph on fixed tensors (a GCM's time loop calls with the same tensors
    every step): replay() runs the call kernel (with its guard) and the recovery again on what the input tensors
    hold then, with one launch, and returns the graph's own output tensors, which the next replay overwrites. The
    graph owns every buffer its kernels address, the mixed form's FP64 flags included. The counts, the
    refusals and check() behave as in an eager call; the queue is ordered by the previous replay's tries, set on the
    host side before each replay. Building it runs the call once, off the graph, without adding to the counts."""

    def __init__(self, solver: CudaSolver, x, T, P, dt, h, *, rtol: float = DEFAULT_RTOL, atol: float = ATOL,
                 max_steps: int = DEFAULT_MAX_STEPS, J=None) -> None:  # fmt: skip
        solver._require_open()
        self.solver = solver
        dev = solver.dev
        self.B = B = check_inputs(dev, solver.ns, x, T, P, dt, h, rtol, atol, max_steps)
        for name, a in (("x", x), ("T", T), ("P", P), ("dt", dt), ("h", h)):
            if not a.is_contiguous():
                raise ValueError(f"{name}: a contiguous tensor expected for a captured call")
        self.jx = check_photolysis(J, B, solver.nj, dev)
        if J is not None and self.jx is not None and self.jx.data_ptr() != J.data_ptr():
            raise ValueError("J: a contiguous tensor expected for a captured call")
        self.inputs = (x, T, P, dt, h)
        self.settings = (rtol, atol, max_steps)
        with torch.cuda.device(dev):
            self.xo, self.ho = torch.empty_like(x), torch.empty_like(h)
            self.done, self.redo, self.lst, self.status = (torch.empty(B, dtype=torch.int32, device=dev)
                                                           for _ in range(4))  # fmt: skip
            self.nok = torch.empty(B, dtype=torch.float64, device=dev)
            self.tries = torch.zeros(B, dtype=torch.float64, device=dev)
            self.order = torch.arange(B, dtype=torch.int32, device=dev)
            # the mixed form's per-cell FP64 flags, owned by the graph: its kernels hold their address
            self.fp64 = torch.zeros(B, dtype=torch.int32, device=dev) if solver.precision == "mixed" else None
            self._ordered = False
            self.graph = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                self._device_part(count=False)  # once eagerly, off the graph: the kernels' first launch
            torch.cuda.current_stream(dev).wait_stream(side)
            torch.cuda.synchronize(dev)
            with torch.cuda.graph(self.graph):
                self._device_part()

    def _device_part(self, count: bool = True) -> None:
        s = self.solver
        x, T, P, dt, h = self.inputs
        rtol, atol, max_steps = self.settings
        self.ho.copy_(h)
        self.tries.zero_()  # fill kernels: no memset node in the graph
        s.count.zero_()
        saved, s.tries = s.tries, self.tries
        stream = torch.cuda.current_stream(s.dev).cuda_stream
        s._launch_all(x, T, P, dt, h, self.xo, self.ho, self.done, self.nok, self.status, self.redo, self.lst,
                      self.jx, self.B, rtol, atol, max_steps, stream, self.order, fp64=self.fp64)  # fmt: skip
        s.tries = saved
        if count:
            s.total += s.count

    def replay(self) -> LaunchResult:
        s = self.solver
        s._require_open()
        s.flag = None
        with torch.cuda.device(s.dev):
            if self._ordered:  # the costliest cells of the previous replay first
                self.order.copy_(torch.argsort(self.tries, descending=True).to(torch.int32))
            self.graph.replay()
            self._ordered = True
            x, T, P, dt, _ = self.inputs
            s.tries = self.tries
            s._refuse(x, T, P, dt, self.jx, self.status)
        s.flag = self.redo
        return LaunchResult(self.xo, self.ho, self.done.bool(), self.nok, self.redo)
