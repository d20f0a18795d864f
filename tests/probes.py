"""The harness kernels of tests/device_probes.hip (the solver's guard, the recovery's dense LU and the step-size law,
each alone on given inputs, compiled with a network's header and the solver's kernel files), and the call kernel
launched without the recovery (GPU)."""

import ctypes
from pathlib import Path

import numpy as np
import torch

from maxichem_amd import _hip, _kernels, _tables

PROBE = Path(__file__).resolve().parent / "device_probes.hip"


class Probes:
    """The harness kernels for one network's tables (compiled once), on the current GPU."""

    def __init__(self, net, photo: bool = False, drift: float = 1e-6):
        self.tab = tab = _tables.build(net, photo)
        src = _kernels.source(tab, 1, drift) + "\n" + PROBE.read_text()
        self.module = _hip.Module(_hip.compile_source(src, "device_probes.hip"), torch.cuda.current_device())
        self.guard = self.module.function("mxc_guard_probe")
        self.dense = self.module.function("mxc_dense_probe")
        self.step = self.module.function("mxc_step_probe")
        host = _kernels.host_tables(tab)  # the solver's own tables, in the kernels' Net field order
        self.tables = [torch.from_numpy(host[k]).cuda() for k in _kernels.NET_FIELDS]
        self.net = torch.tensor([a.data_ptr() for a in self.tables], dtype=torch.int64, device="cuda")

    def close(self) -> None:
        self.module.unload()

    def run_guard(self, x0, xres, done, nok, atol: float = 1e-25):
        """The guard of each row of xres (a result state, network order) against the same row of x0: returns the
        written rows and the guard's bits per row."""
        x0, xres = np.atleast_2d(x0), np.atleast_2d(xres)
        cells = x0.shape[0]
        assert x0.shape == xres.shape == (cells, self.tab.ns)
        d = torch.tensor(np.broadcast_to(np.asarray(done, np.int32), (cells,)).copy(), device="cuda")
        n = torch.tensor(np.broadcast_to(np.asarray(nok, np.float64), (cells,)).copy(), device="cuda")
        a, r = (torch.tensor(v, dtype=torch.float64, device="cuda") for v in (x0, xres))
        out = torch.full_like(a, float("nan"))
        code = torch.full((cells,), -1, dtype=torch.int32, device="cuda")
        p = _hip.ptr
        args = [p(self.net), p(a), p(r), p(d), p(n), p(out), p(code), ctypes.c_double(atol)]
        _hip.launch(self.guard, cells, 64, 0, torch.cuda.current_stream().cuda_stream, args)
        torch.cuda.synchronize()
        return out.cpu().numpy(), code.cpu().numpy()

    def run_dense(self, W, y):
        """The recovery's dense LU and solve of each W [count, NS, NS] with y [count, NS]: returns the factors, the
        original row of each factor row, 1 / U_ii and the solutions."""
        W, y = np.asarray(W, np.float64), np.asarray(y, np.float64)
        count, ns = W.shape[0], self.tab.ns
        assert W.shape == (count, ns, ns) and y.shape == (count, ns)
        w, v = (torch.tensor(np.ascontiguousarray(q), device="cuda") for q in (W, y))
        dinv = torch.empty(count, ns, dtype=torch.float64, device="cuda")
        rows = torch.empty(count, ns, dtype=torch.int32, device="cuda")
        lds = 8 * (ns * ns + 2 * ns) + 4 * ns
        p = _hip.ptr
        _hip.launch(self.dense, count, 64 * _kernels.REC_WAVES, lds, torch.cuda.current_stream().cuda_stream,
                    [p(w), p(dinv), p(rows), p(v)])  # fmt: skip
        torch.cuda.synchronize()
        return w.cpu().numpy(), rows.cpu().numpy(), dinv.cpu().numpy(), v.cpu().numpy()

    def run_step_factor(self, rows):
        """step_factor of each row (en, ok, prej, pen, pha, ha)."""
        a = torch.tensor(np.asarray(rows, np.float64).reshape(-1, 6), device="cuda")
        out = torch.full((a.shape[0],), float("nan"), dtype=torch.float64, device="cuda")
        n = a.shape[0]
        _hip.launch(self.step, -(-n // 64), 64, 0, torch.cuda.current_stream().cuda_stream,
                    [_hip.ptr(a), _hip.ptr(out), ctypes.c_int(n)])  # fmt: skip
        torch.cuda.synchronize()
        return out.cpu().numpy()


def call_kernel_only(solver, x, T, P, dt, h, *, rtol=1e-3, atol=1e-25, max_steps=4000, J=None):
    """The solver's call kernel alone, without the recovery, on numpy inputs: the next step sizes, done flags,
    accepted steps, the guard's redo flags and the tries of each cell."""
    from maxichem_amd.kincuda import WAVE

    x = np.atleast_2d(np.asarray(x, np.float64))
    B = x.shape[0]
    g = lambda a, dt=torch.float64: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")  # noqa: E731
    xd, Td, Pd, dtd = g(x), g(np.broadcast_to(T, (B,))), g(np.broadcast_to(P, (B,))), g(np.broadcast_to(dt, (B,)))
    ho = g(np.broadcast_to(h, (B,)))
    jx = g(J) if solver.nj else None
    xo = torch.empty_like(xd)
    done, status, redo, lst = (torch.zeros(B, dtype=torch.int32, device="cuda") for _ in range(4))
    nok, tries = torch.zeros(B, dtype=torch.float64, device="cuda"), torch.zeros(B, dtype=torch.float64, device="cuda")
    solver.count.zero_()
    p = _hip.ptr
    args = [p(solver._net), p(xd), p(Td), p(Pd), p(dtd), p(ho), p(xo), p(done), p(nok), p(status), p(redo), p(lst),
            p(solver.count), p(tries), p(None), p(jx), ctypes.c_int(B), ctypes.c_double(rtol), ctypes.c_double(atol),
            ctypes.c_int(max_steps), p(None), p(solver._rscratch), p(None)]  # fmt: skip
    _hip.launch(solver._call, max(1, min(B, solver.max_blocks)), WAVE * solver.cpw, solver.lds_call,
                torch.cuda.current_stream().cuda_stream, args)  # fmt: skip
    torch.cuda.synchronize()
    return {"h": ho.cpu().numpy(), "done": done.cpu().numpy().astype(bool), "steps": nok.cpu().numpy(),
            "redo": redo.cpu().numpy(), "tries": tries.cpu().numpy()}  # fmt: skip
