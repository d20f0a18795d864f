"""The allocation-free call (mxc_solver_call_static) against mxc_solver_call on the same export, bit for bit: two
calls in a row, the second over fewer cells, so the static work buffers keep the first call's contents when the
second starts; counts and flags included, the refusal of a batch above the limit, and no device allocation by the
call itself (torch's allocator and the device's free memory unchanged across it) (GPU)."""

import ctypes

import numpy as np
import pytest
import torch

from maxichem_amd import _export, _hip, standin, workloads

pytestmark = pytest.mark.gpu
STATIC_CELLS = 262144  # MXC_STATIC_CELLS of include/maxichem_amd.h


def g(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


class Solver:
    def __init__(self, folder):
        self.lib = _hip.lib()
        self.h = ctypes.c_void_p()
        assert self.lib.mxc_solver_open(str(folder).encode(), 0, ctypes.byref(self.h)) == 0, self.err()
        ns, nj, nel, slots = (ctypes.c_int() for _ in range(4))
        assert self.lib.mxc_solver_sizes(self.h, *map(ctypes.byref, (ns, nj, nel, slots))) == 0
        self.slots = slots.value

    def err(self):
        return self.lib.mxc_last_error().decode()

    def call(self, entry, x, T, P, dt, h, J, max_steps=4000):
        """One call through entry: its code, outputs, counts and message, and whether torch's allocator and the
        device's free memory were the same after the entry point returned as just before it was called."""
        B = x.shape[0]
        out = dict(x=torch.empty_like(x), h=torch.empty_like(h), steps=torch.empty_like(h),
                   done=torch.empty(B, dtype=torch.int32, device="cuda"),
                   flag=torch.empty(B, dtype=torch.int32, device="cuda"))  # fmt: skip
        p = _hip.ptr
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        before = torch.cuda.memory_allocated(), torch.cuda.mem_get_info()[0]
        rc = getattr(self.lib, entry)(self.h, stream, B, p(x), p(T), p(P), p(dt), p(h), p(J), 1e-3, 1e-25, max_steps,
                                      p(out["x"]), p(out["h"]), p(out["done"]), p(out["steps"]), p(out["flag"]))  # fmt: skip
        out["message"] = self.err()  # before the counts' entry point clears it
        torch.cuda.synchronize()
        out["unchanged"] = (torch.cuda.memory_allocated(), torch.cuda.mem_get_info()[0]) == before
        c = (ctypes.c_int * self.slots)()
        assert self.lib.mxc_solver_counts(self.h, c, None) == 0
        out["count"] = np.array(c[:])
        return rc, out

    def close(self):
        assert self.lib.mxc_solver_close(self.h) == 0, self.err()


def inputs(net, cells, seed):
    d = workloads.config1(net.species, len(net.photo), cells=cells, seed=seed)
    return [g(d[k]) for k in ("x", "T", "P", "dt", "h", "J")]


def same(a, b):
    return all(np.array_equal(a[k].cpu().numpy().view(np.uint8), b[k].cpu().numpy().view(np.uint8))
               for k in ("x", "h", "steps", "done", "flag")) and np.array_equal(a["count"], b["count"])  # fmt: skip


def test_static_call_is_the_allocating_call(tmp_path):
    if not torch.cuda.is_available():
        pytest.skip("this test needs a GPU")
    net = standin.load("NCHO")
    folder = _export.export_network(net, tmp_path / "export", photo=True)
    a, b = Solver(folder), Solver(folder)
    try:
        for k, (cells, seed, max_steps) in enumerate(((4096, 3, 4000), (1500, 4, 2))):  # then fewer, recovered
            args = inputs(net, cells, seed)
            rc_s, got = b.call("mxc_solver_call_static", *args, max_steps=max_steps)
            rc_a, want = a.call("mxc_solver_call", *args, max_steps=max_steps)
            print(f"{cells} cells, max_steps {max_steps}: codes {rc_s} {rc_a}, counts {got['count'].tolist()}, "
                  f"memory unchanged by the static call {got['unchanged']}")  # fmt: skip
            assert rc_s == rc_a == 0, (got["message"], want["message"])
            assert same(got, want)
            if k:  # the runtime's own first-use set-up is behind us: the static call allocates nothing
                assert got["unchanged"]
        # a refused cell: the same code and message
        x, T, P, dt, h, J = inputs(net, 256, 5)
        T[7] = 5000.0
        rc_s, got = b.call("mxc_solver_call_static", x, T, P, dt, h, J)
        rc_a, want = a.call("mxc_solver_call", x, T, P, dt, h, J)
        assert rc_s == rc_a == 2 and got["message"] == want["message"] and "cell 7" in got["message"]
        assert same(got, want)
        # above the limit: refused before any device work
        big = STATIC_CELLS + 1
        x = torch.zeros(big, len(net.species), dtype=torch.float64, device="cuda")
        v = torch.zeros(big, dtype=torch.float64, device="cuda")
        Jb = torch.zeros(big, len(net.photo), dtype=torch.float64, device="cuda")
        rc, out = b.call("mxc_solver_call_static", x, v, v, v, v, Jb)
        assert rc == 1 and f"at most {STATIC_CELLS} cells" in out["message"]
    finally:
        a.close()
        b.close()
