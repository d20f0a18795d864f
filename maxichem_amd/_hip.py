"""ctypes binding of libmaxichem_amd.so, the C ABI over hiprtc and the HIP module API (include/maxichem_amd.h).

The reference binds the CUDA driver and NVRTC the same way (maxichem/_cuda.py). PyTorch is imported first, so the
library's libamdhip64 and libhiprtc are the ones PyTorch already loaded: one HIP runtime per process. There is no
fallback: without the library, building a solver raises.
"""

from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
from pathlib import Path

import torch  # noqa: F401  (loads the process's HIP runtime before the library)

from ._kernels import ARCH, OPTIONS

HERE = Path(__file__).resolve().parent
LIB = HERE / "libmaxichem_amd.so"
SOURCES = (HERE / "csrc" / "maxichem_hip.cpp", HERE / "csrc" / "solver.cpp")
INCLUDE = HERE.parent / "include"
ABI_VERSION = 3
CACHE = Path(os.environ.get("MAXICHEM_AMD_CACHE", HERE / "_cache"))

ATTR_REGISTERS, ATTR_PRIVATE_BYTES, ATTR_STATIC_LDS, ATTR_MAX_THREADS = 0, 1, 2, 3


def build_library() -> Path:
    """Compile the C ABI library in place with hipcc (host code; gfx950 is the only offload target)."""
    cmd = [
        "/opt/rocm/bin/hipcc", f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{INCLUDE}",
        "-o", str(LIB) + ".tmp", *map(str, SOURCES), "-lhiprtc",
    ]  # fmt: skip
    subprocess.run(cmd, check=True)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


class HipError(RuntimeError):
    """A failure of hiprtc or the HIP runtime, with its message."""


_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise HipError(f"{LIB} is missing: build it (python -c 'import __graft_entry__ as g; g.build()')")
        h = ctypes.CDLL(str(LIB))
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        h.mxc_version.restype = i
        h.mxc_last_error.restype = ctypes.c_char_p
        h.mxc_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), i,
                                  ctypes.POINTER(vp), ctypes.POINTER(sz), ctypes.c_char_p, sz]  # fmt: skip
        h.mxc_free.argtypes = [vp]
        h.mxc_module_load.argtypes = [vp, sz, i, ctypes.POINTER(vp)]
        h.mxc_module_function.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp)]
        h.mxc_module_unload.argtypes = [vp]
        h.mxc_launch.argtypes = [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, vp, ctypes.POINTER(vp)]
        h.mxc_occupancy.argtypes = [vp, i, sz, ctypes.POINTER(i)]
        h.mxc_function_attribute.argtypes = [vp, i, ctypes.POINTER(i)]
        h.mxc_device_info.argtypes = [i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
        h.mxc_compiler_version.argtypes = [ctypes.POINTER(i), ctypes.POINTER(i)]
        ip, d = ctypes.POINTER(i), ctypes.c_double
        h.mxc_solver_open.argtypes = [ctypes.c_char_p, i, ctypes.POINTER(vp)]
        h.mxc_solver_close.argtypes = [vp]
        h.mxc_solver_sizes.argtypes = [vp, ip, ip, ip, ip]
        h.mxc_solver_defaults.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(d), ip]
        h.mxc_solver_name.argtypes = [vp, i, i]
        h.mxc_solver_name.restype = ctypes.c_char_p
        h.mxc_solver_call.argtypes = [vp, vp, i, vp, vp, vp, vp, vp, vp, d, d, i, vp, vp, vp, vp, vp]
        h.mxc_solver_call_static.argtypes = [vp, vp, i, vp, vp, vp, vp, vp, vp, d, d, i, vp, vp, vp, vp, vp]
        h.mxc_solver_counts.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_longlong)]
        h.mxc_solver_atoms.argtypes = [vp, ctypes.POINTER(d)]
        h.mxc_kernel_arguments.argtypes = [ctypes.c_char_p, ctypes.c_char_p, sz]
        if h.mxc_version() != ABI_VERSION:
            raise HipError(f"{LIB} has ABI version {h.mxc_version()}, {ABI_VERSION} expected: rebuild it")
        _lib = h
    return _lib


def _check(code: int, what: str) -> None:
    if code:
        raise HipError(f"{what}: {lib().mxc_last_error().decode()}")


def compiler_version() -> tuple[int, int]:
    a, b = ctypes.c_int(), ctypes.c_int()
    _check(lib().mxc_compiler_version(ctypes.byref(a), ctypes.byref(b)), "hiprtc version")
    return a.value, b.value


def compile_source(source: str, name: str = "maxichem_amd.hip", options: tuple[str, ...] = OPTIONS) -> bytes:
    """hiprtc-compiled code object of source; kept in the code cache, keyed by source, options and hiprtc."""
    key = hashlib.sha256("\0".join([source, *options, repr(compiler_version())]).encode()).hexdigest()[:32]
    path = CACHE / f"{key}.hsaco"
    if path.exists():
        return path.read_bytes()
    opts = (ctypes.c_char_p * len(options))(*[o.encode() for o in options])
    code, size = ctypes.c_void_p(), ctypes.c_size_t()
    log = ctypes.create_string_buffer(1 << 16)
    rc = lib().mxc_compile(source.encode(), name.encode(), opts, len(options), ctypes.byref(code),
                           ctypes.byref(size), log, len(log))  # fmt: skip
    if rc:
        raise HipError(f"hiprtc: {lib().mxc_last_error().decode()}\n{log.value.decode(errors='replace')}")
    data = ctypes.string_at(code, size.value)
    lib().mxc_free(code)
    try:
        CACHE.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(f".tmp{os.getpid()}")
        tmp.write_bytes(data)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only tree only loses the cache
    return data


def device_info(device: int) -> dict[str, int]:
    cu, lds, wave = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().mxc_device_info(device, ctypes.byref(cu), ctypes.byref(lds), ctypes.byref(wave)), "device info")
    return {"compute_units": cu.value, "lds_per_block": lds.value, "wave": wave.value}


class Module:
    """A loaded code object on one device."""

    def __init__(self, code: bytes, device: int) -> None:
        self._code = ctypes.create_string_buffer(code, len(code))
        self.handle = ctypes.c_void_p()
        self.device = device
        _check(lib().mxc_module_load(self._code, len(code), device, ctypes.byref(self.handle)), "module load")

    def function(self, name: str) -> ctypes.c_void_p:
        f = ctypes.c_void_p()
        _check(lib().mxc_module_function(self.handle, name.encode(), ctypes.byref(f)), f"kernel {name}")
        return f

    def unload(self) -> None:
        if self.handle:
            h, self.handle = self.handle, ctypes.c_void_p()
            _check(lib().mxc_module_unload(h), "module unload")


def launch(function: ctypes.c_void_p, grid: int, block: int, lds: int, stream: int, args: list) -> None:
    """Launch with args as ctypes values (c_void_p for device pointers), in the kernel's parameter order."""
    params = (ctypes.c_void_p * len(args))(*[ctypes.addressof(a) for a in args])
    _check(lib().mxc_launch(function, grid, block, lds, ctypes.c_void_p(stream), params), "launch")


def occupancy(function: ctypes.c_void_p, block: int, lds: int) -> int:
    n = ctypes.c_int()
    _check(lib().mxc_occupancy(function, block, lds, ctypes.byref(n)), "occupancy")
    return n.value


def attribute(function: ctypes.c_void_p, which: int) -> int:
    v = ctypes.c_int()
    _check(lib().mxc_function_attribute(function, which, ctypes.byref(v)), "function attribute")
    return v.value


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    """A tensor's device pointer as a kernel argument (null for None)."""
    return ctypes.c_void_p(0 if t is None else t.data_ptr())
