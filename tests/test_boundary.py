"""The C ABI library and the drop-in interface, without a GPU: the library loads and exports every entry point its
header declares, hiprtc compiles for gfx950 here, and the call's settings are keyword-only as in the reference's
tests/test_call_interface.py."""

import ctypes
import inspect
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_entry_point():
    from maxichem_amd import _hip

    header = (ROOT / "include" / "maxichem_amd.h").read_text()
    declared = sorted(set(re.findall(r"\b(mxc_\w+)\s*\(", header)))
    assert len(declared) >= 10
    lib = ctypes.CDLL(str(_hip.LIB))
    missing = [name for name in declared if not hasattr(lib, name)]
    assert missing == []
    assert _hip.lib().mxc_version() == _hip.ABI_VERSION


def test_only_one_hip_runtime_is_mapped():
    import torch  # noqa: F401

    from maxichem_amd import _hip

    _hip.lib()
    maps = Path("/proc/self/maps").read_text()
    runtimes = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert len(runtimes) == 1, runtimes


def test_hiprtc_compiles_for_gfx950_without_a_gpu():
    from maxichem_amd import _hip

    src = 'extern "C" __global__ void k(double* a) { a[threadIdx.x] *= 2.0; }'
    code = _hip.compile_source(src, "probe.hip")
    assert code[:4] == b"\x7fELF"
    with pytest.raises(_hip.HipError, match="hiprtc"):
        _hip.compile_source("this is not HIP", "bad.hip")


def test_kernel_source_is_specialised_to_the_network(read_network):
    from maxichem_amd import _kernels, _tables

    net = read_network("synthetic")
    a = _kernels.header(_tables.build(net, False), 1, 1e-6)
    b = _kernels.header(_tables.build(net, True), 1, 1e-6)
    assert "constexpr int NJ = 0;" in a and f"constexpr int NJ = {len(net.photo)};" in b
    assert "constexpr double DRIFT = 1e-06;" in a


def test_every_optional_argument_is_keyword_only():
    from maxichem_amd.kincuda import CudaSolver

    for function, positional in ((CudaSolver.__call__, 5), (CudaSolver.launch, 5), (CudaSolver.__init__, 2)):
        params = list(inspect.signature(function).parameters.values())[1:]
        pos = [p for p in params if p.kind is not inspect.Parameter.KEYWORD_ONLY]
        assert [p.name for p in pos if p.default is not p.empty] == []
        assert len(pos) == positional


def test_settings_by_position_are_refused():
    from maxichem_amd.kincuda import CudaSolver

    solver = object.__new__(CudaSolver)
    with pytest.raises(TypeError, match=r"positional arguments? but \d+ were given"):
        CudaSolver.__call__(solver, *([None] * 5), 1e-25)


def test_results_are_named_tuples_in_the_reference_order():
    from maxichem_amd.kincuda import CallResult, LaunchResult

    assert CallResult._fields == ("x", "h", "done", "steps")
    assert LaunchResult._fields == ("x", "h", "done", "steps", "flag")


def test_bad_construction_is_refused_before_any_device_use(read_network):
    from maxichem_amd.kincuda import CudaSolver

    with pytest.raises(ValueError, match="photo=True needs a network with photolysis branches"):
        CudaSolver(read_network("synthetic_falloff"), "cuda", photo=True)
    with pytest.raises(ValueError, match="drift: a finite allowance"):
        CudaSolver(read_network("synthetic"), "cuda", drift=-1.0)
