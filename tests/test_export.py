"""The export and the solver without Python (the reference's tests/test_export_cpp.py and cpp/test_export.cpp):

- without a GPU: the export's files and meta.txt, the kernel argument lists (the kernels' own parameters, the Python
  list and the C ABI's list agree), and mxc_solver_open refusing a bad export before any device work;
- on the GPU: the C ABI's solver (through ctypes) and the C++ class maxichem::Solver (tests/cpp/test_export.cpp,
  a program linked against libtorch) against the Python call, bit for bit, counts and flags included, for thermal
  and photolysis networks, a forced recovery, the rate coefficients in global scratch, and the refusals.
"""

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from maxichem_amd import _export, _hip, _kernels, standin, workloads
from maxichem_amd._constants import COUNT_SLOTS, DEFAULT_MAX_STEPS

ROOT = Path(__file__).resolve().parent.parent
PROGRAM = ROOT / "tests" / "cpp" / "test_export"
CELLS = 2048
FORCED_TRIES = 2  # most cells end unfinished and the recovery recomputes them


def kernel_parameters(kernel: str) -> list[str]:
    """A kernel's parameter names, read from its HIP source."""
    text = "\n".join((_kernels.KERNEL_DIR / f).read_text() for f in _kernels.KERNEL_FILES)
    m = re.search(rf"\b{kernel}\s*\(([^)]*)\)\s*\{{", text)
    assert m, kernel
    return [re.findall(r"(\w+)\s*$", p.strip())[0] for p in m.group(1).split(",")]


def library_arguments(kernel: str) -> list[str]:
    buf = ctypes.create_string_buffer(1024)
    assert _hip.lib().mxc_kernel_arguments(kernel.encode(), buf, len(buf)) == 0, _hip.lib().mxc_last_error()
    return buf.value.decode().split()


def test_argument_lists_agree_with_the_kernels():
    for kernel in _kernels.KERNEL_NAMES:
        assert kernel_parameters(kernel) == list(_kernels.KERNEL_ARGUMENTS[kernel]) == library_arguments(kernel)
    assert _hip.lib().mxc_kernel_arguments(b"mxc_nothing", ctypes.create_string_buffer(8), 8) == 1


def test_net_fields_are_the_kernels_struct():
    text = (_kernels.KERNEL_DIR / "common.hip").read_text()
    body = re.search(r"struct Net \{(.*?)\};", text, re.S).group(1)
    assert re.findall(r"\*\s*(\w+);", body) == list(_kernels.NET_FIELDS)


@pytest.fixture(scope="module")
def ncho_export(tmp_path_factory):
    return _export.export_network(standin.load("NCHO"), tmp_path_factory.mktemp("ncho") / "export", photo=True)


def meta_lines(folder: Path) -> list[str]:
    return (folder / "meta.txt").read_text().splitlines()


def test_export_files_match_meta(ncho_export):
    lines = meta_lines(ncho_export)
    meta = {ln.split()[0]: ln.split()[1:] for ln in lines if not ln.startswith(("table ", "args "))}
    assert meta["export_format"] == [str(_export.EXPORT_FORMAT)] and meta["arch"] == ["gfx950"]
    net = standin.load("NCHO")
    assert meta["species"] == list(net.species) and meta["elements"] == list(net.elements)
    assert (ncho_export / "code.hsaco").read_bytes()[:4] == b"\x7fELF"
    sizes = {"int32": 4, "int64": 8, "float64": 8}
    tables = [ln.split() for ln in lines if ln.startswith("table ")]
    assert [t[1] for t in tables] == list(_kernels.NET_FIELDS) == meta["net"]
    for _, name, dtype, *shape in tables:
        assert (ncho_export / "tables" / f"{name}.bin").stat().st_size == sizes[dtype] * int(np.prod([int(q) for q in shape]))
    atoms = np.fromfile(ncho_export / "tables" / "atoms.bin").reshape(len(net.species), len(net.elements))
    np.testing.assert_array_equal(atoms, net.atoms)


def open_error(folder: Path) -> str:
    s = ctypes.c_void_p()
    rc = _hip.lib().mxc_solver_open(str(folder).encode(), 0, ctypes.byref(s))
    assert rc == 1 and not s.value
    return _hip.lib().mxc_last_error().decode()


EDITS = {  # meta.txt edits mxc_solver_open refuses before any device work, and the message each gives
    "format": ("export_format 2", "export_format 3", "export format 3, this library reads 2"),
    "no recovery workgroup": ("rec_threads 256", "rec_threads 96", "sizes out of range"),
    "unknown": ("args mxc_recover net", "args mxc_recover net2", "kernel mxc_recover takes argument net2, which"),
    "repeated": ("args mxc_guard net x0 xout", "args mxc_guard net x0 x0 xout", "kernel mxc_guard lists argument x0 twice"),
    "out of order": ("args mxc_call net x0 Ta Pa", "args mxc_call net x0 Pa Ta",
                     "kernel mxc_call has argument Pa at position 2, where this library passes Ta"),
    "dropped": ("args mxc_guard net x0 xout doneo noko status redo list count B atol",
                "args mxc_guard net x0 xout doneo noko status redo list count B", "kernel mxc_guard lacks argument atol"),
    "net order": ("net perm rc rf", "net rc perm rf", "net field 0 is rc, the kernels read perm there"),
    "table shape": ("table atoms float64 69 5", "table atoms float64 69 6", "table atoms has 2760 bytes"),
}  # fmt: skip


@pytest.mark.parametrize("edit", EDITS)
def test_open_refuses_a_bad_export_before_device_work(ncho_export, tmp_path, edit):
    old, new, message = EDITS[edit]
    folder = tmp_path / "export"
    subprocess.run(["cp", "-r", str(ncho_export), str(folder)], check=True)
    text = (folder / "meta.txt").read_text()
    assert old in text
    (folder / "meta.txt").write_text(text.replace(old, new, 1))
    assert message in open_error(folder)


def test_open_refuses_a_missing_table(ncho_export, tmp_path):
    folder = tmp_path / "export"
    subprocess.run(["cp", "-r", str(ncho_export), str(folder)], check=True)
    (folder / "tables" / "chw.bin").unlink()
    assert "cannot read" in open_error(folder) and "chw.bin" in _hip.lib().mxc_last_error().decode()
    assert "cannot read" in open_error(tmp_path / "nowhere")


# ---- on the GPU ----


def save(folder: Path, **arrays) -> None:
    folder.mkdir(parents=True, exist_ok=True)
    for name, a in arrays.items():
        np.ascontiguousarray(a).tofile(folder / f"{name}.bin")


def make_case(solver, folder: Path, cells: int, seed: int, max_steps: int, edit: str | None = None) -> dict:
    """Inputs of the first call from cold (solar gas, dt = 300 s, J on the lowest-P quarter), the Python call's
    results or refusal, written for the C++ program."""
    net_species = solver._names[0]
    d = workloads.config1(net_species, max(solver.nj, 1), cells=cells, seed=seed)
    x, T, P, dt, h = d["x"], d["T"], d["P"], d["dt"], d["h"]
    J = d["J"][:, : solver.nj]
    if edit == "sum":
        x[5] *= 1.5
    elif edit == "negative":
        x[5, 0] = -x[5, 0]
    elif edit == "nan":
        T[5] = np.nan
    elif edit == "hot":
        T[5] = 3500.0
    elif edit == "inside":
        T[5] = 3000.0
    elif edit == "straddle":
        T[5], T[9], x[11] = 200.0, 3500.0, 2.0 * x[11]
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    folder.mkdir(parents=True, exist_ok=True)
    save(folder, x=x, T=T, P=P, dt=dt, h=h, **({"J": J} if solver.nj else {}))
    (folder / "B.txt").write_text(f"{cells}\n")
    (folder / "max_steps.txt").write_text(f"{max_steps}\n")
    inputs = dict(x=g(x), T=g(T), P=g(P), dt=g(dt), h=g(h), J=g(J) if solver.nj else None)
    try:
        r = solver.launch(inputs["x"], inputs["T"], inputs["P"], inputs["dt"], inputs["h"], max_steps=max_steps,
                          J=inputs["J"])  # fmt: skip
    except ValueError as e:
        (folder / "refused.txt").write_text(str(e))
        save(folder, counts_out=solver.count.cpu().numpy())
        return {"inputs": inputs, "refused": str(e)}
    save(folder, x_out=r.x.cpu().numpy(), h_out=r.h.cpu().numpy(), steps_out=r.steps.cpu().numpy(),
         done_out=r.done.cpu().numpy().astype(np.int8), counts_out=solver.count.cpu().numpy())  # fmt: skip
    return {"inputs": inputs, "result": r, "count": solver.count.cpu().numpy()}


def run_program(export: Path, case: Path) -> subprocess.CompletedProcess:
    if not PROGRAM.exists():
        raise FileNotFoundError(f"{PROGRAM} is missing: build it (python -c 'import __graft_entry__ as g; g.build()')")
    r = subprocess.run([str(PROGRAM), str(export), str(case)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    return r


class AbiSolver:
    """The C ABI's solver through ctypes."""

    def __init__(self, folder: Path, device: int = 0) -> None:
        self.lib = _hip.lib()
        self.h = ctypes.c_void_p()
        assert self.lib.mxc_solver_open(str(folder).encode(), device, ctypes.byref(self.h)) == 0, self.err()
        ns, nj, nel, slots = (ctypes.c_int() for _ in range(4))
        assert self.lib.mxc_solver_sizes(self.h, *map(ctypes.byref, (ns, nj, nel, slots))) == 0
        self.ns, self.nj, self.nel, self.slots = ns.value, nj.value, nel.value, slots.value

    def err(self) -> str:
        return self.lib.mxc_last_error().decode()

    def call(self, x, T, P, dt, h, J=None, rtol=1e-3, atol=1e-25, max_steps=DEFAULT_MAX_STEPS):
        B = x.shape[0]
        out = dict(x=torch.empty_like(x), h=torch.empty_like(h), steps=torch.empty_like(h),
                   done=torch.empty(B, dtype=torch.int32, device=x.device),
                   flag=torch.empty(B, dtype=torch.int32, device=x.device))  # fmt: skip
        p = _hip.ptr
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = self.lib.mxc_solver_call(self.h, stream, B, p(x), p(T), p(P), p(dt), p(h), p(J), rtol, atol, max_steps,
                                      p(out["x"]), p(out["h"]), p(out["done"]), p(out["steps"]), p(out["flag"]))  # fmt: skip
        return rc, out

    def counts(self):
        c = (ctypes.c_int * self.slots)()
        t = (ctypes.c_longlong * self.slots)()
        assert self.lib.mxc_solver_counts(self.h, c, t) == 0
        return np.array(c[:]), np.array(t[:])

    def close(self) -> None:
        if self.h:
            assert self.lib.mxc_solver_close(self.h) == 0, self.err()
            self.h = ctypes.c_void_p()


def bits(t: torch.Tensor) -> np.ndarray:
    return t.cpu().contiguous().numpy().view(np.uint8)


SOLVERS = {  # network, photolysis, variant
    "NCHO photolysis": ("NCHO", True, (1, 16, 1, 0)),
    "NCHO thermal": ("NCHO", False, (1, 16, 1, 0)),
    "SNCHO photolysis": ("SNCHO", True, (1, 16, 1, 0)),
    "NCHO rates in global scratch": ("NCHO", True, (1, 16, 1, 1)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("max_steps", [DEFAULT_MAX_STEPS, FORCED_TRIES], ids=["default", "recovery"])
@pytest.mark.parametrize("which", SOLVERS)
def test_abi_and_cpp_calls_are_the_python_call(gpu, tmp_path, which, max_steps):
    from maxichem_amd.kincuda import CudaSolver

    shape, photo, variant = SOLVERS[which]
    with CudaSolver(standin.load(shape), "cuda", photo=photo, variant=variant) as solver:
        solver.export(tmp_path / "export")
        case = make_case(solver, tmp_path / "case", CELLS, 17, max_steps)
        want, count = case["result"], case["count"]
    print(f"{which}, max_steps {max_steps}: counts {count.tolist()}")
    if max_steps == FORCED_TRIES:
        assert count[COUNT_SLOTS.index("recomputed")] > CELLS // 2  # the recovery recomputed most cells
    a = AbiSolver(tmp_path / "export")
    try:
        for k in range(2):  # twice: the second call reuses the work buffers
            rc, got = a.call(**case["inputs"], max_steps=max_steps)
            assert rc == 0, a.err()
            for name in ("x", "h", "steps", "flag"):
                assert np.array_equal(bits(got[name]), bits(getattr(want, name))), (k, name)
            assert torch.equal(got["done"].bool(), want.done)
            c, t = a.counts()
            assert np.array_equal(c, count) and np.array_equal(t, (k + 1) * count.astype(np.int64))
    finally:
        a.close()
    r = run_program(tmp_path / "export", tmp_path / "case")
    assert r.returncode == 0 and "bitwise identical to the Python call" in r.stdout
    assert ("(default arguments)" in r.stdout) == (max_steps == DEFAULT_MAX_STEPS)
    assert ("J [B, 0] gives the same bits, J [B, 1] refused" in r.stdout) == (not photo)
    assert "HIP runtimes mapped 1" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("edit", ["sum", "negative", "nan", "hot", "inside", "straddle"])
def test_cpp_refuses_what_python_refuses(gpu, tmp_path, edit):
    from maxichem_amd.kincuda import CudaSolver

    with CudaSolver(standin.load("NCHO"), "cuda", photo=True) as solver:
        solver.export(tmp_path / "export")
        case = make_case(solver, tmp_path / "case", 256, 5, DEFAULT_MAX_STEPS, edit)
    refused = "refused" in case
    assert refused == (edit != "inside")
    if refused:
        assert "(first: cell 5" in case["refused"] or "(first: cell 9" in case["refused"]
        a = AbiSolver(tmp_path / "export")
        try:
            rc, _ = a.call(**case["inputs"])
            assert rc == 2 and a.err() == case["refused"]
        finally:
            a.close()
    r = run_program(tmp_path / "export", tmp_path / "case")
    assert r.returncode == 0
    assert ("refused as the Python call" if refused else "bitwise identical") in r.stdout


@pytest.mark.gpu
def test_cpp_constructor_refuses_another_contract(gpu, ncho_export, tmp_path):
    folder = tmp_path / "export"
    subprocess.run(["cp", "-r", str(ncho_export), str(folder)], check=True)
    text = (folder / "meta.txt").read_text()
    (folder / "meta.txt").write_text(text.replace("args mxc_guard net x0 xout", "args mxc_guard net x0 x0 xout"))
    (tmp_path / "case").mkdir()
    r = run_program(folder, tmp_path / "case")
    assert r.returncode != 0 and "kernel mxc_guard lists argument x0 twice" in r.stderr
    assert "export loaded" not in r.stdout
