"""The export and the solver without Python (the reference's tests/test_export_cpp.py and cpp/test_export.cpp):

- without a GPU: the export's files and meta.txt, the kernel argument lists (the kernels' own parameters, the Python
  list and the C ABI's list agree), and mxc_solver_open refusing a bad export before any device work;
- on the GPU: the C ABI's solver (through ctypes) and the C++ class maxichem::Solver (driven by
  tests/cpp/cpp_caller.cpp, a program linked against libtorch that runs a plan of calls and writes what each
  returned) against the Python call, bit for bit, counts and flags included, for thermal and photolysis networks, a
  forced recovery, the rate coefficients in global scratch, and the refusals.
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
PROGRAM = ROOT / "tests" / "cpp" / "cpp_caller"
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
    assert re.findall(r"gptr<\w+>\s+(\w+);", body) == list(_kernels.NET_FIELDS)


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
    "format": ("export_format 3", "export_format 2", "export format 2, this library reads 3"),
    "no recovery workgroup": ("rec_threads 256", "rec_threads 96", "sizes out of range"),
    "unknown": ("args mxc_recover net", "args mxc_recover net2", "kernel mxc_recover takes argument net2, which"),
    "repeated": ("args mxc_recover net x0 Ta", "args mxc_recover net x0 x0 Ta",
                 "kernel mxc_recover lists argument x0 twice"),
    "out of order": ("args mxc_call net x0 Ta Pa", "args mxc_call net x0 Pa Ta",
                     "kernel mxc_call has argument Pa at position 2, where this library passes Ta"),
    "dropped": ("max_steps rscratch\n", "max_steps\n", "kernel mxc_recover lacks argument rscratch"),
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
    """Inputs of the first call from cold (solar gas, dt = 300 s, J on the lowest-P quarter), written for the C++
    driver, and the Python call's results or refusal."""
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
    inputs = dict(x=g(x), T=g(T), P=g(P), dt=g(dt), h=g(h), J=g(J) if solver.nj else None)
    try:
        r = solver.launch(inputs["x"], inputs["T"], inputs["P"], inputs["dt"], inputs["h"], max_steps=max_steps,
                          J=inputs["J"])  # fmt: skip
    except ValueError as e:
        return {"inputs": inputs, "refused": str(e), "count": solver.count.cpu().numpy()}
    return {"inputs": inputs, "result": r, "count": solver.count.cpu().numpy()}


def run_program(export: Path, case: Path, plan: list[str]) -> tuple[subprocess.CompletedProcess, Path]:
    """Runs the C++ driver on a plan of steps (tests/cpp/cpp_caller.cpp); returns its process and output folder."""
    if not PROGRAM.exists():
        raise FileNotFoundError(f"{PROGRAM} is missing: build it (python -c 'import __graft_entry__ as g; g.build()')")
    (case / "plan.txt").write_text("\n".join(plan) + "\n")
    out = case / "out"
    out.mkdir(exist_ok=True)
    r = subprocess.run([str(PROGRAM), str(export), str(case), str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    return r, out


def program_call(out: Path, tag: str, cells: int, ns: int) -> dict:
    """What one call of the driver's plan returned: the outputs and counts, or the exception's kind and message."""
    res = {"count": np.fromfile(out / f"{tag}.count.bin", np.int32)}
    err = out / f"{tag}.error.txt"
    if err.exists():
        res["kind"], res["message"] = err.read_text().split("\n", 1)
        return res
    res["x"] = np.fromfile(out / f"{tag}.x.bin").reshape(cells, ns)
    for name in ("h", "steps"):
        res[name] = np.fromfile(out / f"{tag}.{name}.bin")
    res["done"] = np.fromfile(out / f"{tag}.done.bin", np.int8).astype(bool)
    return res


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
    # the C++ class: two calls (the work buffers reused), a thermal network's empty and refused J, then the
    # accessors' tensors changed in place before a last call
    steps = "default" if max_steps == DEFAULT_MAX_STEPS else str(max_steps)
    j = "given" if photo else "none"
    plan = [f"call first {steps} {j}", f"call second {steps} {j}"]
    plan += [] if photo else [f"call empty {steps} empty", f"call columns {steps} columns"]
    plan += ["state before", "scramble", f"call last {steps} {j}", "state after"]
    r, out = run_program(tmp_path / "export", tmp_path / "case", plan)
    assert r.returncode == 0
    ns = want.x.shape[1]
    tags = ["first", "second", "last"] + ([] if photo else ["empty"])
    for tag in tags:
        got = program_call(out, tag, CELLS, ns)
        for name in ("x", "h", "steps"):
            assert np.array_equal(got[name].view(np.uint8), bits(getattr(want, name))), (tag, name)
        assert np.array_equal(got["done"], want.done.cpu().numpy()) and np.array_equal(got["count"], count), tag
    if not photo:
        got = program_call(out, "columns", CELLS, ns)
        assert got["kind"] == "c10" and "J: none or an empty" in got["message"]
    before, after = ({k: np.fromfile(out / f"{tag}.{k}.bin", dt) for k, dt in
                      (("count", np.int32), ("total", np.int64), ("atoms", np.float64))} for tag in ("before", "after"))  # fmt: skip
    calls = len(tags)
    assert np.array_equal(before["total"], (calls - 1) * count.astype(np.int64))
    assert np.array_equal(after["total"], calls * count.astype(np.int64))  # the solver's own totals, not the copy
    assert np.array_equal(after["atoms"], before["atoms"]) and np.array_equal(after["count"], count)
    assert np.array_equal(after["atoms"].reshape(ns, -1), standin.load(shape).atoms)
    assert (out / "runtimes.txt").read_text() == "1"  # the library shares torch's HIP runtime


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
    r, out = run_program(tmp_path / "export", tmp_path / "case", ["call first default given"])
    assert r.returncode == 0
    got = program_call(out, "first", 256, 69)
    if refused:
        assert got["kind"] == "invalid_argument" and got["message"] == "maxichem: " + case["refused"]
        assert np.array_equal(got["count"], case["count"])
    else:
        assert np.array_equal(got["x"].view(np.uint8), bits(case["result"].x))


@pytest.mark.gpu
def test_cpp_constructor_refuses_another_contract(gpu, ncho_export, tmp_path):
    folder = tmp_path / "export"
    subprocess.run(["cp", "-r", str(ncho_export), str(folder)], check=True)
    text = (folder / "meta.txt").read_text()
    (folder / "meta.txt").write_text(text.replace("args mxc_recover net x0 Ta", "args mxc_recover net x0 x0 Ta"))
    (tmp_path / "case").mkdir()
    r, _ = run_program(folder, tmp_path / "case", [])
    assert r.returncode == 3 and "kernel mxc_recover lists argument x0 twice" in r.stderr
    assert "opened" not in r.stdout
