"""Write a solver's kernels and tables for a caller without Python.

The folder is what the C ABI's mxc_solver_open (include/maxichem_amd.h) and the C++ class maxichem::Solver
(include/maxichem.hpp) load, as the reference's export feeds its C++ caller (maxichem/_export.py:26-72,
cpp/maxichem.cpp:200-290):
- code.hsaco: the kernels' code object for gfx950 (hiprtc output, the bytes a Python solver loads);
- source/maxichem_amd.hip: the translation unit it was compiled from;
- tables/<name>.bin: every table the kernels read, raw little-endian, in the Net struct's order;
- meta.txt: "key value..." lines: the format, architecture, sizes, launch settings, the call's limits and defaults,
  the count slots, species and element names, the Net struct's fields, each table's dtype and shape and each
  kernel's parameter list.
A call through the export gives the Python call's bits. The code object is for one architecture: the loader refuses
another device.
"""

from __future__ import annotations

from pathlib import Path

from . import _constants as C
from . import _hip, _kernels, _tables

EXPORT_FORMAT = 3  # 2: the recovery workgroup's size (rec_threads); 3: the guard inside the call kernel
CODE_FILE = "code.hsaco"
_DTYPES = {"int32": "int32", "int64": "int64", "float64": "float64"}


def write(folder: str | Path, *, tab: _tables.Tables, source: str, code: bytes, species: list[str],
          elements: list[str], cpw: int, batch: int, wps: int, rates_global: bool, precision: str,
          drift: float) -> Path:  # fmt: skip
    """Write the export of one compiled solver into folder (made if missing); returns the folder."""
    folder = Path(folder)
    (folder / "tables").mkdir(parents=True, exist_ok=True)
    (folder / "source").mkdir(exist_ok=True)
    (folder / "source" / "maxichem_amd.hip").write_text(source)
    (folder / CODE_FILE).write_bytes(code)
    lds_call, lds_rec = _kernels.lds_bytes(tab, cpw, precision == "mixed", rates_global, wps)
    if len(species) != tab.ns or len(elements) != tab.nel:
        raise ValueError(f"{len(species)} species and {len(elements)} elements named, the tables have {tab.ns} and "
                         f"{tab.nel}")  # fmt: skip
    lines = [
        f"export_format {EXPORT_FORMAT}",
        f"arch {_kernels.ARCH}",
        f"code {CODE_FILE}",
        f"ns {tab.ns}",
        f"nj {tab.nj}",
        f"nr {tab.nr}",
        f"nel {tab.nel}",
        f"wave {C.WAVE}",
        f"cpw {cpw}",
        f"batch {batch}",
        f"wps {wps}",
        f"rates_global {int(rates_global)}",
        f"precision {precision}",
        f"lds_call {lds_call}",
        f"lds_rec {lds_rec}",
        f"rec_threads {C.WAVE * _kernels.REC_WAVES}",
        f"rec_steps {C.REC_STEPS}",
        f"atol {C.ATOL!r}",
        f"rtol {C.DEFAULT_RTOL!r}",
        f"max_steps {C.DEFAULT_MAX_STEPS}",
        f"t_min {C.T_MIN!r}",
        f"t_max {C.T_MAX!r}",
        f"x_sum_tol {C.X_SUM_TOL!r}",
        f"drift {float(drift)!r}",
        f"tables_digest {tab.digest()}",
        f"counts {' '.join(C.COUNT_SLOTS)}",
        f"species {' '.join(species)}",
        f"elements {' '.join(elements)}",
        f"net {' '.join(_kernels.NET_FIELDS)}",
    ]
    for name, a in _kernels.host_tables(tab).items():
        a.tofile(folder / "tables" / f"{name}.bin")
        lines.append(f"table {name} {_DTYPES[str(a.dtype)]} {' '.join(str(n) for n in a.shape)}")
    for kernel in _kernels.KERNEL_NAMES:
        lines.append(f"args {kernel} {' '.join(_kernels.KERNEL_ARGUMENTS[kernel])}")
    (folder / "meta.txt").write_text("\n".join(lines) + "\n")
    return folder


def export_network(net, folder: str | Path, *, photo: bool = False, drift: float | None = None,
                   variant: tuple[int, ...] = (1, _kernels.DEFAULT_BATCH, 1, 0), precision: str = "fp64") -> Path:  # fmt: skip
    """The export of a network's solver without building one on a device (hiprtc compiles for gfx950 on any host):
    what CudaSolver(net, device, photo=photo, drift=drift, variant=variant, precision=precision).export(folder)
    writes."""
    if photo and (not net.photo or net.photo_missing):
        raise ValueError("photo=True needs a network with photolysis branches whose species are all in its species list")
    cpw, batch, wps, rg = (tuple(int(q) for q in variant) + (1, 0)[len(variant) - 2 :])[:4]
    tab = _tables.build(net, photo)
    d = float(C.DRIFT if drift is None else drift)
    src = _kernels.source(tab, cpw, d, batch, precision == "mixed", wps, bool(rg))
    return write(folder, tab=tab, source=src, code=_hip.compile_source(src), species=list(net.species),
                 elements=list(net.elements), cpw=cpw, batch=batch, wps=wps, rates_global=bool(rg),
                 precision=precision, drift=d)  # fmt: skip
