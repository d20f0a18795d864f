"""The tracers' composition: molar masses, mass mixing ratios and mole fractions, element totals, the start gas.

The reference's maxichem/coupling/composition.py (names, arguments, results and refusal messages kept, as its callers
and tests match on them). The GCM carries each species as a mass mixing ratio q (kg per kg of gas); the chemistry
takes mole fractions, and each cell keeps its tracer mass through a call. The per-cell conversions run on the GPU as
single-pass HIP kernels (kernels/coupling.hip: a wavefront takes four cells at a time, the input checks are counted
in the same pass and read back once); their tensors live on one GPU, in float64 or float32 (one precision per call). The set-up functions
(molar_masses, start_mole_fractions) return float64 tensors on the CPU, as the reference's.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np
import torch

from .. import _hip, _kernels

__all__ = [
    "ATOMIC_MASS",
    "GAS_CONSTANT",
    "START_PER_H",
    "element_moles",
    "mass_to_mole",
    "molar_masses",
    "mole_to_mass",
    "signed_mole_fractions",
    "start_mole_fractions",
    "tracer_totals",
]

# J mol^-1 K^-1: k_B N_A, both exact in the SI (composition.py:25)
GAS_CONSTANT = 8.31446261815324

# kg/mol: the atomic weights the reference takes from pyharp (composition.py:28)
ATOMIC_MASS = {"H": 1.008e-3, "He": 4.002602e-3, "C": 12.011e-3, "N": 14.007e-3, "O": 15.999e-3, "S": 32.06e-3}

# Elements per H nucleus at the start (composition.py:34-40): solar O, C, N, S to five digits, and He/H 0.09692
START_PER_H = {"O": 6.0618e-4, "C": 2.7759e-4, "N": 8.1846e-5, "S": 1.4099e-5, "He": 0.09692}

# The molecule holding each element at the start (composition.py:43-44): C in CO, the rest of O in H2O, N in N2, S in
# H2S; a network without N or S starts without them
_CARRIER = {"H": "H2", "He": "He", "C": "CO", "O": "H2O", "N": "N2", "S": "H2S"}
_OPTIONAL = ("N", "S")

_SOURCE = "coupling.hip"
_WAVE, _THREADS, _CELLS_PER_BLOCK, _GROUP_CELLS, _MAX_SPECIES, _MAX_ELEMENTS = 64, 256, 16, 32, 256, 8
_INT_MAX = 2**31 - 1
_FLAGS0 = (0, 0, _INT_MAX, 0, 0, 0, _INT_MAX, 0)  # the kernels' flags before a launch (kernels/coupling.hip)
_modules: dict[int, tuple[_hip.Module, dict[str, ctypes.c_void_p]]] = {}


def _function(device: torch.device, name: str) -> ctypes.c_void_p:
    """A coupling kernel on device (compiled once by hiprtc, loaded once per device)."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _modules:
        code = _hip.compile_source((_kernels.KERNEL_DIR / _SOURCE).read_text(), _SOURCE)
        _modules[idx] = (_hip.Module(code, idx), {})
    module, functions = _modules[idx]
    if name not in functions:
        functions[name] = module.function(name)
    return functions[name]


def _suffix(dtype: torch.dtype) -> str:
    return {torch.float64: "f64", torch.float32: "f32"}[dtype]


def _on_gpu(**tensors: torch.Tensor) -> torch.device:
    """The one GPU all the tensors are on; refuses another place."""
    devices = {a.device for a in tensors.values()}
    if len(devices) != 1 or next(iter(devices)).type != "cuda":
        where = ", ".join(f"{k} on {a.device}" for k, a in tensors.items())
        raise ValueError(f"{', '.join(tensors)} must be on one GPU (these conversions run there), got {where}")
    return next(iter(devices))


def _check_form(name: str, a: torch.Tensor, mw: torch.Tensor) -> None:
    """The shape and precision a [cells, S] array and its molar masses must have before a launch."""
    if mw.dim() != 1:
        raise ValueError(f"mw must be [S], got shape {tuple(mw.shape)}")
    if a.dim() != 2 or a.shape[1] != mw.shape[0]:
        raise ValueError(f"{name} must be [cells, {mw.shape[0]}], got shape {tuple(a.shape)}")
    if a.dtype not in (torch.float64, torch.float32):
        raise ValueError(f"{name} must be float64 or float32, got {a.dtype}")
    if mw.shape[0] > _MAX_SPECIES:
        raise ValueError(f"at most {_MAX_SPECIES} species per cell, got {mw.shape[0]}")


def _check_cells_shape(name: str, a: torch.Tensor, cells: int) -> None:
    if a.shape != (cells,):
        raise ValueError(f"{name} must be [{cells}], got shape {tuple(a.shape)}")


def _launch(device: torch.device, name: str, blocks: int, threads: int, args: list) -> None:
    with torch.cuda.device(device):
        _hip.launch(_function(device, name), blocks, threads, 0, torch.cuda.current_stream(device).cuda_stream, args)


def _blocks(cells: int) -> int:
    return max(1, -(-cells // _CELLS_PER_BLOCK))


_flags0: dict[torch.device, torch.Tensor] = {}


def _new_flags(device: torch.device) -> torch.Tensor:
    """The kernels' flags, set from a device copy of their start values (no host-to-device copy per call)."""
    if device not in _flags0:
        _flags0[device] = torch.tensor(_FLAGS0, dtype=torch.int32, device=device)
    return _flags0[device].clone()


def _refuse_mw(flags: list[int], mw: torch.Tensor) -> None:
    if flags[3]:
        raise ValueError(f"mw must be finite and positive, its least value is {float(mw.min())!r} kg/mol")


def _refuse_nonfinite(flags: list[int], name: str) -> None:
    if flags[0]:
        raise ValueError(f"{name} holds {flags[0]} non-finite entries")


def molar_masses(atoms: np.ndarray | torch.Tensor, elements: Sequence[str]) -> torch.Tensor:
    """Each species' molar mass from its atoms (composition.py:72-99): atoms [S, E] (VulcanNetwork.atoms) over the
    elements of its columns; returns [S] kg/mol, float64 on the CPU."""
    missing = [e for e in elements if e not in ATOMIC_MASS]
    if missing:
        raise ValueError(f"no atomic mass for {missing}: add them to maxichem.coupling.composition.ATOMIC_MASS")
    a = torch.as_tensor(atoms, dtype=torch.float64).cpu()
    if a.dim() != 2 or a.shape[1] != len(elements):
        raise ValueError(f"atoms must be [species, {len(elements)}], got shape {tuple(a.shape)}")
    if not bool((torch.isfinite(a) & (a >= 0)).all()):
        raise ValueError("atoms must be finite and not negative: check the network's atoms table")
    weights = torch.tensor([ATOMIC_MASS[e] for e in elements], dtype=torch.float64)
    mw = a @ weights
    empty = [int(i) for i in torch.nonzero(mw <= 0).flatten()]
    if empty:
        raise ValueError(f"species rows {empty} have no atoms, so no molar mass: check the network's atoms table")
    return mw


def _launch_to_mole(q: torch.Tensor, mw: torch.Tensor, x: torch.Tensor, flags: torch.Tensor, signed: bool) -> None:
    """The conversion kernel alone, on checked contiguous tensors (no read of the flags)."""
    cells, S = q.shape
    p = _hip.ptr
    _launch(q.device, f"mxc_to_mole_{_suffix(q.dtype)}", _blocks(cells), _THREADS,
            [p(q), p(mw), p(x), ctypes.c_int(cells), ctypes.c_int(S), ctypes.c_int(int(signed)), p(flags)])  # fmt: skip


def _to_mole(q: torch.Tensor, mw: torch.Tensor, signed: bool) -> torch.Tensor:
    dev = _on_gpu(q=q, mw=mw)
    _check_form("q", q, mw)
    if q.dtype != mw.dtype:
        raise ValueError(f"q and mw must share one precision, got {q.dtype}, {mw.dtype}")
    q, mw = q.contiguous(), mw.contiguous()
    x = torch.empty_like(q)
    flags = _new_flags(dev)
    _launch_to_mole(q, mw, x, flags, signed)
    f = flags.tolist()
    _refuse_mw(f, mw)
    _refuse_nonfinite(f, "q")
    if f[1]:
        raise ValueError(f"q holds negatives, the least {float(q.min()):.3g}: clamp them before the conversion")
    if f[2] != _INT_MAX:
        raise ValueError(f"cell {f[2]} holds no tracer mass, so it has no mole fractions: check its tracers")
    return x


def mass_to_mole(q: torch.Tensor, mw: torch.Tensor) -> torch.Tensor:
    """Mole fractions of the tracers in each cell (composition.py:102-130): q [cells, S] mass mixing ratios (kg per
    kg of gas; finite and not negative: clamp the transport's negatives first), mw [S] kg/mol, on one GPU; returns
    x [cells, S] summing to 1 in each cell. One kernel: n = q / mw, x = n / sum(n)."""
    return _to_mole(q, mw, False)


def signed_mole_fractions(q: torch.Tensor, mw: torch.Tensor) -> torch.Tensor:
    """Mole fractions for output that keep the transport's negatives (composition.py:133-159): n / sum(n), n = q /
    mw; never feed these to the chemistry (mass_to_mole refuses negative q)."""
    return _to_mole(q, mw, True)


def mole_to_mass(x: torch.Tensor, mw: torch.Tensor, mass: torch.Tensor) -> torch.Tensor:
    """Mass mixing ratios holding each cell's tracer mass (composition.py:162-192): x [cells, S] mole fractions
    (finite, not negative), mw [S] kg/mol, mass [cells] each cell's sum of q (finite, not negative), on one GPU;
    returns q [cells, S] summing to mass in each cell. One kernel: q = x mw (mass / sum(x mw))."""
    dev = _on_gpu(x=x, mw=mw, mass=mass)
    _check_form("x", x, mw)
    _check_cells_shape("mass", mass, x.shape[0])
    if not (x.dtype == mw.dtype == mass.dtype):
        raise ValueError(f"x, mw and mass must share one precision, got {x.dtype}, {mw.dtype}, {mass.dtype}")
    x, mw, mass = x.contiguous(), mw.contiguous(), mass.contiguous()
    cells, S = x.shape
    q = torch.empty_like(x)
    flags = _new_flags(dev)
    p = _hip.ptr
    _launch(dev, f"mxc_to_mass_{_suffix(x.dtype)}", _blocks(cells), _THREADS,
            [p(x), p(mw), p(mass), p(q), ctypes.c_int(cells), ctypes.c_int(S), p(flags)])  # fmt: skip
    f = flags.tolist()
    _refuse_mw(f, mw)
    _refuse_nonfinite(f, "x")
    if f[4]:
        raise ValueError(f"mass must be finite and not negative, its least value is {float(mass.min())!r}")
    if f[1]:
        raise ValueError(f"x holds negatives, the least {float(x.min()):.3g}: the call's result is not usable")
    if f[2] != _INT_MAX:
        raise ValueError(f"cell {f[2]} has mole fractions of no mass (sum of x * mw <= 0): check the call's result")
    return q


def _moles(q, mw, atoms, cell_mass=None):
    """element_moles' checks and kernel; with cell_mass, the tracers' totals instead."""
    tensors = {"q": q, "mw": mw, "atoms": atoms, **({} if cell_mass is None else {"cell_mass": cell_mass})}
    dev = _on_gpu(**tensors)
    _check_form("q", q, mw)
    if atoms.dim() != 2 or atoms.shape[0] != mw.shape[0]:
        raise ValueError(f"atoms must be [{mw.shape[0]}, elements], got shape {tuple(atoms.shape)}")
    if not (q.dtype == mw.dtype == atoms.dtype):
        raise ValueError(f"q, mw and atoms must share one precision, got {q.dtype}, {mw.dtype}, {atoms.dtype}")
    cells, S = q.shape
    E = atoms.shape[1]
    if E > _MAX_ELEMENTS:
        raise ValueError(f"at most {_MAX_ELEMENTS} elements, got {E}")
    if cell_mass is not None:
        _check_cells_shape("cell_mass", cell_mass, cells)
        if cell_mass.dtype not in (torch.float64, torch.float32):
            raise ValueError(f"cell_mass must be float64 or float32, got {cell_mass.dtype}")
        cell_mass = cell_mass.to(q.dtype).contiguous()
    q, mw, atoms = q.contiguous(), mw.contiguous(), atoms.contiguous()
    blocks = max(1, -(-cells // _GROUP_CELLS))  # 8 lanes per cell; the totals' order depends on the cells alone
    out = torch.empty(cells, E, dtype=q.dtype, device=dev) if cell_mass is None else None
    partial = None if cell_mass is None else torch.empty(blocks, 1 + E, dtype=q.dtype, device=dev)
    flags = _new_flags(dev)
    p = _hip.ptr
    sfx = _suffix(q.dtype)
    _launch(dev, f"mxc_element_moles_{sfx}", blocks, _THREADS,
            [p(q), p(mw), p(atoms), p(cell_mass), p(out), p(partial), ctypes.c_int(cells), ctypes.c_int(S),
             ctypes.c_int(E), p(flags)])  # fmt: skip
    totals = None
    if partial is not None:
        totals = torch.empty(1 + E, dtype=q.dtype, device=dev)
        _launch(dev, f"mxc_column_sums_{sfx}", 1 + E, _THREADS,
                [p(partial), ctypes.c_int(blocks), ctypes.c_int(1 + E), p(totals)])  # fmt: skip
    f = flags.tolist()
    _refuse_mw(f, mw)
    _refuse_nonfinite(f, "q")
    if f[5]:
        raise ValueError("atoms holds non-finite entries: check the network's atoms table")
    if f[4]:
        raise ValueError(f"cell_mass must be finite and not negative, its least value is {float(cell_mass.min())!r}")
    return out if totals is None else totals


def element_moles(q: torch.Tensor, mw: torch.Tensor, atoms: torch.Tensor) -> torch.Tensor:
    """Each cell's moles of each element per kg of gas (composition.py:195-217): q [cells, S] mass mixing ratios
    (finite; the transport's negatives count as they are), mw [S] kg/mol, atoms [S, E], on one GPU; returns
    [cells, E] mol per kg: (q / mw) atoms, one kernel."""
    return _moles(q, mw, atoms)


def tracer_totals(q: torch.Tensor, cell_mass: torch.Tensor, mw: torch.Tensor, atoms: torch.Tensor) -> torch.Tensor:
    """The tracers' mass and each element's moles summed over the cells (composition.py:220-233): q [cells, S],
    cell_mass [cells] kg (finite, not negative), mw [S], atoms [S, E], on one GPU; returns [1 + E]: the tracer mass
    in kg, then the moles of each element in atoms' column order. Two kernels: each block's cells weighted by their
    mass, then the blocks' sums, both in a fixed order (the same bits on every run)."""
    return _moles(q, mw, atoms, cell_mass)


def start_mole_fractions(species: Sequence[str], elements: Sequence[str]) -> torch.Tensor:
    """The start gas (composition.py:236-275): H2 and He with START_PER_H of each element in its carrier molecule
    (C in CO, the rest of O in H2O, N in N2, S in H2S; a network without N or S starts without them), which a chemistry
    call relaxes to the cell's state. Returns [S] mole fractions in the species' order, float64 on the CPU, summing
    to 1."""
    unknown = [e for e in elements if e not in _CARRIER]
    if unknown:
        raise ValueError(
            f"the network has elements {unknown} with no start carrier: add one to maxichem.coupling.composition"
        )
    present = [e for e in _CARRIER if e in elements or e not in _OPTIONAL]
    lacking = [f"{_CARRIER[e]} (for {e})" for e in present if _CARRIER[e] not in species]
    if lacking:
        raise ValueError(f"the network lacks the start species {lacking}: the start mixture needs them")
    per_h = {e: START_PER_H[e] if e in present else 0.0 for e in START_PER_H}
    # H2 molecules per H nucleus: each H nucleus is in H2, H2O (O beyond C) or H2S, two per molecule
    h2_per_h = 0.5 - (per_h["O"] - per_h["C"]) - per_h["S"]
    molecules = {
        "H2": h2_per_h,
        "He": per_h["He"],
        "CO": per_h["C"],
        "H2O": per_h["O"] - per_h["C"],
        "N2": 0.5 * per_h["N"],
        "H2S": per_h["S"],
    }
    x = torch.zeros(len(species), dtype=torch.float64)
    order = list(species)
    for name, amount in molecules.items():
        if amount:
            x[order.index(name)] = amount
    return x / x.sum()
