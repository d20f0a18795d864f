"""The call kernel's autotuner: cells (wavefronts) per workgroup, the stream batch, the register budget (wavefronts
per SIMD) and the rate coefficients' place (LDS or global scratch), per network, device and workload.

Each candidate is compiled, loaded and timed by HIP events around the call kernel on a sample of the workload:
"uneven", one synthetic call from cold (solar gas, dt log-uniform 1e1-1e4 s, half of the cells lit), or "settled",
calls of dt = 300 s after a dark relaxation (the GPU FP64 case's sequence, a quarter of the cells lit). The fastest
is stored in the code cache (tune.json) under the network's table digest, the photolysis switch, the device name and
the workload, and later solvers of that network on that device take the uneven choice unless told the workload.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np
import torch

from . import _hip, _kernels, _tables, workloads

VARIANTS = ((1, 16, 1, 0), (1, 8, 1, 0), (2, 16, 1, 0), (4, 16, 1, 0), (1, 16, 2, 0), (1, 16, 1, 1), (1, 16, 2, 1))
DEFAULT = (1, _kernels.DEFAULT_BATCH, 1, 0)
STORE: Path | None = None  # where the choices are kept; None: tune.json in the code cache
CU_LDS = 160 * 1024  # LDS per compute unit (gfx950)
CU_WAVES = 4  # call wavefronts per compute unit at one per SIMD


def default_variant(tab, mixed: bool = False) -> tuple[int, int, int, int]:
    """The variant without a stored choice, with the same bits as every other: four cells per workgroup when the
    stream words then fit in the workgroup's LDS at as many cells per compute unit (on MI355X 2.6 % faster on NCHO:
    the words come with the LDS's latency instead of the caches'); else two when that keeps as many cells per
    compute unit as one (about 2 % faster: single-wavefront workgroups leave the SIMDs less evenly loaded); else
    one."""

    def cells(cpw: int) -> int:
        return min(CU_WAVES, CU_LDS // _kernels.lds_bytes(tab, cpw, mixed)[0] * cpw)

    best = max(cells(c) for c in (1, 2, 4))
    if _kernels.words_in_lds(tab, 4, mixed) and cells(4) >= best:
        return (4, _kernels.DEFAULT_BATCH, 1, 0)
    return (2 if cells(2) >= cells(1) else 1, _kernels.DEFAULT_BATCH, 1, 0)


def _store() -> Path:
    return STORE if STORE is not None else _hip.CACHE / "tune.json"


WORKLOADS = ("uneven", "settled")


def _key(net, photo: bool, device: torch.device, workload: str = "uneven") -> str:
    key = f"{_tables.build(net, photo).digest()}|{int(photo)}|{torch.cuda.get_device_name(device)}"
    return key if workload == "uneven" else f"{key}|{workload}"


def stored(net, device, photo: bool, workload: str = "uneven") -> dict | None:
    try:
        return json.loads(_store().read_text()).get(_key(net, photo, device, workload))
    except (OSError, ValueError):
        return None


def choose(net, device, photo: bool, drift: float, run: bool = False, workload: str = "uneven",
           mixed: bool = False) -> tuple[int, ...]:  # fmt: skip
    """The variant for this network, device and workload: tuned now when run, else the stored choice, else
    default_variant."""
    if run:
        return tuple(tune(net, device, photo, drift, workload=workload)["best"])
    s = stored(net, device, photo, workload)
    return tuple(s["best"]) if s else default_variant(_tables.build(net, photo), mixed)


def tune(net, device, photo: bool, drift: float, cells: int = 8192, seed: int = 7, repeats: int = 2,
         workload: str = "uneven") -> dict:  # fmt: skip
    """Time every variant on the workload's sample; store and return {"best", "ms": {variant: ms}, "private_bytes":
    {variant: bytes per lane the variant spills}}."""
    from .kincuda import CudaSolver

    if workload not in WORKLOADS:
        raise ValueError(f"workload: one of {WORKLOADS} expected, got {workload!r}")
    dev = torch.device(device)
    nj = len(net.photo) if photo else 0

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    if workload == "uneven":
        d = workloads.config3(net.species, len(net.photo), cells=cells, seed=seed)
        rng = np.random.default_rng(seed)
        dt = 10.0 ** rng.uniform(1.0, 4.0, cells)
        x, T, P, h, dtt = g(d["x"]), g(d["T"]), g(d["P"]), g(d["h"]), g(dt)
        J = g(d["J"][:, :nj]) if photo else None
    else:
        d = workloads.config1(net.species, len(net.photo), cells=cells, seed=seed)
        T, P = g(d["T"]), g(d["P"])
        J = g(d["J"][:, :nj]) if photo else None
        with CudaSolver(net, dev, photo=photo, drift=drift, variant=DEFAULT) as s:
            r = s(g(d["x"]), T, P, g(d["dt_relax"]), g(d["h"]), J=None if J is None else torch.zeros_like(J))
            x, h, dtt = r.x, r.h, g(d["dt"])
    ms, spills = {}, {}
    for v in VARIANTS:
        with CudaSolver(net, dev, photo=photo, drift=drift, variant=v) as s:
            s(x, T, P, dtt, h, J=J)  # the first call pays for loading
            s.timing = []
            for _ in range(repeats):
                s(x, T, P, dtt, h, J=J)
            torch.cuda.synchronize(dev)
            key = ",".join(map(str, v))
            ms[key] = float(np.median([a.elapsed_time(b) for a, b in s.timing]))
            spills[key] = int(s.info["private_bytes"])
            s.timing = None
    best = min(ms, key=ms.get)
    out = {"best": [int(q) for q in best.split(",")], "ms": ms, "private_bytes": spills, "cells": cells,
           "workload": workload}  # fmt: skip
    path = _store()
    try:
        data = json.loads(path.read_text()) if path.exists() else {}
    except ValueError:
        data = {}
    data[_key(net, photo, dev, workload)] = out
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(f".tmp{os.getpid()}")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass
    return out
