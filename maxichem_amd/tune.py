"""The call kernel's autotuner: cells (wavefronts) per workgroup, the stream batch, the register budget (wavefronts
per SIMD) and the rate coefficients' place (LDS or global scratch), per network and device.

Each candidate is compiled, loaded and timed on one synthetic uneven call (solar gas from cold, dt log-uniform
1e1-1e4 s, half of the cells lit), by HIP events around the call kernel; the fastest is stored in the code cache
(tune.json) under the network's table digest, the photolysis switch and the device name, and later solvers of that
network on that device take it.
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


def _store() -> Path:
    return STORE if STORE is not None else _hip.CACHE / "tune.json"


def _key(net, photo: bool, device: torch.device) -> str:
    return f"{_tables.build(net, photo).digest()}|{int(photo)}|{torch.cuda.get_device_name(device)}"


def stored(net, device, photo: bool) -> dict | None:
    try:
        return json.loads(_store().read_text()).get(_key(net, photo, device))
    except (OSError, ValueError):
        return None


def choose(net, device, photo: bool, drift: float, run: bool = False) -> tuple[int, int]:
    """The variant for this network on this device: tuned now when run, else the stored choice, else DEFAULT."""
    if run:
        return tuple(tune(net, device, photo, drift)["best"])
    s = stored(net, device, photo)
    return tuple(s["best"]) if s else DEFAULT


def tune(net, device, photo: bool, drift: float, cells: int = 8192, seed: int = 7, repeats: int = 2) -> dict:
    """Time every variant on one uneven call; store and return {"best", "ms": {variant: ms}}."""
    from .kincuda import CudaSolver

    dev = torch.device(device)
    nj = len(net.photo) if photo else 0
    d = workloads.config3(net.species, len(net.photo), cells=cells, seed=seed)
    rng = np.random.default_rng(seed)
    dt = 10.0 ** rng.uniform(1.0, 4.0, cells)

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    x, T, P, h, dtt = g(d["x"]), g(d["T"]), g(d["P"]), g(d["h"]), g(dt)
    J = g(d["J"][:, :nj]) if photo else None
    ms = {}
    for v in VARIANTS:
        with CudaSolver(net, dev, photo=photo, drift=drift, variant=v) as s:
            s(x, T, P, dtt, h, J=J)  # the first call pays for loading
            s.timing = []
            for _ in range(repeats):
                s(x, T, P, dtt, h, J=J)
            torch.cuda.synchronize(dev)
            ms[",".join(map(str, v))] = float(np.median([a.elapsed_time(b) for a, b in s.timing]))
            s.timing = None
    best = min(ms, key=ms.get)
    out = {"best": [int(q) for q in best.split(",")], "ms": ms, "cells": cells}
    path = _store()
    try:
        data = json.loads(path.read_text()) if path.exists() else {}
    except ValueError:
        data = {}
    data[_key(net, photo, dev)] = out
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(f".tmp{os.getpid()}")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass
    return out
