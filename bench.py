#!/usr/bin/env python3
"""Cells/s per call of the batched Rodas3 chemistry step on one MI355X.

Workload (the GPU FP64 case): 65,536 cells of the 69-species C-H-N-O network (390 reversible thermal reactions,
48 photolysis branches; the seeded stand-in of that shape unless VULCAN_THERMO holds the real files). Per cell, from
numpy.random.default_rng(1): T log-uniform 300-3000 K, P log-uniform 1e2-1e9 dyn cm^-2, a solar gas. One dark call of
dt = 1e8 s relaxes the cells (set-up, untimed); then W untimed and K timed calls of dt = 300 s follow, each starting
from the previous call's x and h, with per-branch J log-uniform 1e-9 to 1e-3 s^-1 on the lowest-P quarter of cells
and J = 0 elsewhere. All float64, inputs resident in HBM.

A step is one call, solver(x, T, P, dt, h, J=J): the call kernel, the guard, the recovery and the count read.
--form all (the default) times, from the same start, the plain FP64 form, the mixed-precision form (FP32 trailing LU,
FP64 refinement), the learned form (a surrogate trained first on calls of the plain solver over other seeds) and the
learned form with FP32 LU; value = cells x K / wall time of the K calls of the fastest form whose first timed call
agrees with the plain form's to the parity bar (species above 1e-20 within relative 1e-2, the same done flags), and
the line holds every form's figures. The roofline is the call kernel's: its algorithmic FP64 flops per launch
(per-try work counted from the network's tables, times each cell's tries, which the kernel counts) over its average
launch duration from HIP events on the launch stream. The CPU baseline times the FP64 oracle (OpenMP, one cell per
thread) on a sample of the first timed call's cells and checks the GPU's result on them.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
FP64_PEAK = 78.6  # TFLOP/s, MI355X FP64 vector
# HBM bytes the plain FP64 call kernel (its guard included) moves per cell-call on the 69-species workload: a constant
# from the FETCH_SIZE and WRITE_SIZE counters of separate profiled runs of this bench (timed calls), scaled by what
# they read for a 512 MiB copy (x2 and x1), profiles/mxc_call_ncho_65536_traffic.md; not measured in this run
TRAFFIC_PER_CELL_CALL = {"NCHO": 2048}
HEADLINE_MARGIN = 1.02  # another form heads the line only when it beats the plain one by more than this
TRAFFIC_SOURCE = "profiles/mxc_call_ncho_65536_traffic.md (rocprofv3 FETCH_SIZE, WRITE_SIZE; a constant, not this run)"


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed calls")
    ap.add_argument("--warmup", type=int, default=1, help="untimed calls of dt = 300 s after the dark relaxation")
    ap.add_argument("--cells", type=int, default=65536)
    ap.add_argument("--network", default="NCHO", choices=("NCHO", "SNCHO"))
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="leave out the CPU baseline (profiling runs)")
    ap.add_argument("--variant", default=None,
                    help="the call kernel's variant 'cells per workgroup,batch,waves per SIMD' (default: the tuned or "
                         "the default one)")
    ap.add_argument("--tune", action="store_true",
                    help="tune the call kernel's variant on a sample of settled calls first (the calls workload)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed calls, copy a tensor of this many MiB once: a kernel of known traffic that "
                         "scales hardware byte counters")
    ap.add_argument("--form", default="all", choices=("plain", "mixed", "learned", "learned_mixed", "all"),
                    help="plain FP64; mixed (FP32 trailing LU, FP64 refinement); learned (the learned warm start, FP64);"
                         " learned_mixed (learned + FP32 LU, the metric's form); all: each from the same start")
    ap.add_argument("--train-cells", type=int, default=16384, help="cells per training sequence")
    ap.add_argument("--train-sequences", type=int, default=6, help="training sequences (seeds 101..)")
    ap.add_argument("--train-epochs", type=float, default=4.0)
    ap.add_argument("--train-pairs", type=int, default=2_000_000,
                    help="plain calls the learned workload's surrogate is trained on")
    ap.add_argument("--oracle-sample", type=int, default=256, help="cells of the learned workload the oracle checks")
    ap.add_argument("--directions", type=int, default=8, help="tangent directions per jvp (--workload tangent)")
    ap.add_argument("--workload", default="calls", choices=("calls", "uneven", "coupling", "learned", "tangent"),
                    help="calls: the 65,536-cell call sequence (the metric's); uneven: 262,144 cells, one call each "
                         "from cold, half lit, dt log-uniform 1e1-1e6 s, by graph replay with the tuned variant; "
                         "coupling: the GCM-side conversion kernels on 65,536 cells, against HBM bandwidth; learned: "
                         "the learned warm start on 262,144 cold cells of the uneven case against the plain form")
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(tab, start: dict, first, cells: int, seconds: float, calls: int) -> dict:
    """The oracle over an evenly spaced sample of the first timed call's cells, and the GPU's error on them. When
    the whole first call takes less than the target time, the oracle goes on with the sequence's next timed calls
    (each from its own previous result, as the GPU's), up to the timed calls' number, and the rate is over them all."""
    import numpy as np

    sys.path.insert(0, str(ROOT))
    import oracle

    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    orc = oracle.Oracle(tab)

    def run(n: int):
        idx = np.linspace(0, cells - 1, n).round().astype(np.int64)
        a = {k: v[idx] for k, v in start.items()}
        t0 = time.perf_counter()
        out = orc.call(a["x"], a["T"], a["P"], a["dt"], a["h"], J=a["J"], threads=threads)
        return idx, out, time.perf_counter() - t0

    n = 4 * threads
    idx, out, t = run(n)
    if t < seconds / 4:
        n = int(min(cells, max(n, n * seconds / max(t, 1e-3))))
        idx, out, t = run(n)
    done_calls, total = 1, t
    if n == cells:  # the whole call: the sequence's next calls until the target time
        x, h = out["x"], out["h"]
        while total < seconds and done_calls < calls:
            t0 = time.perf_counter()
            nxt = orc.call(x, start["T"], start["P"], start["dt"], h, J=start["J"], threads=threads)
            total += time.perf_counter() - t0
            done_calls += 1
            x, h = nxt["x"], nxt["h"]
    xg, dg = first["x"][idx], first["done"][idx]
    big = out["x"] > 1e-20
    rel = float(np.max(np.abs(xg[big] / out["x"][big] - 1.0))) if big.any() else 0.0
    sh = lambda x: (x @ tab.atoms) / (x @ tab.atoms).sum(1, keepdims=True)  # noqa: E731
    share = float(np.max(np.abs(sh(xg) - sh(out["x"]))))
    sample = (f"the FP64 oracle (C, OpenMP) on the first {done_calls} timed calls of {n} cells each (dt = 300 s, each "
              f"from its own previous result), {total:.1f} s" if done_calls > 1 else
              f"the FP64 oracle (C, OpenMP) on {n} evenly spaced cells of the first timed call, dt = 300 s, {t:.1f} s")
    model = ""
    try:
        model = next(ln.split(":", 1)[1].strip() for ln in Path("/proc/cpuinfo").read_text().splitlines()
                     if ln.startswith("model name"))  # fmt: skip
    except (OSError, StopIteration):
        pass
    return {
        "value": n * done_calls / total, "unit": "cells/s", "cores": threads, "cpu_model": model, "kind": "port",
        "sample": sample,
        "gpu_vs_oracle": {"max_rel_species_above_1e-20": rel, "max_share_diff": share,
                          "done_match": bool((dg == out["done"]).all()), "cells": n},
    }  # fmt: skip


def timed(solver, x, T, P, dt, h, J, steps: int, flops_of) -> dict:
    """steps calls from (x, h), each from the previous one's result: wall time, the call kernel's events and flops."""
    import numpy as np
    import torch

    solver.timing = []
    tries, first = [], None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = solver(x, T, P, dt, h, J=J)
        x, h = r.x, r.h
        tries.append(solver.tries)
        if first is None:
            first = r
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    events, solver.timing = solver.timing, None
    kernel_ms = np.array([a.elapsed_time(b) for a, b in events])
    tries = [t.cpu().numpy() for t in tries]
    per_launch = np.array([flops_of(t) for t in tries])
    out = {"wall": wall, "kernel_ms": kernel_ms, "tries": tries, "first": first, "last": r,
           "achieved": float(per_launch.sum() / (kernel_ms.sum() * 1e-3) / 1e12)}  # fmt: skip
    fp64 = getattr(solver, "_fp64", None)
    if fp64 is not None:  # the mixed form: cells whose FP32 refinement failed (they go on in FP64)
        out["fp64_fallback_cells"] = float(fp64.float().mean())
    return out


def train_surrogate(net, plain, args, log_fn):
    """The learned form's surrogate, trained on calls of the plain solver over sequences drawn as the workload but
    from other seeds (101, 102, ...): a dark relaxation and calls of dt = 300 s, and cold calls of dt log-uniform
    1e1-1e6 s."""
    import numpy as np

    from maxichem_amd import workloads
    from maxichem_amd.learned import Surrogate, generate

    nj = len(net.photo)
    seqs = []
    for k in range(args.train_sequences):
        if k % 3 == 2:
            d = workloads.config3(net.species, nj, cells=args.train_cells, seed=101 + k)
            seqs.append({"x": d["x"], "T": d["T"], "P": d["P"], "h": d["h"], "calls": [(d["dt"], d["J"])]})
        else:
            d = workloads.config1(net.species, nj, cells=args.train_cells, seed=101 + k)
            dark = np.zeros_like(d["J"])
            seqs.append({"x": d["x"], "T": d["T"], "P": d["P"], "h": d["h"],
                         "calls": [(d["dt_relax"], dark)] + [(d["dt"], d["J"])] * 8})  # fmt: skip
    t0 = time.perf_counter()
    data = generate(plain, seqs)
    n = int(data["x0"].shape[0])
    log_fn(f"surrogate data: {n} calls in {time.perf_counter() - t0:.1f} s")
    sur = Surrogate(net, plain.dev, seed=0)
    sur.standardise(data)
    t0 = time.perf_counter()
    losses = sur.fit(data, epochs=args.train_epochs, batch=8192, lr=2e-3)
    log_fn(f"surrogate training: {time.perf_counter() - t0:.1f} s, epoch losses {[round(v, 4) for v in losses]}, "
           f"mean abs errors {sur.evaluate(data)}")  # fmt: skip
    return sur, n


def solver_drift() -> float:
    from maxichem_amd._constants import DRIFT

    return DRIFT


def run_uneven(args, net, dev) -> int:
    """The uneven batch: one call per step on the same cold inputs, replayed from a captured graph."""
    import numpy as np
    import torch

    from maxichem_amd import flops, workloads
    from maxichem_amd.kincuda import CudaSolver

    cells = args.cells if args.cells != 65536 else 262144
    d = workloads.config3(net.species, len(net.photo), cells=cells)

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    x, T, P, dt, h, J = (g(d[k]) for k in ("x", "T", "P", "dt", "h", "J"))
    from maxichem_amd import tune as _tune

    t0 = time.perf_counter()
    tuned = _tune.tune(net, dev, True, solver_drift(), workload="uneven")
    solver = CudaSolver(net, dev, photo=True, variant=tuple(tuned["best"]))
    log(f"autotuned on the uneven sample in {time.perf_counter() - t0:.1f} s: variant {solver.info['variant']}; "
        f"ms {tuned['ms']}, spilled bytes per lane {tuned['private_bytes']}")  # fmt: skip
    graph = solver.capture(x, T, P, dt, h, J=J)
    graph.replay()  # warm-up: also orders the next replay's queue by this one's tries
    torch.cuda.synchronize()
    log("warm-up replay done")
    wall, tries = [], []
    for k in range(args.steps):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r = graph.replay()
        torch.cuda.synchronize()
        wall.append(time.perf_counter() - t1)
        tries.append(graph.tries.cpu().numpy())
        log(f"replay {k + 1}/{args.steps}: {wall[-1] * 1e3:.1f} ms, done {int(r.done.sum())}/{cells}")
    total = float(np.sum(wall))
    tr = tries[-1]
    work = flops.work(solver.tab)
    # one eager call with events: the call kernel's and the recovery kernel's shares
    solver.timing, solver.timing_recover = [], []
    r = solver.launch(x, T, P, dt, h, J=J)
    torch.cuda.synchronize()
    call_ms = float(sum(a.elapsed_time(b) for a, b in solver.timing))
    rec_ms = float(sum(a.elapsed_time(b) for a, b in solver.timing_recover))
    solver.timing = None
    flag = r.flag.cpu().numpy()
    counts = solver.fallback
    log(f"eager call: call kernel {call_ms:.1f} ms, recovery {rec_ms:.1f} ms; {counts}")
    # the oracle on the cells the guard failed and the costliest ones
    sys.path.insert(0, str(ROOT))
    import oracle

    from maxichem_amd import _tables

    pick = np.unique(np.concatenate([np.flatnonzero(flag != 0), np.argsort(-tr, kind="stable")[:16]]))
    ref = oracle.Oracle(_tables.build(net, True)).call(d["x"][pick], d["T"][pick], d["P"][pick], d["dt"][pick],
                                                       d["h"][pick], J=d["J"][pick])  # fmt: skip
    xg = r.x.cpu().numpy()[pick]
    ok = flag[pick] != 2
    rb = (ref["x"] > 1e-20) & ok[:, None]
    check = {"cells": int(len(pick)), "recovered_cells": int((flag == 1).sum()), "still_failed_cells":
             np.flatnonzero(flag == 2).tolist(), "flags_match": bool(np.array_equal(flag[pick], ref["flag"])),
             "done_match": bool(np.array_equal(r.done.cpu().numpy()[pick], ref["done"])),
             "tries_match": bool(np.array_equal(solver.tries.cpu().numpy()[pick], ref["tries"])),
             "max_rel_species_above_1e-20_passed_cells": float(np.abs(xg[rb] / ref["x"][rb] - 1).max()) if rb.any()
             else 0.0}  # fmt: skip
    log(f"oracle on the recomputed and costliest cells: {check}")
    achieved = float(np.mean([flops.call_flops(solver.tab, q) for q in tries]) / (total / args.steps) / 1e12)
    line = {
        "metric": "cells/s per call (NCHO, 64K cells): learned+FP32-LU vs plain FP64 at same error",
        "value": cells * args.steps / total, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": total / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"uneven: {args.network} {cells} cells, one call from cold (h = 1e-8 s), half lit, dt "
                               "log-uniform 1e1-1e6 s; graph replays with the autotuned variant",
                   "variant": list(solver.info["variant"]), "seed": 3, "form": "plain"},
        "tries_per_cell": {"mean": float(tr.mean()), "p50": float(np.percentile(tr, 50)),
                           "p99": float(np.percentile(tr, 99)), "max": float(tr.max())},
        "tuning": tuned,
        "kernel_share": {"call_ms": call_ms, "recovery_ms": rec_ms, "recovery_share": rec_ms / (call_ms + rec_ms),
                         "counts": counts},
        "oracle_on_recomputed_and_costliest": check,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK, "traffic": None, "kernel": "mxc_call (whole replay)",
                     "flop_per_try": work.per_try},
        "cpu_baseline": None,
    }  # fmt: skip
    solver.close()
    print(json.dumps(line), flush=True)
    return 0


def learned_training_sequences(net, pairs: int, cells: int) -> list[dict]:
    """Plain-solver call sequences for the surrogate, from the GPU FP64 case's and the uneven case's distributions
    on other seeds (101, 102, ...), until about pairs calls: per config1 draw a dark relaxation and eight calls of
    dt = 300 s, per config3 draw one cold call of dt log-uniform 1e1-1e6 s, alternately."""
    import numpy as np

    from maxichem_amd import workloads

    nj = len(net.photo)
    seqs, n, k = [], 0, 0
    while n < pairs:
        if k % 2:
            d = workloads.config3(net.species, nj, cells=cells, seed=101 + k)
            seqs.append({"x": d["x"], "T": d["T"], "P": d["P"], "h": d["h"], "calls": [(d["dt"], d["J"])]})
            n += cells
        else:
            d = workloads.config1(net.species, nj, cells=cells, seed=101 + k)
            seqs.append({"x": d["x"], "T": d["T"], "P": d["P"], "h": d["h"],
                         "calls": [(d["dt_relax"], np.zeros_like(d["J"]))] + [(d["dt"], d["J"])] * 8})  # fmt: skip
            n += 9 * cells
        k += 1
    return seqs


def run_learned(args, net, dev) -> int:
    """The learned warm start at its own size: 262,144 cold cells of the uneven distribution (seed 3), one call each,
    plain FP64 against the learned forms, each call from the same cold inputs (a solver with no previous call: the
    plain queue is in index order, the learned one by predicted cost, and the learned form starts the cold cells at
    its predicted step). The surrogate is trained first on about --train-pairs plain calls of the GPU FP64 case's
    and the uneven case's distributions (other seeds). Reported: tries per cell, cells/s, speedup, the learned
    forms' error against the plain form and the oracle's on a sample."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT))
    import oracle

    from maxichem_amd import _tables, workloads
    from maxichem_amd.kincuda import CudaSolver
    from maxichem_amd.learned import LearnedSolver, Surrogate, generate

    cells = args.cells if args.cells != 65536 else 262144
    d = workloads.config3(net.species, len(net.photo), cells=cells, seed=3)

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    x, T, P, dt, h, J = (g(d[k]) for k in ("x", "T", "P", "dt", "h", "J"))
    plain = CudaSolver(net, dev, photo=True, variant=(1, 16, 1, 0))
    t0 = time.perf_counter()
    data = generate(plain, learned_training_sequences(net, args.train_pairs, args.train_cells))
    npairs = int(data["x0"].shape[0])
    log(f"surrogate data: {npairs} calls in {time.perf_counter() - t0:.1f} s")
    sur = Surrogate(net, dev, seed=0)
    sur.standardise(data)
    t0 = time.perf_counter()
    losses = sur.fit(data, epochs=args.train_epochs, batch=8192, lr=2e-3)
    train_s = time.perf_counter() - t0
    quality = sur.evaluate(data)
    log(f"surrogate training: {train_s:.1f} s, epoch losses {[round(v, 4) for v in losses]}, mean abs errors {quality}")
    del data
    forms = {"plain": plain,
             "learned": LearnedSolver(net, dev, sur, photo=True),
             "learned_proposals": LearnedSolver(net, dev, sur, photo=True, propose=True)}  # fmt: skip
    out, results = {}, {}
    for name, s in forms.items():
        walls, tries = [], None
        for k in range(args.steps):
            s.tries = None  # a cold call: no previous call's tries order the queue
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r = s.launch(x, T, P, dt, h, J=J)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t1)
            tries = s.tries.cpu().numpy()
            log(f"{name} call {k + 1}/{args.steps}: {walls[-1]:.2f} s, done {int(r.done.sum())}/{cells}, "
                f"tries mean {tries.mean():.1f}")  # fmt: skip
        results[name] = r
        wall = float(np.median(walls))
        out[name] = {"cells_per_s": cells / wall, "s_per_call": wall, "calls_timed": args.steps,
                     "tries_per_cell": {"mean": float(tries.mean()), "p50": float(np.percentile(tries, 50)),
                                        "p99": float(np.percentile(tries, 99)), "max": float(tries.max())},
                     "recomputed": s.fallback["recomputed"], "still_failed": s.fallback["still_failed"]}  # fmt: skip
    want = results["plain"]
    big = want.x > 1e-20
    for name in ("learned", "learned_proposals"):
        got = results[name]
        rel = (got.x[big] / want.x[big] - 1.0).abs()
        out[name]["vs_plain"] = {"max_rel_species_above_1e-20": float(rel.max()),
                                 "cells_beyond_1e-2": int(((got.x / want.x - 1.0).abs() * big > 1e-2).any(1).sum()),
                                 "done_match": bool(torch.equal(got.done, want.done)),
                                 "done_differ": int((got.done != want.done).sum())}  # fmt: skip
        out[name]["speedup_vs_plain"] = out["plain"]["s_per_call"] / out[name]["s_per_call"]
    # the oracle on an evenly spaced sample, against every form
    idx = np.linspace(0, cells - 1, args.oracle_sample).astype(np.int64)
    orc = oracle.Oracle(_tables.build(net, True))
    ref = orc.call(d["x"][idx], d["T"][idx], d["P"][idx], d["dt"][idx], d["h"][idx], J=d["J"][idx])
    rb = ref["x"] > 1e-20
    for name, r in results.items():
        xg = r.x.cpu().numpy()[idx]
        out[name]["vs_oracle_sample"] = {"cells": int(len(idx)),
                                         "max_rel_species_above_1e-20": float(np.abs(xg[rb] / ref["x"][rb] - 1).max()),
                                         "done_match": bool(np.array_equal(r.done.cpu().numpy()[idx], ref["done"]))}  # fmt: skip
    # the headline: the fastest form at the plain form's error (every cell's species above 1e-20 within 1e-2 of it,
    # the same done flags); the plain form itself when no learned form meets that
    meets = ["plain"] + [k for k in ("learned", "learned_proposals")
                         if out[k]["vs_plain"]["cells_beyond_1e-2"] == 0 and out[k]["vs_plain"]["done_match"]]  # fmt: skip
    best = max(meets, key=lambda k: out[k]["cells_per_s"])
    line = {
        "metric": "cells/s per call (NCHO, 64K cells): learned+FP32-LU vs plain FP64 at same error",
        "value": out[best]["cells_per_s"], "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": out[best]["s_per_call"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"learned: {args.network} {cells} cold cells of the uneven case (h = 1e-8 s, half lit, "
                               "dt log-uniform 1e1-1e6 s), one call each from cold, plain FP64 against the learned "
                               "warm start (4 x 256 GELU MLP, BF16 MFMA)", "seed": 3, "form": best,
                   "training_pairs": npairs, "training_s": train_s, "surrogate_mean_abs_error": quality},
        "forms": out,
        "roofline": None,
        "cpu_baseline": None,
    }  # fmt: skip
    for s in forms.values():
        s.close()
    print(json.dumps(line), flush=True)
    return 0


HBM_PEAK = 8000.0  # GB/s, MI355X HBM3E


def run_coupling(args, net, dev) -> int:
    """The GCM-side conversions (maxichem_amd.coupling.composition) on one cell batch: each kernel's time by HIP
    events on the stream it runs on, against its algorithmic bytes (each [cells, S] array read or written once)."""
    import numpy as np
    import torch

    from maxichem_amd import workloads
    from maxichem_amd.coupling import composition as cc

    cells, S, E = args.cells, len(net.species), len(net.elements)
    d = workloads.config1(net.species, len(net.photo), cells)
    rng = np.random.default_rng(5)
    mw = torch.from_numpy(cc.molar_masses(net.atoms, net.elements).numpy()).to(dev)
    atoms = torch.from_numpy(np.asarray(net.atoms, np.float64)).to(dev)
    x = torch.from_numpy(d["x"]).to(dev)
    mass = torch.from_numpy(rng.uniform(0.5, 1.5, cells)).to(dev)
    cell_mass = torch.from_numpy(rng.uniform(1e3, 1e6, cells)).to(dev)
    q = cc.mole_to_mass(x, mw, mass)
    kernels = {  # name: (call, algorithmic bytes per call)
        "mxc_to_mole_f64": (lambda: cc.mass_to_mole(q, mw), 16 * cells * S),
        "mxc_to_mass_f64": (lambda: cc.mole_to_mass(x, mw, mass), 16 * cells * S + 8 * cells),
        "mxc_element_moles_f64": (lambda: cc.element_moles(q, mw, atoms), 8 * cells * (S + E)),
        "mxc_element_moles_f64 + mxc_column_sums_f64 (tracer_totals)": (
            lambda: cc.tracer_totals(q, cell_mass, mw, atoms), 8 * cells * (S + 1)),
    }
    kernels["copy of [cells, S] (torch clone: the same bytes as mxc_to_mole)"] = (lambda: q.clone(), 16 * cells * S)
    out = {}
    for name, (call, nbytes) in kernels.items():
        for _ in range(max(1, args.warmup)):
            call()
        torch.cuda.synchronize()
        ms, walls = [], []
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            a.record()
            call()  # the conversions read their flags back: such a call ends synchronized
            b.record()
            b.synchronize()
            walls.append(time.perf_counter() - t0)
            ms.append(a.elapsed_time(b))
        gbs = nbytes / (float(np.median(ms)) * 1e-3) / 1e9
        out[name] = {"ms": float(np.median(ms)), "wall_ms": float(np.median(walls)) * 1e3, "bytes": nbytes,
                     "GB/s": gbs, "frac": gbs / HBM_PEAK}  # fmt: skip
        log(f"{name}: {out[name]['ms'] * 1e3:.1f} us between events (flags fill, kernel, flags read), "
            f"{out[name]['wall_ms'] * 1e3:.1f} us wall; {gbs:.0f} GB/s = {gbs / HBM_PEAK:.1%} of HBM")
    # the roofline's kernel alone between events on its stream (the calls above include the flags' fill and read)
    x_out, flags = torch.empty_like(q), cc._new_flags(dev)
    stream = torch.cuda.current_stream(dev)
    ka = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    cc._launch_to_mole(q, mw, x_out, flags, False)
    ka[0].record(stream)
    for k in range(args.steps):
        cc._launch_to_mole(q, mw, x_out, flags, False)
        ka[k + 1].record(stream)
    ka[-1].synchronize()
    kernel_ms = float(np.mean([ka[k].elapsed_time(ka[k + 1]) for k in range(args.steps)]))
    head = dict(out["mxc_to_mole_f64"])
    head["GB/s"] = 16 * cells * S / (kernel_ms * 1e-3) / 1e9
    head["frac"] = head["GB/s"] / HBM_PEAK
    log(f"mxc_to_mole_f64 alone: {kernel_ms * 1e3:.1f} us per launch, {head['GB/s']:.0f} GB/s = {head['frac']:.1%} "
        "of HBM")  # fmt: skip
    line = {
        "metric": "GB/s (GCM-side conversions, NCHO, 64K cells)", "value": head["GB/s"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": kernel_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"coupling: {args.network} {cells} cells x {S} species, mass to mole fractions and "
                               "back, element moles, totals"},
        "kernels": out,
        "roofline": {"bound": "hbm", "achieved": head["GB/s"], "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": head["frac"], "traffic": None, "kernel": "mxc_to_mole_f64"},
        "cpu_baseline": None,
    }  # fmt: skip
    print(json.dumps(line), flush=True)
    return 0


def run_tangent(args, net, dev) -> int:
    """ADSolver.jvp with --directions random state directions per call (adaptive contract, the reference's default),
    timed by HIP events around each jvp on its stream (the tangent kernel, the recovery of listed cells and the count
    read), on the bench's cells after the dark relaxation and one lit call; beside it the plain call on the same
    inputs, and for at most 8,192 cells the vjp by columns without the rate cotangents."""
    import numpy as np
    import torch

    from maxichem_amd import workloads
    from maxichem_amd.kincuda import CudaSolver
    from maxichem_amd.kincuda_ad import ADSolver

    B, D = args.cells, args.directions
    d = workloads.config1(net.species, len(net.photo), B)

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    def events(call, n):
        ms = []
        for _ in range(n):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            call()
            b.record()
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.mean(ms))

    T, P, dt, J = g(d["T"]), g(d["P"]), g(d["dt"]), g(d["J"])
    with CudaSolver(net, dev, photo=True) as s:
        r = s(g(d["x"]), T, P, g(d["dt_relax"]), g(d["h"]), J=torch.zeros_like(J))
        r = s(r.x, T, P, dt, r.h, J=J)
        x, h = r.x, r.h
        call_ms = events(lambda: s(x, T, P, dt, h, J=J), max(1, args.steps))
    rng = np.random.default_rng(0)
    dx = g(rng.standard_normal((D, B, len(net.species))) * 1e-3) * x
    z = torch.zeros(D, B, dtype=torch.float64, device=dev)
    with ADSolver(net, dev, photo=True) as ad:
        log(f"tangent solver: {ad.info}; tangent kernel LDS {ad.lds_jvp} B per cell, {ad.max_blocks_jvp} cells at once")

        def jvp():
            ad.jvp(x, T, P, dt, h, dx, z, z, z, z, J=J)

        for _ in range(max(1, args.warmup)):
            jvp()
        jvp_ms = events(jvp, max(1, args.steps))
        vjp_ms = None
        if B <= 8192:
            gx = g(rng.standard_normal((B, len(net.species))))
            vjp_ms = events(lambda: ad.vjp(x, T, P, dt, h, gx, J=J, rates=False), 1)
    log(f"{B} cells: call {call_ms:.1f} ms, jvp with {D} directions {jvp_ms:.1f} ms ({jvp_ms / D:.1f} ms a "
        f"direction)" + ("" if vjp_ms is None else f", vjp without the rate cotangents {vjp_ms:.1f} ms"))  # fmt: skip
    line = {
        "metric": f"cell-directions/s of the call's tangent ({args.network}, {B} cells, {D} directions)",
        "value": B * D / (jvp_ms * 1e-3), "unit": "cell-directions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": jvp_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"tangent: {args.network} {B} cells after the dark relaxation and one lit call of "
                               f"dt = 300 s, jvp along {D} random state directions (adapt=True)"},
        "call_ms": call_ms, "jvp_ms": jvp_ms, "vjp_ms": vjp_ms, "roofline": None, "cpu_baseline": None,
    }  # fmt: skip
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU: the cells of a call are independent and this build measures one GPU's "
              f"share (asked for --gpus {args.gpus}, WORLD_SIZE {world})", file=sys.stderr)  # fmt: skip
        return 2
    import numpy as np
    import torch

    from maxichem_amd import flops, standin, workloads
    from maxichem_amd.kincuda import CudaSolver

    dev = torch.device("cuda", 0)
    net = standin.real_or_standin(args.network)
    nj = len(net.photo)
    if args.workload == "uneven":
        return run_uneven(args, net, dev)
    if args.workload == "coupling":
        return run_coupling(args, net, dev)
    if args.workload == "learned":
        return run_learned(args, net, dev)
    if args.workload == "tangent":
        return run_tangent(args, net, dev)
    d = workloads.config1(net.species, nj, args.cells)
    log(f"network {args.network}: {len(net.species)} species, {len(net.reac)} thermal, {nj} photolysis; "
        f"{args.cells} cells")  # fmt: skip

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    if args.variant:
        variant = tuple(int(q) for q in args.variant.split(","))
    else:  # the autotuner's choice for these settled calls (tuned now with --tune), else the default
        from maxichem_amd import tune as _tune

        variant = _tune.choose(net, dev, True, solver_drift(), run=args.tune, workload="settled")
    solver = CudaSolver(net, dev, photo=True, variant=variant)
    log(f"solver: {solver.info}")
    x, T, P, h, J = g(d["x"]), g(d["T"]), g(d["P"]), g(d["h"]), g(d["J"])
    dt_relax, dt = g(d["dt_relax"]), g(d["dt"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = solver(x, T, P, dt_relax, h, J=torch.zeros_like(J))
    torch.cuda.synchronize()
    log(f"dark relaxation dt = 1e8 s: {time.perf_counter() - t0:.2f} s, done {int(r.done.sum())}/{args.cells}, "
        f"fallback {solver.fallback}")  # fmt: skip
    x, h = r.x, r.h
    for w in range(args.warmup):
        r = solver(x, T, P, dt, h, J=J)
        x, h = r.x, r.h
        log(f"warmup call {w + 1}/{args.warmup} done")
    start = {"x": x, "T": T, "P": P, "dt": dt, "h": h, "J": J}
    start_np = {k: v.cpu().numpy() for k, v in start.items()}
    work = flops.work(solver.tab)

    def flops_of(t):
        return flops.call_flops(solver.tab, t)

    runs = {}
    forms = ("plain", "mixed", "learned", "learned_mixed") if args.form == "all" else (args.form,)
    if "plain" in forms:
        runs["plain"] = timed(solver, x, T, P, dt, h, J, args.steps, flops_of)
    if "mixed" in forms:
        mixed = CudaSolver(net, dev, photo=True, precision="mixed")
        mixed(x, T, P, dt, h, J=J)  # warm-up (its first call from the same start)
        runs["mixed"] = timed(mixed, x, T, P, dt, h, J, args.steps, flops_of)
        mixed.close()
    ndata = 0
    if "learned" in forms or "learned_mixed" in forms:
        from maxichem_amd.learned import LearnedSolver

        sur, ndata = train_surrogate(net, solver, args, log)
        for name, precision in (("learned", "fp64"), ("learned_mixed", "mixed")):
            if name in forms:
                ls = LearnedSolver(net, dev, sur, photo=True, precision=precision)
                ls(x, T, P, dt, h, J=J)  # warm-up of the learned path (its first call from the same start)
                runs[name] = timed(ls, x, T, P, dt, h, J, args.steps, flops_of)
                ls.close()
    for name, rr in runs.items():
        log(f"{name}: {len(rr['tries'])} timed calls {rr['wall']:.3f} s; call kernel {rr['kernel_ms'].mean():.3f} ms "
            f"average; mean tries per cell-call {np.mean([t.mean() for t in rr['tries']]):.2f}; last call done "
            f"{int(rr['last'].done.sum())}/{args.cells}")  # fmt: skip

    # every form against the plain one on the same start (its first timed call): the parity bar (species above
    # 1e-20 within relative 1e-2, done flags equal)
    agree = {}
    if "plain" in runs:
        a = runs["plain"]["first"]
        big = a.x > 1e-20
        sh = lambda x: (x @ solver._tables["atoms"]) / (x @ solver._tables["atoms"]).sum(1, keepdim=True)  # noqa: E731
        for name, rr in runs.items():
            if name != "plain":
                b = rr["first"]
                agree[name] = {"max_rel_species_above_1e-20": float((b.x[big] / a.x[big] - 1.0).abs().max()),
                               "max_share_diff": float((sh(a.x) - sh(b.x)).abs().max()),
                               "done_match": bool(torch.equal(a.done, b.done))}  # fmt: skip
    # the headline: the fastest form at the plain form's error
    ok = {n for n in runs if n == "plain" or (n in agree and agree[n]["done_match"]
                                              and agree[n]["max_rel_species_above_1e-20"] <= 1e-2)}  # fmt: skip
    if not ok:  # no plain form timed (--form learned, ...): nothing to check against; the fastest form measured
        ok = set(runs)
    per_call = {n: runs[n]["wall"] / len(runs[n]["tries"]) for n in runs}
    headline = min(ok, key=per_call.get)
    if "plain" in runs and per_call[headline] > per_call["plain"] / HEADLINE_MARGIN:
        headline = "plain"  # a form within the run-to-run spread of the plain one does not replace it
    main_run = runs[headline]
    # the constant measures the FP64 call kernel, which the plain and learned forms share (the learned form changes
    # only the starting h); the mixed forms run another kernel
    fp64_call = headline in ("plain", "learned") and args.network in TRAFFIC_PER_CELL_CALL
    line = {
        "metric": "cells/s per call (NCHO, 64K cells): learned+FP32-LU vs plain FP64 at same error",
        "value": args.cells * len(main_run["tries"]) / main_run["wall"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": len(main_run["tries"]),
        "warmup": args.warmup,
        "ms_per_step": main_run["wall"] / len(main_run["tries"]) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.network} {args.cells} cells: dark relaxation dt = 1e8 s (untimed), then calls of "
                        "dt = 300 s carrying x and h, J log-uniform 1e-9..1e-3 s^-1 on the lowest-P quarter",
            "network": f"{'stand-in ' if getattr(net, 'standin', True) else ''}{args.network}: "
                       f"{len(net.species)} species, {len(net.reac)} thermal, {nj} photolysis",
            "cells": args.cells, "seed": 1, "form": headline, "forms_timed": list(runs), "rtol": 1e-3, "atol": 1e-25,
        },
        "forms": {},
        "roofline": {
            "bound": "fp64_valu",
            "compute": "FP64 vector ALU (VALU; MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s, and the "
                       "kernel issues no MFMA)",
            "achieved": main_run["achieved"], "peak": FP64_PEAK, "unit": "TFLOP/s",
            "frac": main_run["achieved"] / FP64_PEAK,
            "traffic": TRAFFIC_PER_CELL_CALL[args.network] * args.cells if fp64_call else None,
            "traffic_source": TRAFFIC_SOURCE if fp64_call else None,
            "kernel": "mxc_call", "kernel_ms": float(main_run["kernel_ms"].mean()),
            "flop_per_try": work.per_try, "flop_rates_per_cell": work.rates,
            "tries_per_launch": float(np.mean([t.sum() for t in main_run["tries"]])),
        },  # fmt: skip
        "cpu_baseline": None,
    }
    for name, rr in runs.items():
        k = len(rr["tries"])
        line["forms"][name] = {
            "cells_per_s": args.cells * k / rr["wall"], "ms_per_call": rr["wall"] / k * 1e3, "calls_timed": k,
            "kernel_ms": float(rr["kernel_ms"].mean()), "fp64_tflops": rr["achieved"],
            "frac_fp64_peak": rr["achieved"] / FP64_PEAK,
            "tries_per_cell_call": float(np.mean([t.mean() for t in rr["tries"]])),
        }  # fmt: skip
        if "fp64_fallback_cells" in rr:
            line["forms"][name]["fp64_fallback_cells"] = rr["fp64_fallback_cells"]
            line["forms"][name]["fp32_converged_cells"] = 1.0 - rr["fp64_fallback_cells"]
    for name, a in agree.items():
        line["forms"][name]["vs_plain_first_call"] = a
        line["forms"][name]["speedup_vs_plain"] = (runs["plain"]["wall"] / len(runs["plain"]["tries"])) / (
            runs[name]["wall"] / len(runs[name]["tries"]))
    if ndata:
        line["forms"]["training_calls"] = ndata
    if args.copy_mib:
        src = torch.empty(args.copy_mib << 17, dtype=torch.float64, device=dev).fill_(1.0)
        dst = src.clone()
        torch.cuda.synchronize()
        line["copy_probe_mib"] = args.copy_mib
        del src, dst
    if not args.no_cpu:
        first = runs["plain"]["first"] if "plain" in runs else main_run["first"]
        first_np = {"x": first.x.cpu().numpy(), "done": first.done.cpu().numpy()}
        log("CPU baseline: the oracle on a sample of the first timed call")
        line["cpu_baseline"] = cpu_baseline(solver.tab, start_np, first_np, args.cells, args.cpu_seconds, args.steps)
    solver.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
