"""The HIP path at the bench's sizes (GPU), through what does not depend on the batch size.

Two batches: 65,536 cells of the GPU FP64 case (the dark relaxation of dt = 1e8 s, then a lit call of dt = 300 s)
and 262,144 cells of the uneven case (cold cells, half lit, dt log-uniform 1e1-1e6 s). The x row of every cell the
guard passed (flag 0 or 1) is non-negative and sums to 1, and its element shares move by at most the parity bar's
1e-10 relative; a cell the guard failed (flag 2) is not done. The whole batch, shuffled, gives every cell the same
bits. An evenly spaced sample of 256 cells, the 64 costliest cells and every cell the guard failed, against the
oracle, meet the parity bar (species above 1e-20 within relative 1e-2, the same done flags) and have the oracle's
flags; on the even sample the accepted steps and tries are the oracle's (all but rare cells where rounding tips a
try) and then so is the next step size, to 1e-2, on all but a tail.
"""

import numpy as np
import oracle
import pytest
import torch

from maxichem_amd import _tables, standin, workloads
from maxichem_amd.kincuda import CudaSolver

pytestmark = pytest.mark.gpu
REL = 1e-2
FLOOR = 1e-20
SHARE_DRIFT = 1e-10
SAMPLE = 256
COSTLY = 64  # the cells of most tries, compared with the oracle beside the even sample
STEPS_SAME = 0.9  # share of the sample whose accepted steps and tries are the oracle's
H_SAME = 0.95  # share of those whose next step size is the oracle's to REL (the last error estimate is noisy where
# trace species sit near atol)


def t(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def check_batch(name, solver, orc, atoms, inp, dt, J):
    cells = inp["x"].shape[0]
    args = [t(inp["x"]), t(inp["T"]), t(inp["P"]), t(dt), t(inp["h"])]
    Jg = t(J)
    r = solver.launch(*args, J=Jg)
    x, done, flag = r.x.cpu().numpy(), r.done.cpu().numpy(), r.flag.cpu().numpy()
    tries = solver.tries.cpu().numpy()  # this call's, before the shuffled call below
    passed = flag != 2
    assert (x[passed] >= 0.0).all() and np.abs(x[passed].sum(1) - 1.0).max() <= 1e-12
    e0, e1 = inp["x"] @ atoms, x @ atoms
    s0, s1 = e0 / e0.sum(1, keepdims=True), e1 / e1.sum(1, keepdims=True)
    drift = np.where(s0 > 1e-12, np.abs(s1 - s0) / np.maximum(s0, 1e-300), 0.0).max(1)
    assert drift[passed].max() <= SHARE_DRIFT
    assert not done[~passed].any()
    # the same cells in another order: the same bits for each
    perm = torch.randperm(cells, generator=torch.Generator().manual_seed(5)).cuda()
    s = solver.launch(*(a[perm] for a in args), J=Jg[perm])
    inverse = torch.empty_like(perm)
    inverse[perm] = torch.arange(cells, device="cuda")
    assert all(torch.equal(p, q[inverse]) for p, q in zip(r, s))
    # an evenly spaced sample against the oracle, with the costliest cells and every cell the guard failed
    idx = np.linspace(0, cells - 1, SAMPLE).astype(np.int64)
    costly = np.argsort(-tries, kind="stable")[:COSTLY]
    failed = np.flatnonzero(flag == 2)
    extra = np.setdiff1d(np.union1d(costly, failed), idx)
    both = np.concatenate([idx, extra])
    want = orc.call(inp["x"][both], inp["T"][both], inp["P"][both], dt[both], inp["h"][both], J=J[both])
    big = want["x"] > FLOOR
    rel = np.abs(x[both][big] / want["x"][big] - 1.0)
    steps, h = r.steps.cpu().numpy(), r.h.cpu().numpy()
    same = (steps[idx] == want["steps"][:SAMPLE]) & (tries[idx] == want["tries"][:SAMPLE])
    hrel = np.abs(h[idx][same] / want["h"][:SAMPLE][same] - 1.0)
    print(f"{name}: {cells} cells, done {int(done.sum())}, flagged {int((flag == 2).sum())}, share drift "
          f"{drift[passed].max():.1e}; tries max {tries.max():.0f}; {both.size} cells against the oracle (even "
          f"sample {SAMPLE}, the {COSTLY} costliest, flagged {failed.size}): max rel {rel.max():.2e}, flags "
          f"{flag[both].tolist().count(2)}/{want['flag'].tolist().count(2)} flagged; even sample: steps equal on "
          f"{int(same.sum())} with the same tries, h rel max {hrel.max() if hrel.size else 0.0:.2e}")  # fmt: skip
    assert np.array_equal(done[both].astype(np.int32), want["done"])
    assert np.array_equal(flag[both], want["flag"])
    assert rel.max() <= REL
    assert same.mean() >= STEPS_SAME and (hrel.size == 0 or (hrel <= REL).mean() >= H_SAME)
    return x, h


def test_gpu_fp64_case_at_65536_cells():
    net = standin.load("NCHO")
    inp = workloads.config1(net.species, len(net.photo), cells=65536, seed=1)
    orc = oracle.Oracle(_tables.build(net, True))
    atoms = np.asarray(net.atoms, np.float64)
    with CudaSolver(net, "cuda", photo=True) as solver:
        x, h = check_batch("relaxation", solver, orc, atoms, inp, inp["dt_relax"], np.zeros_like(inp["J"]))
        lit = dict(inp, x=x, h=h)
        check_batch("lit call", solver, orc, atoms, lit, inp["dt"], inp["J"])


def test_uneven_case_at_262144_cells():
    net = standin.load("NCHO")
    inp = workloads.config3(net.species, len(net.photo), cells=262144, seed=3)
    orc = oracle.Oracle(_tables.build(net, True))
    atoms = np.asarray(net.atoms, np.float64)
    with CudaSolver(net, "cuda", photo=True) as solver:
        check_batch("uneven", solver, orc, atoms, inp, inp["dt"], inp["J"])
