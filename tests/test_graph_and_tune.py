"""Graph-captured calls and the autotuner (GPU): a replay equals an eager call on the same inputs bit for bit, with
eager work between replays and new input contents; refusals come through a replay; every kernel variant gives the
same bits, and the tuned choice is stored and taken."""

import numpy as np
import pytest
import torch

from maxichem_amd import standin, tune, workloads
from maxichem_amd.kincuda import CudaSolver

pytestmark = pytest.mark.gpu


def t(a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


@pytest.fixture(scope="module")
def ncho():
    if not torch.cuda.is_available():
        pytest.skip("this test needs a GPU")
    return standin.load("NCHO")


def inputs(net, cells, seed):
    d = workloads.config1(net.species, len(net.photo), cells=cells, seed=seed)
    return [t(d[k]) for k in ("x", "T", "P")] + [t(np.full(cells, 50.0)), t(np.full(cells, 1e-6)), t(d["J"])]


def same(a, b):
    return all(torch.equal(p, q) for p, q in zip(a, b))


def test_replays_equal_eager_calls(ncho):
    with CudaSolver(ncho, "cuda", photo=True) as solver:
        x, T, P, dt, h, J = inputs(ncho, 2048, 5)
        graph = solver.capture(x, T, P, dt, h, J=J)
        for k in range(3):
            if k:  # new contents in the same tensors, and eager work between replays
                x2, T2, P2, _, _, J2 = inputs(ncho, 2048, 5 + k)
                for a, b in ((x, x2), (T, T2), (P, P2), (J, J2)):
                    a.copy_(b)
                other = inputs(ncho, 512, 40 + k)
                solver(*other[:5], J=other[5])
                torch.cuda.synchronize()
            before = solver.total.clone()
            got = graph.replay()
            got = tuple(v.clone() for v in got)
            counted = solver.total - before
            want = solver.launch(x, T, P, dt, h, J=J)
            print(f"replay {k}: done {int(got[2].sum())}, steps {float(got[3].sum()):.0f}, counts {counted.tolist()}")
            assert same(got, want)
            assert torch.equal(counted, solver.count.to(torch.int64))


def test_a_replay_refuses_as_the_eager_call(ncho):
    with CudaSolver(ncho, "cuda", photo=True) as solver:
        x, T, P, dt, h, J = inputs(ncho, 256, 9)
        graph = solver.capture(x, T, P, dt, h, J=J)
        graph.replay()
        x[3, 5] = float("nan")
        with pytest.raises(ValueError, match=r"1 cells of bad input: .*\(first: cell 3: sum\(x\) - 1 = nan"):
            graph.replay()
        assert solver.flag is None
        x[3, 5] = 1e-30
        graph.replay()
        assert solver.fallback["bad_input"] == 0


def test_every_variant_gives_the_same_bits(ncho):
    x, T, P, dt, h, J = inputs(ncho, 1024, 3)
    results = {}
    for v in ((1, 16, 1, 0), (1, 8, 1, 0), (2, 4, 1, 0), (4, 16, 1, 0), (1, 16, 2, 0), (1, 16, 1, 1), (1, 16, 2, 1)):
        with CudaSolver(ncho, "cuda", photo=True, variant=v) as solver:
            results[v] = solver.launch(x, T, P, dt, h, J=J)
            assert solver.info["variant"] == v
    first = results[(1, 16, 1, 0)]
    assert all(same(first, r) for r in results.values())


def test_the_tuned_choice_is_stored_and_taken(ncho, tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "STORE", tmp_path / "tune.json")  # a scratch store
    out = tune.tune(ncho, torch.device("cuda"), True, 1e-6, cells=2048)
    print(f"autotuner: {out}")
    assert tuple(out["best"]) in tune.VARIANTS and len(out["ms"]) == len(tune.VARIANTS)
    assert tuple(tune.stored(ncho, torch.device("cuda"), True)["best"]) == tuple(out["best"])
    with CudaSolver(ncho, "cuda", photo=True) as solver:
        assert solver.info["variant"] == tuple(out["best"])
    # the settled workload's choice is stored apart, and taken when that workload is named
    settled = tune.tune(ncho, torch.device("cuda"), True, 1e-6, cells=2048, workload="settled")
    print(f"autotuner, settled calls: {settled}")
    assert settled["workload"] == "settled" and set(settled["private_bytes"]) == set(settled["ms"])
    assert tune.choose(ncho, torch.device("cuda"), True, 1e-6, workload="settled") == tuple(settled["best"])
    assert tune.choose(ncho, torch.device("cuda"), True, 1e-6) == tuple(out["best"])
    with pytest.raises(ValueError, match="workload"):
        tune.tune(ncho, torch.device("cuda"), True, 1e-6, cells=64, workload="busy")
