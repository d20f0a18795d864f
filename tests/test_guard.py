"""The guard, the recovery and the refusals on the GPU: counts, flags, messages and what the other cells get.

Each case follows one of the reference's guard tests (maxichem tests/test_guard.py:26-424) in behaviour: what is
refused, how it is counted and named, and that the cells left alone are bitwise what they are in a batch without the
refused ones. Where the recovery recomputes a cell, its result is compared with the oracle's recovery of that cell.
"""

import ctypes

import numpy as np
import oracle
import pytest
import torch

from maxichem_amd import _hip, _kernels, _tables, standin, workloads
from maxichem_amd._constants import T_MAX, T_MIN, X_SUM_TOL
from maxichem_amd.kincuda import CallResult, CudaSolver, LaunchResult

pytestmark = pytest.mark.gpu


def t(a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


def full(n, v):
    return torch.full((n,), float(v), dtype=torch.float64, device="cuda")


@pytest.fixture(scope="module")
def ncho():
    if not torch.cuda.is_available():
        pytest.skip("this test needs a GPU")
    net = standin.load("NCHO")
    with CudaSolver(net, "cuda") as solver:
        yield net, solver


def starts(net, solver, cells, seed, settle=100.0):
    """Solar gas states at seeded T and P, settled by a call of `settle` seconds (its x and h)."""
    inp = workloads.config0(net.species, 0, cells=cells, seed=seed)
    x, T, P = t(inp["x"]), t(inp["T"]), t(inp["P"])
    r = solver(x, T, P, full(cells, settle), full(cells, 1e-8))
    return r.x, T, P, r.h


def same(a, b):
    return all(torch.equal(p, q) for p, q in zip(a, b))


def test_nan_input_is_refused(ncho, monkeypatch):
    net, solver = ncho
    cells = 64
    x, T, P, h = starts(net, solver, cells, 1)
    x[3, 5] = float("nan")
    dt = full(cells, 1.0)
    with pytest.raises(ValueError, match=r"1 cells of bad input: .*\(first: cell 3: sum\(x\) - 1 = nan"):
        solver(x, T, P, dt, h)
    assert solver.flag is None
    with monkeypatch.context() as m:  # what the kernels returned
        m.setattr(solver, "_refuse", lambda *_: None)
        y, _, d, _ = solver(x, T, P, dt, h)
    counts = solver.fallback
    keep = torch.arange(cells, device="cuda") != 3
    alone = solver(x[keep], T[keep], P[keep], dt[keep], h[keep])
    print(f"counts {counts}")
    assert counts["bad_input"] == 1 and counts["recomputed"] == 0
    assert not bool(d[3]) and not bool(torch.isfinite(y[3]).all())
    assert torch.equal(y[keep], alone.x) and bool(d[keep].all())


def test_forced_failure_is_recomputed_as_the_oracle_does(ncho):
    net, solver = ncho
    cells = 64
    x, T, P, h = starts(net, solver, cells, 1)
    inp = workloads.config0(net.species, 0, cells=cells, seed=1)
    dt = full(cells, 1.0)
    x[7], h[7], dt[7] = t(inp["x"][7]), 1e-10, 500.0  # cell 7: a cold start over 500 s in 10 tries
    r = solver.launch(x, T, P, dt, h, max_steps=10)
    counts = solver.fallback
    print(f"counts {counts}; cell 7 flag {int(r.flag[7])} done {bool(r.done[7])} steps {float(r.steps[7])}")
    assert counts["unfinished"] == 1 and counts["recomputed"] == 1
    assert counts["recovered"] + counts["still_failed"] == 1
    assert int(r.flag[7]) == (1 if bool(r.done[7]) else 2)
    others = torch.arange(cells, device="cuda") != 7
    assert bool(r.done[others].all()) and int(r.flag[others].abs().sum()) == 0
    orc = oracle.Oracle(_tables.build(net, False))
    c = [7]
    want = orc.call(x[c].cpu().numpy(), T[c].cpu().numpy(), P[c].cpu().numpy(), dt[c].cpu().numpy(),
                    h[c].cpu().numpy(), max_steps=10)  # fmt: skip
    got = r.x[7].cpu().numpy()
    big = want["x"][0] > 1e-20
    rel = np.abs(got[big] / want["x"][0][big] - 1.0).max()
    print(f"recovered cell against the oracle's recovery: max rel {rel:.2e}, steps {want['steps'][0]}, "
          f"flag {want['flag'][0]}")  # fmt: skip
    assert int(want["flag"][0]) == int(r.flag[7]) and bool(want["done"][0]) == bool(r.done[7])
    assert rel <= 1e-2


def test_element_drift_is_recomputed(ncho):
    # allowance 0 flags every rounding change of an element share; atol 1e-18 keeps the clamping allowance small
    net, solver = ncho
    cells = 1024
    inp = workloads.config0(net.species, 0, cells=cells, seed=2)
    args = (t(inp["x"]), t(inp["T"]), t(inp["P"]), full(cells, 500.0), full(cells, 1e-10))
    with CudaSolver(net, "cuda", drift=0.0) as strict:
        y, _, dz, _, flag = strict.launch(*args, atol=1e-18)
        counts = strict.fallback
    yk, _, _, _, flag0 = solver.launch(*args, atol=1e-18)
    passed = flag == 0
    print(f"allowance 0: counts {counts}; default allowance flags {int(flag0.sum())}")
    assert counts["drift"] > 0 and counts["recomputed"] == counts["drift"]
    assert counts["recovered"] + counts["still_failed"] == counts["drift"]
    assert int(flag0.sum()) == 0
    assert torch.equal(y[passed], yk[passed])
    assert not bool(dz[flag == 2].any())


def test_temperature_out_of_range_is_refused(ncho, monkeypatch):
    net, solver = ncho
    cells = 256
    x, T, P, h = starts(net, solver, cells, 2)
    T = T.clone()
    T[0], T[1] = 250.0, T_MAX + 0.5
    T[2], T[3] = T_MIN, T_MAX  # the ends of the range are accepted
    args = (x, T, P, full(cells, 500.0), full(cells, 1e-10))
    refused = r"1 cells below 300 K, .*\(first: cell 0, T = 250 K\).*; 1 cells above 3000 K, .*cell 1, T = 3000.5 K\)"
    with CudaSolver(net, "cuda") as fresh:
        assert fresh.check()["cold"] == 0
        with pytest.raises(ValueError, match=refused):
            fresh(*args)
        with pytest.raises(ValueError, match="1 cells below 300 K and 1 above 3000 K"):
            fresh.check()
    with monkeypatch.context() as m:
        m.setattr(solver, "_refuse", lambda *_: None)
        _, _, d, _ = solver(*args)
    counts = solver.fallback
    assert counts["cold"] == 1 and counts["hot"] == 1
    assert not bool(d[:2].any()) and bool(d[2:4].all())


def test_bad_inputs_are_refused_and_the_rest_untouched(ncho, monkeypatch):
    net, solver = ncho
    cells = 64
    inp = workloads.config0(net.species, 0, cells=cells, seed=1)
    xb, T, Pb = t(inp["x"]), t(inp["T"]), t(inp["P"])
    dtb, hb = full(cells, 50.0), full(cells, 1e-10)
    xb[3] = 0.0
    xb[7, 0] = -1e-3
    Pb[5] = 0.0
    dtb[9] = -1.0
    xb[11] *= 0.9
    xb[13] *= 1.0 + 0.9 * X_SUM_TOL  # just inside the tolerance on either side: accepted
    xb[15] *= 1.0 - 0.9 * X_SUM_TOL
    bad = [3, 5, 7, 9, 11]
    with pytest.raises(ValueError, match=r"5 cells of bad input: .*\(first: cell 3: sum\(x\) - 1 = -1 "):
        solver(xb, T, Pb, dtb, hb)
    with pytest.raises(ValueError, match=r"1 cells of bad input: .*\(first: cell 0: sum\(x\) - 1 = -0.1 "):
        solver(xb[11:12], T[11:12], Pb[11:12], dtb[11:12], hb[11:12])
    with monkeypatch.context() as m:
        m.setattr(solver, "_refuse", lambda *_: None)
        y, _, d, _ = solver(xb, T, Pb, dtb, hb)
    counts = solver.fallback
    good = torch.ones(cells, dtype=torch.bool, device="cuda")
    good[bad] = False
    yr, _, dr, _ = solver(xb[good], T[good], Pb[good], dtb[good], hb[good])
    print(f"counts {counts}")
    assert counts["bad_input"] == len(bad) and counts["nonfinite"] == 0 and counts["recomputed"] == 0
    assert not bool(d[bad].any()) and bool(d[[13, 15]].all())
    assert torch.equal(y[good], yr) and torch.equal(d[good], dr)


def test_non_finite_and_overflowing_pressures_are_named(ncho):
    net, solver = ncho
    inp = workloads.config0(net.species, 0, cells=4, seed=1)
    P = t(inp["P"])
    P[1:] = t([np.inf, np.nan, 1e300])
    named = r"^the call refused cells: 3 cells of bad input: .*P non-finite, not above 0 or so large that P / \(k_B T\)"
    with pytest.raises(ValueError, match=rf"{named} overflows.*first: cell 1:"):
        solver(t(inp["x"]), t(inp["T"]), P, full(4, 50.0), full(4, 1e-10))
    assert solver.fallback["bad_input"] == 3


def test_one_step_outside_the_range_is_refused(ncho, monkeypatch):
    net, solver = ncho
    inp = workloads.config0(net.species, 0, cells=4, seed=2)
    cold, hot = np.nextafter(T_MIN, 0.0), np.nextafter(T_MAX, np.inf)
    T = t([cold, hot, cold, hot])
    dt, h = t([0.0, 0.0, 500.0, 500.0]), full(4, 1e-10)
    x, P = t(inp["x"]), t(inp["P"])
    refused = r"2 cells below 300 K, .*\(first: cell 0, T = 300 K\).*; 2 cells above 3000 K, .*\(first: cell 1, "
    with pytest.raises(ValueError, match=refused):
        solver(x, T, P, dt, h)
    with monkeypatch.context() as m:
        m.setattr(solver, "_refuse", lambda *_: None)
        _, _, done, _ = solver(x, T, P, dt, h)
    counts = solver.fallback
    assert counts["cold"] == 2 and counts["hot"] == 2 and not bool(done.any())


def test_non_finite_step_acts_as_dt(ncho):
    net, solver = ncho
    cells = 64
    inp = workloads.config0(net.species, 0, cells=cells, seed=4)
    x, T, P = t(inp["x"]), t(inp["T"]), t(inp["P"])
    dt = full(cells, 50.0)
    ref = solver(x, T, P, dt, dt.clone())
    assert solver.fallback["nonfinite_h"] == 0
    for q in (np.nan, -np.inf, np.inf):
        out = solver(x, T, P, dt, full(cells, q))
        assert same(out, ref), q
        assert solver.fallback["nonfinite_h"] == cells


def test_zero_dt_leaves_the_state(ncho):
    net, solver = ncho
    cells = 16
    x, T, P, h = starts(net, solver, cells, 5)
    r = solver.launch(x, T, P, full(cells, 0.0), h)
    s = x / x.sum(1, keepdim=True)
    assert bool(r.done.all()) and int(r.steps.sum()) == 0
    assert torch.allclose(r.x, s, rtol=1e-12, atol=0.0)


def test_launch_adds_to_the_total(ncho):
    net, solver = ncho
    x, T, P, h = starts(net, solver, 8, 2)
    before = solver.total.clone()
    solver.launch(x, T, P, full(8, 1.0), full(8, 1.0))
    assert torch.equal(solver.total - before, solver.count.to(torch.int64))


def test_call_and_launch_name_their_results(ncho):
    net, solver = ncho
    inp = workloads.config0(net.species, 0, cells=8, seed=2)
    args = (t(inp["x"]), t(inp["T"]), t(inp["P"]), full(8, 50.0), full(8, 1e-10))
    launched = solver.launch(*args, max_steps=2)  # few tries: the guard recomputes cells
    assert isinstance(launched, LaunchResult) and solver.flag is launched.flag
    assert bool((launched.flag != 0).any())
    called = solver(*args, max_steps=2)
    assert isinstance(called, CallResult)
    assert all(torch.equal(getattr(called, k), getattr(launched, k)) for k in called._fields)
    assert torch.equal(solver.flag, launched.flag)


def test_flag_is_none_after_a_call_that_raised(ncho):
    net, solver = ncho
    x, T, P, h = starts(net, solver, 8, 2)
    dt = full(8, 1.0)
    bad = x.clone()
    bad[3, 0] = float("nan")
    solver(x, T, P, dt, h)
    with pytest.raises(ValueError, match="refused cells"):
        solver.launch(bad, T, P, dt, h)
    assert solver.flag is None
    solver(x, T, P, dt, h)
    with pytest.raises(ValueError, match=r"^T: "):  # on the host, before any launch
        solver(x[:5], T[:5].cpu(), P[:5], dt[:5], h[:5])
    assert solver.flag is None


def test_wrong_inputs_raise_before_any_launch(ncho):
    net, solver = ncho
    x, T, P, h = starts(net, solver, 8, 2)
    dt = full(8, 50.0)
    with pytest.raises(ValueError, match=r"^T: .* on cpu$"):
        solver(x, T.cpu(), P, dt, h)
    with pytest.raises(ValueError, match=r"^max_steps from 1"):
        solver(x, T, P, dt, h, max_steps=0)
    with pytest.raises(ValueError, match=r"^J: none or an empty"):
        solver(x, T, P, dt, h, J=torch.zeros(8, 3, dtype=torch.float64, device="cuda"))
    with CudaSolver(net, "cuda", photo=True) as lit:
        with pytest.raises(ValueError, match="required"):
            lit(x, T, P, dt, h)


def test_closed_solver_refuses_a_call(ncho):
    net, _ = ncho
    with CudaSolver(net, "cuda") as solver:
        assert solver.fallback["cold"] == 0
    solver.close()  # twice does nothing
    inp = workloads.config0(net.species, 0, cells=8, seed=2)
    with pytest.raises(RuntimeError, match="closed"):
        solver(t(inp["x"]), t(inp["T"]), t(inp["P"]), full(8, 1.0), full(8, 1.0))


ALLOWANCE = 1e-6
ATOMS = {"H": {"H": 1}, "H2": {"H": 2}, "OH": {"H": 1, "O": 1}, "H2O": {"H": 2, "O": 1}, "O": {"O": 1}, "He": {"He": 1}}
DRIFTS = (0.5, 0.99, 1.01, 2.0)  # the largest relative change of an element share, in allowances


def test_guard_flags_drift_beyond_the_allowance(read_network, gpu):
    # the call kernel's guard alone (tests/recovery_probe.hip), on states whose He share moved by r allowances; the
    # oracle's check_cell agrees
    from probes import Probes

    net = read_network("synthetic_small")
    x0 = np.array([{"H": 0.1, "H2": 0.6, "OH": 0.05, "H2O": 0.1, "O": 0.05, "He": 0.1}[q] for q in net.species])
    atoms = np.array([[ATOMS[q].get(e, 0) for e in ("H", "O", "He")] for q in net.species], dtype=float)
    share = (x0 @ atoms) / (x0 @ atoms).sum()
    he, f = net.species.index("He"), share[2]
    results = []
    for r in DRIFTS:
        d = r * ALLOWANCE / (1.0 - f - f * r * ALLOWANCE)
        x = x0.copy()
        x[he] *= 1.0 + d
        moved = (x @ atoms) / (x @ atoms).sum() / share - 1.0
        assert np.abs(moved).max() == pytest.approx(r * ALLOWANCE, rel=1e-6)
        results.append(x / x.sum())
    probes = Probes(net, drift=ALLOWANCE)
    try:
        _, code = probes.run_guard(np.array([x0] * len(DRIFTS)), np.array(results), 1, 1.0)
    finally:
        probes.close()
    orc = oracle.Oracle(_tables.build(net, False), ALLOWANCE)
    want = [orc.check_cell(x0, x, True, 1.0)[0] for x in results]
    print(f"drift {DRIFTS} allowances: guard bits {code.tolist()}, oracle {want}")
    assert code.tolist() == want == [16 * int(r > 1.0) for r in DRIFTS]


GUARD_CASES = 64


def test_guard_matches_the_oracle(gpu):
    # the guard on NCHO results (finished and not, drifted, non-finite) against the oracle's check_cell: the same
    # bits, and the corrected shares within rounding (the sums run in another order); the lanes' sums of the
    # correction restore the input's element shares to 1e-14
    from probes import Probes

    net = standin.load("NCHO")
    tab = _tables.build(net, False)
    rng = np.random.default_rng(3)
    inp = workloads.config0(net.species, 0, cells=GUARD_CASES, seed=4)
    x0 = inp["x"]
    res = x0 * np.exp(rng.normal(0.0, 1e-7, x0.shape))  # small moves: corrected
    res[8:16] = x0[8:16] * np.exp(rng.normal(0.0, 1e-2, (8, tab.ns)))  # large moves: drift
    res[16, 5] = np.inf
    res[17, 0] = np.nan
    done = np.ones(GUARD_CASES, np.int32)
    done[20:24] = 0
    nok = rng.integers(1, 50, GUARD_CASES).astype(np.float64)
    probes = Probes(net)
    try:
        out, code = probes.run_guard(x0, res, done, nok)
    finally:
        probes.close()
    orc = oracle.Oracle(tab)
    atoms = tab.atoms
    for b in range(GUARD_CASES):
        want_code, want = orc.check_cell(x0[b], res[b] / res[b].sum(), bool(done[b]), nok[b])
        assert code[b] == want_code, b
        if want_code == 0:
            np.testing.assert_allclose(out[b], want, rtol=1e-13, atol=0)
            s0, s1 = x0[b] @ atoms, out[b] @ atoms
            assert np.abs(s1 / s1.sum() - s0 / s0.sum()).max() <= 1e-14 * (s0 / s0.sum()).max()
    assert sorted(set(code.tolist())) == [0, 4, 8, 16]


def kernel_sum(row, perm, k0, nt):
    """The call kernel's screen sum of x (kernels/call.hip): lane l adds its lead entry x[perm[l]] (l < K0) and its
    trailing entry x[perm[K0 + l]] (l < NT); the lanes are added in pairs, quads, half-rows and rows, then the four
    rows as (row0 + row1) + (row2 + row3)."""
    lanes = [(row[perm[l]] if l < k0 else 0.0) + (row[perm[k0 + l]] if l < nt else 0.0) for l in range(64)]
    for width in (2, 4, 8, 16):
        lanes = _fold(lanes, width)
    r = [lanes[16 * q] for q in range(4)]
    return (r[0] + r[1]) + (r[2] + r[3])


def _fold(lanes, width):
    """One DPP step of the kernel's reduction: every lane of each group of `width` gets the sum of its half and the
    other half (a + b, which rounds as b + a)."""
    out = list(lanes)
    half = width // 2
    for g in range(0, 64, width):
        for i in range(g, g + width):
            j = i ^ half if width <= 4 else (g + width - 1 - (i - g))  # pairs and quads by xor; halves and rows mirrored
            out[i] = lanes[i] + lanes[j]
    return out


def on_the_tolerance(rows, perm, k0, nt, kernel_above):
    """A row of the given states moved so that its sum is on the tolerance: above it in the kernel's order and not
    above in index order (kernel_above), or the other way round."""
    for row in rows:
        base = np.asarray(row, np.float64)
        j = int(np.argmax(base))
        target = 1.0 + X_SUM_TOL - (sum(base.tolist()) - base[j])
        for step in range(-400, 400):
            r = base.copy()
            r[j] = np.nextafter(target, np.inf) if step == 0 else target + step * np.spacing(target)
            ks = kernel_sum(r.tolist(), perm, k0, nt) - 1.0
            ns = sum(r.tolist()) - 1.0
            if kernel_above and ks > X_SUM_TOL and not ns > X_SUM_TOL:
                return r
            if not kernel_above and not ks > X_SUM_TOL and ns > X_SUM_TOL:
                return r
    return None


def test_sum_on_the_tolerance_is_judged_once(ncho):
    # only the kernel's own sum of x is judged: a sum above X_SUM_TOL in its order and not in index order is
    # refused; the opposite cell is accepted and stepped (the reference's test_guard.py:222-241)
    net, solver = ncho
    inp = workloads.config0(net.species, 0, cells=16, seed=3)
    perm = solver.tab.perm.tolist()
    above = on_the_tolerance(inp["x"], perm, solver.tab.k0, solver.tab.nt, True)
    below = on_the_tolerance(inp["x"], perm, solver.tab.k0, solver.tab.nt, False)
    assert above is not None and below is not None
    T, P = t(inp["T"][:1]), t(inp["P"][:1])
    dt, h = full(1, 50.0), full(1, 1e-10)
    with pytest.raises(ValueError, match=r"1 cells of bad input: .*\(first: cell 0: sum\(x\) - 1 = "):
        solver(t([above]), T, P, dt, h)
    assert solver.fallback["bad_input"] == 1
    r = solver(t([below]), T, P, dt, h)
    assert bool(r.done.all()) and solver.fallback["bad_input"] == 0


@pytest.mark.parametrize("variant", [(1, 16, 1, 0), (1, 16, 1, 1)], ids=["rates in LDS", "rates in global scratch"])
def test_close_releases_the_gpu_memory(read_network, gpu, variant):
    before = torch.cuda.memory_allocated()
    solver = CudaSolver(read_network("synthetic_small"), "cuda", variant=variant)
    held = torch.cuda.memory_allocated() - before
    solver.close()
    left = torch.cuda.memory_allocated() - before
    print(f"holds {held} B, {left} B after close (the counts stay readable)")
    assert left <= 4096 < held or left < held // 4


def test_an_empty_batch_is_a_call(ncho):
    net, solver = ncho
    ns = len(net.species)
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    r = solver.launch(torch.empty(0, ns, dtype=torch.float64, device="cuda"), e, e, e, e)
    assert r.x.shape == (0, ns) and r.h.shape == (0,) and r.done.shape == (0,) and r.flag.shape == (0,)
    assert int(solver.count[1:-1].abs().sum()) == 0


def test_an_empty_j_is_none_without_photolysis(read_network, gpu):
    # without photolysis, a J of no columns is the same call as no J; one with columns is refused by name
    net = read_network("synthetic_small")
    cells = 4
    x = torch.zeros(cells, len(net.species), dtype=torch.float64, device=gpu)
    for name, value in (("H", 0.1), ("H2", 0.6), ("OH", 0.05), ("H2O", 0.1), ("O", 0.05), ("He", 0.1)):
        x[:, net.species.index(name)] = value
    T, P, dt, h = (torch.full((cells,), q, dtype=torch.float64, device=gpu) for q in (1500.0, 1e6, 1e-3, 1e-3))
    empty = torch.zeros(cells, 0, dtype=torch.float64, device=gpu)
    with CudaSolver(net, gpu) as solver:
        assert all(map(torch.equal, solver(x, T, P, dt, h), solver(x, T, P, dt, h, J=empty)))
        with pytest.raises(ValueError, match=r"^J: none or an empty \[4, 0\] one expected"):
            solver(x, T, P, dt, h, J=torch.zeros(cells, 1, dtype=torch.float64, device=gpu))
