"""The step-size law, pinned by known answers, on the oracle and on the GPU (the reference's step_factor and
recovery_start, maxichem/kernels/common.cuh:87-124, and docs/method.md's step control).

- The factor alone on inputs whose answer is exact by hand (err a perfect cube, the clamps, the floor, a non-finite
  err, the secant after two rejections with a fitted order of 1, 3 and beyond, and its floor), on the oracle's
  step_factor and on the kernels' own (tests/device_probes.hip).
- Through a call: A -> 2 B by photolysis from N = (1, 0), whose first Rodas3 try is known in exact rational
  arithmetic, so its error is too: an accepted try over the whole dt returns h = dt 0.9 err^(-1/3) (and 20 dt when
  the error is tiny); two rejected tries return the secant step from the two exact errors.
- The recovery's passes for a result that only drifted start at dt 0.5^k: with a drift allowance of 0, every pass
  drifts, and the last one is the recovery started at min(h, dt / 8).
"""

import math
from fractions import Fraction

import helpers
import numpy as np
import oracle
import pytest
from helpers import BACKENDS

from maxichem_amd._constants import REC_RETRIES, REC_RETRY_SPLIT, REC_STEPS

T_TEST, P_TEST, J_TEST = 1500.0, 1e6, 2.0
ATOL = 1e-25

# (err, accepted, previous rejected, previous err, previous step, step) -> factor
KNOWN = [
    ((0.125, 1, 0, 0.0, 0.0, 1.0), 1.8),  # 0.9 / cbrt(1/8)
    ((1.0, 1, 0, 0.0, 0.0, 1.0), 0.9),
    ((0.001, 1, 0, 0.0, 0.0, 1.0), 9.0),  # 0.9 / 0.1
    ((1e-6, 1, 0, 0.0, 0.0, 1.0), 20.0),  # 90, held at FACMAX
    ((0.0, 1, 0, 0.0, 0.0, 1.0), 20.0),  # err floored at 1e-10: 1939, held at FACMAX
    ((729.0, 1, 0, 0.0, 0.0, 1.0), 0.2),  # 0.1, held at FACACC
    ((8.0, 0, 0, 0.0, 0.0, 1.0), 0.45),  # rejected: 0.9 / 2
    ((1.728, 0, 0, 0.0, 0.0, 1.0), 0.5),  # 0.75, held at FAC_AFTER_REJECT
    ((1000.0, 0, 0, 0.0, 0.0, 1.0), 0.2),  # 0.09, held at FACREJ
    ((math.nan, 0, 0, 0.0, 0.0, 1.0), 0.2),  # non-finite: 0.1, then FACREJ
    ((math.inf, 0, 1, 8.0, 2.0, 1.0), 0.2),  # non-finite: no secant
    ((2.0, 0, 1, 8.0, 2.0, 1.0), 0.5),  # secant, order log(8/2)/log(2) = 2: (0.5/2)^(1/2) = 0.5
    ((8.0, 0, 1, 64.0, 2.0, 1.0), (1.0 / 16.0) ** (1.0 / 3.0)),  # order 3: (0.5/8)^(1/3)
    ((4.0, 0, 1, 16.0, 4.0, 1.0), 0.125),  # order 1: 0.5/4
    ((16.0, 0, 1, 16.0 * 2**10, 2.0, 1.0), 2.0 ** (-5.0 / 3.0)),  # order 10, held at 3: (0.5/16)^(1/3)
    ((16.0, 0, 1, 16.0 * 1.01, 2.0, 1.0), 1e-4),  # order 0.014 below 0.05: SECMIN
    ((1e12, 0, 1, 1e13, 10.0, 1.0), 1e-4),  # order 1: 5e-13, held at SECMIN
    ((8.0, 0, 0, 64.0, 2.0, 1.0), 0.45),  # the previous try was accepted: no secant
    ((8.0, 0, 1, 0.5, 2.0, 1.0), 0.45),  # the previous err was not above 1: no secant
    ((8.0, 0, 1, 64.0, 1.0, 1.0), 0.45),  # the previous step was not longer: no secant
    ((0.5, 1, 1, 64.0, 2.0, 1.0), 0.9 / 0.5 ** (1.0 / 3.0)),  # accepted: no secant
]


@pytest.mark.parametrize("kind", BACKENDS)
def test_step_factor_known_answers(read_network, kind):
    rows = [q for q, _ in KNOWN]
    if kind == "oracle":
        got = np.array([oracle.step_factor(*q) for q in rows])
    else:
        from probes import Probes

        probes = Probes(read_network("synthetic_decay"), photo=True)
        try:
            got = probes.run_step_factor(rows)
        finally:
            probes.close()
    want = np.array([f for _, f in KNOWN])
    print(np.c_[got, want])
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)


def first_try(z: Fraction):
    """N' and K4 of one Rodas3 try of A -> 2 B from N = (1, 0) at J h = z, exactly (K in units of the state).

    h W = I / gamma - h dF/dN with dF/dN = [[-J, 0], [2 J, 0]]; the stages as in docs/method.md."""
    lower = (2 + z, -2 * z, Fraction(2))  # h W, lower triangular

    def solve(b):
        a = b[0] / lower[0]
        return [a, (b[1] - lower[1] * a) / lower[2]]

    def hf(y):  # h F(y)
        return [-z * y[0], 2 * z * y[0]]

    n = [Fraction(1), Fraction(0)]
    k1 = solve(hf(n))
    k2 = solve([hf(n)[i] + 4 * k1[i] for i in range(2)])
    y3 = [n[i] + 2 * k1[i] for i in range(2)]
    k3 = solve([hf(y3)[i] + k1[i] - k2[i] for i in range(2)])
    y4 = [y3[i] + k3[i] for i in range(2)]
    k4 = solve([hf(y4)[i] + k1[i] - k2[i] - Fraction(8, 3) * k3[i] for i in range(2)])
    return [y4[i] + k4[i] for i in range(2)], k4


def first_error(h: float, rtol: float, ns: int) -> float:
    """The error of the first try of size h from N = (1, 0), exactly, rounded once: before any accepted step the
    larger of the RMS over all ns species and the largest scaled K4."""
    new, k4 = first_try(Fraction(J_TEST) * Fraction(h))
    start = [Fraction(1), Fraction(0)]
    q = [k4[i] / (Fraction(ATOL) + Fraction(rtol) * max(abs(start[i]), abs(new[i]))) for i in range(2)]
    largest = max(abs(v) for v in q)
    rms_squared = sum(v * v for v in q) / ns
    assert rms_squared < largest * largest  # so the error is the largest, a rational
    return float(largest)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("target", [0.125, 0.6, 1e-9])
def test_an_accepted_try_sets_the_next_step(read_network, kind, target):
    net = read_network("synthetic_decay")
    ns = len(net.species)
    dt = 0.05  # J dt = 1/10: the first try covers the call
    rtol = first_error(dt, 1.0, ns) / target
    en = first_error(dt, rtol, ns)
    want = dt * min(max(0.9 / np.cbrt(max(en, 1e-10)), 0.2), 20.0)
    run = helpers.backend(kind, net, photo=True)
    out = run(helpers.state(net, 1, N2_1=1.0), T_TEST, P_TEST, dt, dt, rtol=rtol, atol=ATOL,
              J=np.full((1, len(net.photo)), J_TEST))  # fmt: skip
    run.close()
    print(f"err {en:.6g}: h {out['h'][0]!r}, closed form {want!r}")
    assert out["done"].all() and out["steps"][0] == 1.0
    assert out["h"][0] == pytest.approx(want, rel=1e-14)
    if target < 1e-6:
        assert out["h"][0] == 20.0 * dt  # held at FACMAX, exactly


@pytest.mark.parametrize("kind", BACKENDS)
def test_two_rejections_take_the_secant_step(read_network, kind):
    net = read_network("synthetic_decay")
    ns = len(net.species)
    h1 = 0.5  # J h = 1
    h2 = h1 * 0.2  # the first error is so large that its factor is held at FACREJ
    rtol = first_error(h2, 1.0, ns) / 10.0  # the second try's error is 10
    e1, e2 = first_error(h1, rtol, ns), first_error(h2, rtol, ns)
    assert 0.9 / np.cbrt(e1) < 0.2 and e2 > 1.0
    order = math.log(e1 / e2) / math.log(h1 / h2)
    want = h2 * min(max((0.5 / e2) ** (1.0 / min(order, 3.0)), 1e-4), 0.5)
    x0 = helpers.state(net, 1, N2_1=1.0)[0]
    J = np.full(len(net.photo), J_TEST)
    if kind == "oracle":
        from maxichem_amd import _tables

        got = oracle.Oracle(_tables.build(net, True)).try_loop(x0, T_TEST, P_TEST, h1, h1, rtol=rtol, atol=ATOL,
                                                               max_steps=2, J=J)  # fmt: skip
        h, steps, tries = got["h"], got["steps"], got["tries"]
    else:
        from probes import call_kernel_only

        from maxichem_amd.kincuda import CudaSolver

        with CudaSolver(net, "cuda", photo=True) as solver:
            got = call_kernel_only(solver, x0, T_TEST, P_TEST, h1, h1, rtol=rtol, atol=ATOL, max_steps=2, J=J[None])
        h, steps, tries = got["h"][0], got["steps"][0], got["tries"][0]
    print(f"errors {e1:.6g}, {e2:.6g}, fitted order {order:.4f}: h {h!r}, closed form {want!r}")
    assert steps == 0.0 and tries == 2.0
    assert h == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("kind", BACKENDS)
def test_recovery_passes_start_at_halved_steps(read_network, kind):
    # with no drift allowed every finished result drifts (rounding moves the shares), so the recovery runs all its
    # passes; the last one starts at min(h, dt / 2^REC_RETRIES), and its tries add to the call's and the others'
    net = read_network("synthetic_small")
    from maxichem_amd import _tables

    orc = oracle.Oracle(_tables.build(net, False), 0.0)
    x0 = helpers.state(net, 3, H=0.1, H2=0.6, OH=0.05, H2O=0.1, O=0.05, He=0.1)
    T, P, dt, h = 1500.0, 1e6, np.array([1.0, 10.0, 100.0]), np.array([1.0, 1e-3, 1e-9])
    run = helpers.backend(kind, net, drift=0.0)
    out = run(x0, T, P, dt, h)
    run.close()
    for b in range(3):
        passes = [orc.try_loop(x0[b], T, P, dt[b], max(min(h[b], dt[b]) if k == 0 else
                                                         min(h[b], dt[b], dt[b] * REC_RETRY_SPLIT**k), 1e-12),
                               max_steps=REC_STEPS * 4000, dense=True)
                  for k in range(REC_RETRIES + 1)]  # fmt: skip
        call = orc.try_loop(x0[b], T, P, dt[b], h[b])
        print(f"cell {b}: flag {out['flag'][b]}, h {out['h'][b]!r}, last pass {passes[-1]['h']!r}, tries "
              f"{out['tries'][b]} = {call['tries']} + {[q['tries'] for q in passes]}")  # fmt: skip
        assert out["flag"][b] == 2 and not out["done"][b]
        assert out["steps"][b] == passes[-1]["steps"]
        assert out["h"][b] == pytest.approx(passes[-1]["h"], rel=1e-6 if kind == "gpu" else 0.0)
        assert out["tries"][b] == call["tries"] + sum(q["tries"] for q in passes)
