"""The recovery kernel's dense LU with threshold partial pivoting and its solve, run alone on the GPU
(tests/recovery_probe.hip) on matrices chosen to exercise the rule, against the rule written out in numpy and the
oracle's LU (the reference checks its recovery LU the same way: maxichem tests/test_dense_lu.py:36-115).

Rows are exchanged only where |W_kk| is below PIVTAU times the column's largest |entry| at or below k, a tie going
to the lower row. The networks fix the size: 42 species (one row per lane), the NCHO stand-in's 69 (a second row on
five lanes, factor rows held in registers by the solve) and the SNCHO stand-in's 91 (factor rows read from LDS)."""

import helpers
import numpy as np
import oracle
import pytest

from maxichem_amd import standin

pytestmark = pytest.mark.gpu

PIVTAU = 0.1
FACTOR_TOLERANCE = 1e-13  # |P A - L U| / (|L| |U|) per entry: rounding of the updates
SOLVE_TOLERANCE = 1e-10  # relative to numpy's solve and factors: rounding times these matrices' condition


@pytest.mark.parametrize("network", ["synthetic", "NCHO", "SNCHO"])
def test_dense_lu_pivots_by_the_threshold_rule(read_network, gpu, network):
    from probes import Probes

    net = read_network(network) if network == "synthetic" else standin.load(network)
    probes = Probes(net)
    try:
        ns = probes.tab.ns
        rng = np.random.default_rng(11)
        mats = helpers.adversarial_matrices(ns, rng)
        y = rng.standard_normal((len(mats), ns))
        factors, rows, dinv, solved = probes.run_dense(mats, y)
    finally:
        probes.close()
    exchanged, kept_rows = [], []
    for m, a in enumerate(mats):
        order, want, exchanges, kept = helpers.threshold_lu_numpy(a, PIVTAU)
        exchanged.append(exchanges)
        kept_rows.append(kept)
        got = factors[m]
        lower, upper = np.tril(got, -1) + np.eye(ns), np.triu(got)
        residual = np.abs(lower @ upper - a[order]) / (np.abs(lower) @ np.abs(upper))
        solution = np.linalg.solve(a, y[m])
        err = np.abs(solved[m] - solution).max() / np.abs(solution).max()
        o_rows, o_factors, o_dinv, o_sol = oracle.threshold_lu(a, y[m])
        print(f"{network} matrix {m}: {exchanges} exchanges, {kept} kept, residual {residual.max():.1e}, "
              f"solve {err:.1e}")  # fmt: skip
        assert rows[m].tolist() == order == o_rows.tolist()
        assert residual.max() <= FACTOR_TOLERANCE
        np.testing.assert_allclose(got, want, rtol=SOLVE_TOLERANCE, atol=SOLVE_TOLERANCE * np.abs(want).max())
        np.testing.assert_allclose(got, o_factors, rtol=SOLVE_TOLERANCE, atol=SOLVE_TOLERANCE * np.abs(want).max())
        assert np.array_equal(dinv[m], 1.0 / np.diag(upper))
        assert err <= SOLVE_TOLERANCE
        assert np.abs(solved[m] - o_sol[0]).max() / np.abs(solution).max() <= SOLVE_TOLERANCE
    # the cases do what they are for: exchanges at most pivots, rows kept by the threshold, the tie to the lower row
    assert min(exchanged[:8]) >= ns // 2
    assert min(kept_rows[8:16]) >= 1
    assert rows[-1][0] == 3
