"""The oracle's pieces against what the mathematics fixes: its constants against the package's, the threshold LU
against the rule written out in numpy, the Jacobian against finite differences, one Rodas3 step against a dense
numpy solve of W, and the guard's drift test and share correction. CPU only."""

import helpers
import numpy as np
import oracle
import pytest

from maxichem_amd import _constants, _tables

PIVTAU = 0.1


def test_oracle_constants_are_the_packages():
    for name in ("T_MIN", "T_MAX", "X_SUM_TOL", "REC_STEPS", "H_MIN", "CALL_END", "STEP_SAFETY", "ERR_FLOOR",
                 "FAC_NONFINITE", "FACACC", "FACREJ", "FACMAX", "FAC_AFTER_REJECT", "SECMIN", "SECANT_MIN_ORDER",
                 "SECANT_MAX_ORDER", "SECANT_TARGET", "PIVTAU", "ELEMENT_PIVOT_MIN", "REC_RETRIES",
                 "REC_RETRY_SPLIT", "GMAX", "KB", "P0", "NASA_SPLIT_T", "K_COLL", "FL_FALLOFF", "FL_SECOND_K0",
                 "FL_THREE_BODY", "FL_TROE", "FL_EFF", "FL_VM11", "FL_PHOTO", "FL_NF_SHIFT", "FL_NR_SHIFT"):  # fmt: skip
        assert oracle.constant(name) == float(getattr(_constants, name)), name


def test_threshold_lu_follows_the_rule():
    rng = np.random.default_rng(11)
    n = 42
    mats = helpers.adversarial_matrices(n, rng)
    exchanged, kept_rows = [], []
    for m, a in enumerate(mats):
        y = rng.standard_normal(n)
        pr, got, dinv, sol = oracle.threshold_lu(a, y)
        order, want, exchanges, kept = helpers.threshold_lu_numpy(a, PIVTAU)
        exchanged.append(exchanges)
        kept_rows.append(kept)
        lower, upper = np.tril(got, -1) + np.eye(n), np.triu(got)
        residual = np.abs(lower @ upper - a[order]) / (np.abs(lower) @ np.abs(upper))
        solution = np.linalg.solve(a, y)
        assert pr.tolist() == order
        assert residual.max() <= 1e-13
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
        assert np.array_equal(dinv, 1.0 / np.diag(upper))
        assert np.abs(sol[0] - solution).max() / np.abs(solution).max() <= 1e-10
    assert min(exchanged[:8]) >= n // 2  # shuffled rows: exchanges at most pivots
    assert min(kept_rows[8:16]) >= 1  # some rows kept by the threshold
    assert oracle.threshold_lu(mats[-1])[0][0] == 3  # the tie went to the lower row


def _random_state(net, rng):
    x = 10.0 ** rng.uniform(-8, 0, len(net.species))
    return x / x.sum()


@pytest.mark.parametrize("stem", ["synthetic", "synthetic_small"])
def test_jacobian_matches_central_differences(read_network, stem):
    net = read_network(stem)
    tab = _tables.build(net, photo=True)
    orc = oracle.Oracle(tab)
    rng = np.random.default_rng(5)
    N = _random_state(net, rng) * 1.3  # N need not sum to 1
    J = np.full(len(net.photo), 1e-3)
    T, P = 1700.0, 3e5
    F, JN = orc.rhs_jacobian(T, P, N, J)
    for j in range(len(N)):
        e = abs(N[j]) * 1e-4 + 1e-9  # central differences: truncation ~e^2, rounding ~eps |F| / e
        up, down = N.copy(), N.copy()
        up[j] += e
        down[j] -= e
        fd = (orc.rhs_jacobian(T, P, up, J)[0] - orc.rhs_jacobian(T, P, down, J)[0]) / (2 * e)
        scale = np.abs(fd).max() + np.abs(JN[:, j]).max() + 1e-300
        assert np.abs(fd - JN[:, j]).max() <= 1e-5 * scale, j


def test_one_step_matches_a_dense_solve(read_network):
    # a single accepted step (rtol large) from a random state: the oracle's sparse-order LU and Sherman-Morrison
    # against numpy.linalg.solve of the dense W = I / (gamma h) - dF/dN, with the stages of Rodas3
    net = read_network("synthetic")
    tab = _tables.build(net, photo=True)
    orc = oracle.Oracle(tab)
    rng = np.random.default_rng(9)
    x = _random_state(net, rng)
    J = np.full(len(net.photo), 1e-4)
    T, P, h = 1800.0, 1e6, 1e-7
    out = orc.call(x[None], T, P, h, h, rtol=1e6, max_steps=1, J=J[None])
    assert out["done"][0] and out["steps"][0] == 1
    F, JN = orc.rhs_jacobian(T, P, x, J)
    W = np.eye(len(x)) / (0.5 * h) - JN
    k1 = np.linalg.solve(W, F)
    k2 = np.linalg.solve(W, F + 4.0 * k1 / h)
    f3 = orc.rhs_jacobian(T, P, x + 2 * k1, J)[0]
    k3 = np.linalg.solve(W, f3 + (k1 - k2) / h)
    f4 = orc.rhs_jacobian(T, P, x + 2 * k1 + k3, J)[0]
    k4 = np.linalg.solve(W, f4 + (k1 - k2 - 8.0 / 3.0 * k3) / h)
    y = np.maximum(x + 2 * k1 + k3 + k4, 0.0)
    want = y / y.sum()
    big = want > 1e-12
    np.testing.assert_allclose(out["x"][0][big], want[big], rtol=1e-7)


ATOMS = {"H": {"H": 1}, "H2": {"H": 2}, "OH": {"H": 1, "O": 1}, "H2O": {"H": 2, "O": 1}, "O": {"O": 1}, "He": {"He": 1}}
DRIFTS = (0.5, 0.99, 1.01, 2.0)


def test_drift_beyond_the_allowance_is_flagged(read_network):
    # hand-made results whose He share moved by r times the allowance: flagged only above 1
    net = read_network("synthetic_small")
    allowance = 1e-6
    orc = oracle.Oracle(_tables.build(net), allowance)
    x0 = np.array([{"H": 0.1, "H2": 0.6, "OH": 0.05, "H2O": 0.1, "O": 0.05, "He": 0.1}[q] for q in net.species])
    atoms = np.array([[ATOMS[q].get(e, 0) for e in net.elements] for q in net.species], dtype=float)
    share = (x0 @ atoms) / (x0 @ atoms).sum()
    f = share[net.elements.index("He")]
    codes = []
    for r in DRIFTS:
        d = r * allowance / (1.0 - f - f * r * allowance)
        x = x0.copy()
        x[net.species.index("He")] *= 1.0 + d
        moved = (x @ atoms) / (x @ atoms).sum() / share - 1.0
        assert np.abs(moved).max() == pytest.approx(r * allowance, rel=1e-6)
        codes.append(orc.check_cell(x0, x, True, 1.0)[0])
    assert codes == [_constants.GUARD_DRIFT * (r > 1.0) for r in DRIFTS]


def test_share_correction_restores_the_shares(read_network):
    net = read_network("synthetic_small")
    orc = oracle.Oracle(_tables.build(net))
    rng = np.random.default_rng(2)
    x0 = _random_state(net, rng)
    x = x0 * (1.0 + 1e-9 * rng.standard_normal(len(x0)))  # a result whose shares moved by rounding-sized amounts
    x /= x.sum()
    code, fixed = orc.check_cell(x0, x, True, 1.0)
    assert code == 0
    s0, s1 = (q @ net.atoms for q in (x0, fixed))
    s0, s1 = s0 / s0.sum(), s1 / s1.sum()
    assert np.abs(s1 / s0 - 1.0).max() < 1e-14
    assert abs(fixed.sum() - 1.0) < 1e-15
