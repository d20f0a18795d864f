"""Seeded synthetic inputs of the benchmark configurations (all float64).

Per cell, from numpy.random.default_rng(seed): T log-uniform 300-3000 K, P log-uniform 1e2-1e9 dyn cm^-2, dt and J as
each configuration says, and a solar gas: the elemental ratios per H atom of the reference's START_PER_H
(maxichem/coupling/composition.py:33-39) placed in H2, He, H2O, CH4, CO, CO2, NH3, N2 (and H2S on a sulfur network).
Carbon is split CH4 : CO : CO2 = u : 0.9 (1 - u) : 0.1 (1 - u) and nitrogen NH3 : N2 atoms = w : 1 - w with u, w
uniform per cell from the same generator; every other species starts at 1e-30 and the row is renormalised in float64.
"""

from __future__ import annotations

import numpy as np

# atoms per H atom (maxichem/coupling/composition.py:33-39)
START_PER_H = {"O": 6.0618e-4, "C": 2.7759e-4, "N": 8.1846e-5, "He": 0.09692, "S": 1.4099e-5}
TRACE = 1e-30


def log_uniform(rng: np.random.Generator, lo: float, hi: float, n: int) -> np.ndarray:
    return 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), n)


def solar_gas(species: list[str], rng: np.random.Generator, cells: int) -> np.ndarray:
    """Mole fractions [cells, species] of the solar start gas with seeded C and N splits."""
    idx = {q: i for i, q in enumerate(species)}
    u, w = rng.uniform(0.0, 1.0, cells), rng.uniform(0.0, 1.0, cells)
    c, o, n = START_PER_H["C"], START_PER_H["O"], START_PER_H["N"]
    s = START_PER_H["S"] if "H2S" in idx else 0.0
    ch4, co, co2 = c * u, c * 0.9 * (1.0 - u), c * 0.1 * (1.0 - u)
    h2o = o - co - 2.0 * co2
    nh3, n2 = n * w, n * (1.0 - w) / 2.0
    h2 = (1.0 - 2.0 * h2o - 4.0 * ch4 - 3.0 * nh3 - 2.0 * s) / 2.0
    moles = {"H2": h2, "He": np.full(cells, START_PER_H["He"]), "H2O": h2o, "CH4": ch4, "CO": co, "CO2": co2,
             "NH3": nh3, "N2": n2}  # fmt: skip
    if s:
        moles["H2S"] = np.full(cells, s)
    x = np.full((cells, len(species)), TRACE)
    for q, m in moles.items():
        x[:, idx[q]] = m
    return x / x.sum(1, keepdims=True)


def forcing(rng: np.random.Generator, cells: int) -> tuple[np.ndarray, np.ndarray]:
    """T (K) log-uniform 300-3000 and P (dyn cm^-2) log-uniform 1e2-1e9."""
    return log_uniform(rng, 300.0, 3000.0, cells), log_uniform(rng, 1e2, 1e9, cells)


def lit_quarter(P: np.ndarray, nj: int, rng: np.random.Generator) -> np.ndarray:
    """J [cells, nj]: per-branch log-uniform 1e-9 to 1e-3 s^-1 on the lowest-P quarter (by P rank), 0 elsewhere."""
    cells = len(P)
    J = np.zeros((cells, nj))
    if nj:
        lit = np.argsort(P, kind="stable")[: cells // 4]
        J[lit] = log_uniform(rng, 1e-9, 1e-3, len(lit) * nj).reshape(len(lit), nj)
    return J


def config0(species: list[str], nj: int, cells: int = 256, seed: int = 0) -> dict[str, np.ndarray]:
    """Plain FP64 CPU case: dt log-uniform 1e2-1e5 s, h = 1e-8 s, J = 0 (dark relaxation)."""
    rng = np.random.default_rng(seed)
    T, P = forcing(rng, cells)
    x = solar_gas(species, rng, cells)
    dt = log_uniform(rng, 1e2, 1e5, cells)
    return {"x": x, "T": T, "P": P, "dt": dt, "h": np.full(cells, 1e-8), "J": np.zeros((cells, nj))}


def config1(species: list[str], nj: int, cells: int = 65536, seed: int = 1) -> dict[str, np.ndarray]:
    """GPU FP64 case: one dark call of dt = 1e8 s, then calls of dt = 300 s with J on the lowest-P quarter."""
    rng = np.random.default_rng(seed)
    T, P = forcing(rng, cells)
    x = solar_gas(species, rng, cells)
    J = lit_quarter(P, nj, rng)
    return {"x": x, "T": T, "P": P, "dt_relax": np.full(cells, 1e8), "dt": np.full(cells, 300.0),
            "h": np.full(cells, 1e-8), "J": J}  # fmt: skip


def config3(species: list[str], nj: int, cells: int = 262144, seed: int = 3) -> dict[str, np.ndarray]:
    """Uneven batching: half day (J as config1), half night; dt log-uniform 1e1-1e6 s."""
    rng = np.random.default_rng(seed)
    T, P = forcing(rng, cells)
    x = solar_gas(species, rng, cells)
    J = np.zeros((cells, nj))
    day = rng.permutation(cells)[: cells // 2]
    if nj:
        J[day] = log_uniform(rng, 1e-9, 1e-3, len(day) * nj).reshape(len(day), nj)
    dt = log_uniform(rng, 1e1, 1e6, cells)
    return {"x": x, "T": T, "P": P, "dt": dt, "h": np.full(cells, 1e-8), "J": J}
