"""The oracle of the GCM-side conversions: TEST INFRASTRUCTURE ONLY (tests/ may use it as the checker; the product
package never imports it).

A float64 numpy restatement of the reference's maxichem/coupling/composition.py per-cell arithmetic, written from the
formulas (each function cites the reference lines it follows); tests/test_coupling.py pins it against the reference's
hand-worked values (tests/test_coupling_composition.py:36-99 of the reference) and against exact arithmetic in
fractions on a small case.
"""

from __future__ import annotations

import numpy as np


def mass_to_mole(q: np.ndarray, mw: np.ndarray) -> np.ndarray:
    """x = n / sum(n), n = q / mw, per row (composition.py:126-130)."""
    n = q / mw[None, :]
    return n / n.sum(1, keepdims=True)


def mole_to_mass(x: np.ndarray, mw: np.ndarray, mass: np.ndarray) -> np.ndarray:
    """q = x mw (mass / sum(x mw)), per row (composition.py:186-192)."""
    xm = x * mw[None, :]
    return xm * (mass / xm.sum(1))[:, None]


def element_moles(q: np.ndarray, mw: np.ndarray, atoms: np.ndarray) -> np.ndarray:
    """(q / mw) atoms: [cells, E] mol per kg of gas (composition.py:217)."""
    return (q / mw[None, :]) @ atoms


def tracer_totals(q: np.ndarray, cell_mass: np.ndarray, mw: np.ndarray, atoms: np.ndarray) -> np.ndarray:
    """[1 + E]: sum of q cell_mass, then each element's moles times cell_mass summed over the cells
    (composition.py:231-233)."""
    moles = element_moles(q, mw, atoms)
    return np.concatenate([[np.sum(q * cell_mass[:, None])], (moles * cell_mass[:, None]).sum(0)])
