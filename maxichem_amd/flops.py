"""Algorithmic floating-point work of the chemistry step, counted from a network's tables.

The unit is one Rodas3 try of one cell; a cell-call adds its rate coefficients once. The counts are of the work the
method needs, not of what a kernel executes (a dense block does more): the sparse LU of M0 on its filled pattern in
the elimination order, its triangular solves, three right-hand sides and one Jacobian per try. Transcendentals count
1 each. The split is the reference's (rates, rhs, Jacobian, u, LU, solves, Sherman-Morrison, step control; the call
kernel maxichem/kernels/call.cuh:51-167 and common.cuh:222-599).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._tables import Tables


@dataclass
class Work:
    rhs: int
    jacobian: int
    jx: int
    lu: int
    solve: int
    sherman_morrison: int
    control: int
    rates: int

    @property
    def per_try(self) -> int:
        return 3 * self.rhs + self.jacobian + self.jx + self.lu + 5 * self.solve + self.sherman_morrison + self.control


def symbolic_lu(pattern: set[tuple[int, int]], n: int) -> tuple[int, int, int, int]:
    """(scalings, updates, nnz(L strictly lower), nnz(U upper incl. diagonal)) of the unpivoted LU on pattern."""
    rows_of = [set() for _ in range(n)]  # column k: rows i > k with an entry
    cols_of = [set() for _ in range(n)]  # row k: columns j > k with an entry
    for i, j in pattern:
        if i > j:
            rows_of[j].add(i)
        elif j > i:
            cols_of[i].add(j)
    scal = upd = 0
    for k in range(n):
        rk, ck = rows_of[k], cols_of[k]
        scal += len(rk)
        upd += len(rk) * len(ck)
        for i in rk:  # fill: (i, j) for j in ck
            for j in ck:
                if i > j:
                    rows_of[j].add(i)
                elif j > i:
                    cols_of[i].add(j)
    nnz_l = sum(len(r) for r in rows_of)
    nnz_u = sum(len(c) for c in cols_of) + n
    return scal, upd, nnz_l, nnz_u


def work(tab: Tables) -> Work:
    ns, nr = tab.ns, tab.nr
    pos = np.empty(ns, np.int64)
    pos[tab.perm] = np.arange(ns)
    # -Jn's pattern: a reaction's species p against the species q its rate depends on (a photolysis branch: its
    # reactant only)
    entries = set()
    for r in range(nr):
        x = tab.rx[r]
        species = [int(s) for s in x if s >= 0]
        sides = (x[:3], x[3:]) if r < tab.nrt else (x[:3],)
        qs = {int(s) for side in sides for s in side if s >= 0}
        for p in species:
            for q in qs:
                entries.add((int(pos[p]), int(pos[q])))
    pattern = entries | {(i, i) for i in range(ns)}
    scal, upd, nnz_l, nnz_u = symbolic_lu(pattern, ns)
    mols = sum(int((tab.rx[r, :3] >= 0).sum() + (r < tab.nrt) * (tab.rx[r, 3:] >= 0).sum()) + 1 for r in range(nr))
    thermal = tab.nrt
    return Work(
        rhs=4 * ns + 2 * tab.sv_terms + mols,
        jacobian=4 * tab.jac_terms,
        jx=2 * len(entries),
        lu=ns + scal + 2 * upd,
        solve=2 * (nnz_l + nnz_u) + ns,
        sherman_morrison=4 * 3 * ns + ns,
        control=15 * ns,
        rates=30 * ns + thermal * 18,
    )


def call_flops(tab: Tables, tries: np.ndarray) -> float:
    """Algorithmic flops of one call over its cells, from each cell's tries."""
    w = work(tab)
    return float(len(tries) * w.rates + np.sum(tries) * w.per_try)
