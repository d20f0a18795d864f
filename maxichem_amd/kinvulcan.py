"""Reader of VULCAN reaction lists and NASA-9 thermochemistry files.

Reads the same files, with the same section rules and refusals, as the reference reader
(maxichem/kinvulcan.py:258-357): thermal sections "# two-body", "# 3-body reactions without high-pressure",
"# 3-body", "# special", then "# reverse stops", photolysis branches under "# photo..." and condensation lines that are
skipped. Each thermal reaction becomes a dict holding its law: ``kind`` (2 two-body, 3 falloff, 5 low-pressure only,
6 special), the Arrhenius triples and optional ``troe``, ``width`` and ``eff``.
"""

from __future__ import annotations

import ast
import re
from pathlib import Path
from typing import Any

import numpy as np

from ._constants import (
    KIND_FALLOFF,
    KIND_LOW_PRESSURE,
    KIND_SPECIAL,
    KIND_TWO_BODY,
    MAX_GIBBS_JUMP,
    NASA_SPLIT_T,
    THREE_BODY_KINDS,
)

__all__ = ["SPECIAL", "VulcanNetwork", "formula", "gibbs_rt"]

# Laws VULCAN keeps in its code rather than its lists (maxichem/kinvulcan.py:41-55): OH + CH3 + M -> CH3OH + M, the
# fits of Jasper et al. (2007) in the form of Visscher & Moses (2011): two k0 terms, k_inf and a Troe factor without
# Troe's shift terms ("vm11").
SPECIAL: dict[str, dict[str, Any]] = {
    "OH+CH3+M->CH3OH+M": {
        "A0": 1.932e3, "B0": -9.88, "C0": 7544.0,
        "A1": 5.109e-11, "B1": -6.25, "C1": 1433.0,
        "Ai": 1.031e-10, "Bi": -0.018, "Ci": -16.74,
        "troe": (0.8145, 155.8, 1675.0, 4531.0),
        "width": "vm11",
    }
}  # fmt: skip

_ENDINGS = ("# reverse stops", "# condensation", "# photo", "# re_end")
# Tested in this order, since "# 3-body" begins the heading of the low-pressure section.
_THERMAL = (
    ("# two-body", "two"),
    ("# 3-body reactions without high-pressure", "low"),
    ("# 3-body", "three"),
    ("# special", "special"),
    ("# irreversible", "irrev"),
)
KNOWN_HEADINGS = (*_ENDINGS, *(h for h, _ in _THERMAL))
_TITLES = "# id"
_NASA_SIZES = (20, 30, 40)


def formula(name: str) -> dict[str, int]:
    """Element counts of a species name (CH3OH -> C 1, H 4, O 1); a suffix after "_" is a label, not atoms."""
    counts: dict[str, int] = {}
    for element, n in re.findall(r"([A-Z][a-z]?)(\d*)", name.split("_")[0]):
        counts[element] = counts.get(element, 0) + int(n or 1)
    return counts


def gibbs_rt(a: np.ndarray, T: float) -> float:
    """G/RT at T from one NASA-9 range a1..a7, b1, b2 (McBride et al. 2002): H/RT - S/R."""
    lt = np.log(T)
    h = -a[0] / T**2 + a[1] * lt / T + a[2] + a[3] * T / 2 + a[4] * T**2 / 3 + a[5] * T**3 / 4 + a[6] * T**4 / 5
    s = -a[0] / T**2 / 2 - a[1] / T + a[2] * lt + a[3] * T + a[4] * T**2 / 2 + a[5] * T**3 / 3 + a[6] * T**4 / 4
    return float(h + a[7] / T - s - a[8])


def _split(line: str, where: str) -> tuple[str, list[str]]:
    """``id [ A + B -> C ] fields`` -> (equation, fields)."""
    if "[" not in line or "]" not in line.split("[", 1)[1]:
        raise ValueError(f"{where}: a reaction line needs [ reaction ]: {line}")
    equation, rest = line.split("[", 1)[1].split("]", 1)
    if "->" not in equation:
        raise ValueError(f"{where}: a reaction needs ->: {line}")
    return equation.strip(), rest.split()


def _sides(equation: str) -> tuple[list[str], list[str]]:
    left, right = equation.split("->")
    return [q.strip() for q in left.split("+")], [q.strip() for q in right.split("+")]


def _coefficients(fields: list[str], n: int, where: str, line: str) -> list[float]:
    try:
        values = [float(v) for v in fields[:n]]
    except ValueError:
        raise ValueError(f"{where}: {n} rate coefficients expected: {line}") from None
    if len(values) < n:
        raise ValueError(f"{where}: {n} rate coefficients expected, found {len(values)}: {line}")
    if not np.all(np.isfinite(values)):
        raise ValueError(f"{where}: finite rate coefficients expected, got {values}: {line}")
    return values


def _check_m(equation: str, kind: int, where: str) -> None:
    """M must be on both sides of a three-body law and on neither side of a two-body one."""
    left, right = _sides(equation)
    has = ["M" in left, "M" in right]
    want = kind in THREE_BODY_KINDS
    if has != [want, want]:
        sides = " and ".join(s for s, m in zip(("reactant", "product"), has) if m) or "no"
        need = "on both sides (a three-body law)" if want else "on neither side (a two-body law)"
        raise ValueError(f"{where}: M on {sides} side, but this reaction's law needs M {need}: {equation}")


def _thermal(line: str, where: str, section: str) -> dict[str, Any]:
    equation, fields = _split(line, where)
    left, right = _sides(equation)
    r: dict[str, Any] = {"eq": equation, "names": ([q for q in left if q != "M"], [q for q in right if q != "M"])}
    if section == "two":
        r["kind"] = KIND_TWO_BODY
        r["A"], r["B"], r["C"] = _coefficients(fields, 3, where, line)
    elif section == "three":
        r["kind"] = KIND_FALLOFF
        r.update(zip(("A0", "B0", "C0", "Ai", "Bi", "Ci"), _coefficients(fields, 6, where, line)))
    elif section == "low":
        r["kind"] = KIND_LOW_PRESSURE
        r["A0"], r["B0"], r["C0"] = _coefficients(fields, 3, where, line)
    elif section == "irrev":
        raise ValueError(f"no reverse law for the irreversible reaction {equation}: the reader has none")
    else:
        key = equation.replace(" ", "")
        if key not in SPECIAL:
            raise ValueError(f"unknown special reaction {equation}")
        r["kind"] = KIND_SPECIAL
        r.update(SPECIAL[key])
    _check_m(equation, r["kind"], where)
    return r


def _photo(line: str, where: str, seen: set[tuple[str, int]]) -> dict[str, Any]:
    equation, fields = _split(line, where)
    if len(fields) < 2:
        raise ValueError(f"{where}: a photolysis line needs its absorber and branch index: {line}")
    left, right = _sides(equation)
    if left != [fields[0]]:
        raise ValueError(f"{where}: a photolysis line whose reactant is not its absorber {fields[0]}: {line}")
    try:
        branch = int(fields[1])
    except ValueError:
        raise ValueError(f"{where}: a photolysis branch index must be an integer: {line}") from None
    if (fields[0], branch) in seen:
        raise ValueError(f"{where}: two photolysis lines for absorber {fields[0]}, branch {branch}")
    seen.add((fields[0], branch))
    return {"eq": equation, "names": (left, right), "absorber": fields[0], "branch": branch}


def _extra(line: str, where: str, reactions: list[dict[str, Any]], bulk: str) -> None:
    """A "Troe [a, T3, T1, T2]" or "effi {...}" line qualifies the reaction above it."""
    if not reactions:
        raise ValueError(f"{where}: this line must follow a reaction: {line}")
    try:
        value = ast.literal_eval(line[4:].strip())
    except (ValueError, SyntaxError):
        raise ValueError(f"{where}: this line's list or dict cannot be read: {line}") from None
    if line.startswith("Troe"):
        if len(value) != 4:
            raise ValueError(f"{where}: a Troe line holds 4 numbers (a, T3, T1, T2), got {len(value)}: {line}")
        troe = tuple(float(v) for v in value)
        if not np.all(np.isfinite(troe)):
            raise ValueError(f"{where}: finite rate coefficients expected, got {list(troe)}: {line}")
        reactions[-1]["troe"] = troe
        return
    if bulk not in value:
        raise ValueError(f"{where}: no efficiency for the bulk gas {bulk}; pass another bulk: {line}")
    eff = float(value[bulk])
    if not np.isfinite(eff):
        raise ValueError(f"{where}: finite rate coefficients expected, got {[eff]}: {line}")
    reactions[-1]["eff"] = eff


def read_reactions(path: str | Path, bulk: str = "H2") -> tuple[list[dict], list[dict], int]:
    """Thermal reactions, photolysis branches and the number of condensation lines skipped
    (maxichem/kinvulcan.py:191-223)."""
    section: str | None = None
    ended = False
    skipped = 0
    reactions: list[dict[str, Any]] = []
    photo: list[dict[str, Any]] = []
    seen: set[tuple[str, int]] = set()
    for number, raw in enumerate(Path(path).read_text().split("\n"), 1):
        line, where = raw.strip(), f"{path}:{number}"
        if line.startswith("#"):
            low = line.lower()
            if section is not None and not ended:
                # a comment between thermal sections must be a heading, column titles, a commented reaction or empty
                if not (low.startswith((*KNOWN_HEADINGS, _TITLES)) or "[" in low or not low.strip("# ")):
                    raise ValueError(
                        f"{where}: unknown heading: {line}. The headings are {KNOWN_HEADINGS}; free text goes "
                        "above the first one"
                    )
            if low.startswith(_ENDINGS):
                ended = True
                if low.startswith("# photo"):
                    section = "photo"
                elif not low.startswith("# reverse stops"):
                    section = "skip"
            elif not ended:
                for head, name in _THERMAL:
                    if low.startswith(head):
                        section = name
                        break
            continue
        if ended:
            if "[" not in line:
                continue
            if section == "photo":
                photo.append(_photo(line, where, seen))
            elif section == "skip":
                skipped += 1
            else:
                raise ValueError(f"reaction after '# reverse stops' outside photolysis / condensation: {line}")
        elif line.startswith(("Troe", "effi")):
            _extra(line, where, reactions, bulk)
        elif "[" in line:
            if section is None:
                raise ValueError(f"{where}: a reaction above the first heading has no rate law: {line}")
            reactions.append(_thermal(line, where, section))
    return reactions, photo, skipped


def read_nasa9(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    """The two ranges a1..a7, b1, b2 of a NASA-9 file (ten numbers per range, the eighth unused;
    maxichem/kinvulcan.py:313-331)."""
    a = np.loadtxt(path).ravel()
    if a.size not in _NASA_SIZES:
        raise ValueError(f"{path}: {a.size} numbers, a NASA-9 file holds {_NASA_SIZES}; check the file")
    keep = [0, 1, 2, 3, 4, 5, 6, 8, 9]
    return a[keep].astype(np.float64), a[[10 + k for k in keep]].astype(np.float64)


class VulcanNetwork:
    """A VULCAN reaction list with its NASA-9 data (maxichem/kinvulcan.py:258-299).

    Attributes: ``species``, ``elements``, ``atoms`` [species, elements], ``lo``/``hi`` [species, 9], ``reac``
    (thermal reaction dicts with ``re``/``pr`` species indices), ``S`` [species, reactions], ``photo`` and ``Sp``,
    ``photo_missing``, ``skipped``, ``duplicates``.
    """

    def __init__(
        self, network_file: str | Path, nasa_dir: str | Path, species: list[str] | None = None, bulk: str = "H2"
    ) -> None:
        reactions, photo, self.skipped = read_reactions(network_file, bulk)
        if species is None:
            order: list[str] = []
            for r in reactions:
                for q in r["names"][0] + r["names"][1]:
                    if q not in order:
                        order.append(q)
            if "He" not in order:
                order.append("He")
            species = order
        self.species = list(species)
        twice = sorted({q for q in self.species if self.species.count(q) > 1})
        if twice:
            raise ValueError(f"species names {twice} more than once: a name is one row of the state; drop repeats")
        lo, hi = [], []
        for name in self.species:
            path = Path(nasa_dir) / f"{name}.txt"
            if not path.exists():
                raise FileNotFoundError(f"no NASA-9 file for {name}: {path}; pass the folder that holds it")
            low, high = read_nasa9(path)
            jump = abs(gibbs_rt(low, NASA_SPLIT_T) - gibbs_rt(high, NASA_SPLIT_T))
            if not jump <= MAX_GIBBS_JUMP:
                raise ValueError(
                    f"{path}: G / RT jumps by {jump:.3g} at {NASA_SPLIT_T:g} K, above {MAX_GIBBS_JUMP:g}: the two "
                    "ranges read do not meet; check that the file's first two ranges are the ones below and above it"
                )
            lo.append(low)
            hi.append(high)
        self.lo, self.hi = np.array(lo, np.float64), np.array(hi, np.float64)
        self.reac = reactions
        self._index(reactions, network_file)
        self.duplicates = self._find_duplicates(reactions)
        self.S = self._stoichiometry(reactions)
        counts = [formula(q) for q in self.species]
        self.elements = sorted({e for c in counts for e in c})
        self.atoms = np.array([[c.get(e, 0) for e in self.elements] for c in counts], np.float64)
        bad = np.flatnonzero(np.abs(self.atoms.T @ self.S).max(0) > 0) if self.S.size else []
        if len(bad):
            raise ValueError(f"reactions that do not balance their atoms: {[reactions[j]['eq'] for j in bad]}")
        self.photo_missing = sorted(
            {q for r in photo for q in r["names"][0] + r["names"][1] if q not in self.species}
        )
        for r in photo:
            change: dict[str, int] = {}
            for side, sign in ((r["names"][0], -1), (r["names"][1], 1)):
                for q in side:
                    for e, n in formula(q).items():
                        change[e] = change.get(e, 0) + sign * n
            if any(change.values()):
                raise ValueError(f"photolysis branch {r['eq']} does not balance its atoms")
        if not self.photo_missing:
            self._index(photo, network_file)
        self.photo = photo
        self.Sp = self._stoichiometry(photo)

    def _index(self, reactions: list[dict[str, Any]], where: str | Path) -> None:
        position = {q: i for i, q in enumerate(self.species)}
        for r in reactions:
            missing = [q for q in r["names"][0] + r["names"][1] if q not in position]
            if missing:
                raise ValueError(f"{where}: reaction {r['eq']} names {missing}, not in the species list")
            r["re"] = [position[q] for q in r["names"][0]]
            r["pr"] = [position[q] for q in r["names"][1]]

    def _stoichiometry(self, reactions: list[dict[str, Any]]) -> np.ndarray:
        S = np.zeros((len(self.species), len(reactions)), np.float64)
        for j, r in enumerate(reactions):
            for i in r.get("re", ()):
                S[i, j] -= 1.0
            for i in r.get("pr", ()):
                S[i, j] += 1.0
        return S

    @staticmethod
    def _find_duplicates(reactions: list[dict[str, Any]]) -> list[tuple[str, str]]:
        """Thermal entries with the same species on both sides, either orientation; their rates add."""
        first: dict[Any, str] = {}
        pairs = []
        for r in reactions:
            m = ("M",) if r["kind"] != KIND_TWO_BODY else ()
            key = tuple(sorted((tuple(sorted(r["names"][0])) + m, tuple(sorted(r["names"][1])) + m)))
            if key in first:
                pairs.append((first[key], r["eq"]))
            first.setdefault(key, r["eq"])
        return pairs
