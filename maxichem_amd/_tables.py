"""A network's tables: rate rows, elimination order and the HIP kernel's per-network streams.

The reference builds its tables in maxichem/_tables.py:377-443. What must stay alike is the numerics: the reactions'
rate laws and flags, and a minimum-degree elimination order for the unpivoted LU of M0 = I / (gamma h) - Jn. The
kernel layout here is this build's own, for one 64-lane wavefront per cell:

- the trailing NT = min(NS, 64) rows and columns of the elimination order form a dense block held in registers, lane t
  owning row K0 + t, with K0 = NS - NT; the leading K0 rows (all columns) and the trailing rows' first K0 columns are
  dense in LDS;
- each lane owns the reactions r = lane + 64 q: their rate coefficients stay in its registers for the call, and it
  writes their rates of progress v and their partial derivatives dv/dn ("dv") to LDS;
- S v is gathered per species row from load-balanced pieces; the Jacobian's entries (-Jn) are sums of dv terms,
  gathered into whole entries per lane (greedy longest-first over the 64 lanes), leading entries straight to LDS and
  the trailing block through a small staging area, CHUNK columns at a time.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from ._constants import (
    FALLOFF_KINDS,
    FL_EFF,
    FL_FALLOFF,
    FL_INERT,
    FL_NF_SHIFT,
    FL_NR_SHIFT,
    FL_PHOTO,
    FL_SECOND_K0,
    FL_THREE_BODY,
    FL_TROE,
    FL_VM11,
    KIND_TWO_BODY,
    RATE_ROW,
    THREE_BODY_KINDS,
    WAVE,
)
from .kinvulcan import VulcanNetwork, _check_m

MAX_SPECIES = 128  # the lead part of a state vector is one register per lane: K0 = NS - 64 <= 64
CHUNK = 8  # trailing-block columns assembled per staging pass
# Term word of the Jacobian stream (uint32): dv index [0, 16), end of entry 16, x2 17, destination (LDS double
# offset) [18, 31), sign 31. The sign sits in the word's sign bit, where a kernel takes it with one AND to flip a
# double's sign; x2 is an exponent step. A term is then exact without a compare or select.
T_IDX_BITS, T_END, T_DBL, T_DEST, T_NEG = 16, 16, 17, 18, 31
T_DEST_BITS = 13
# S v stream word (uint32): reaction [0, 11), end of piece 11, x2 12, piece index [13, 31), sign 31
SV_R_BITS, SV_END, SV_DBL, SV_PIECE, SV_NEG = 11, 11, 12, 13, 31
SV_PIECE_BITS = 18


def _log(a: float) -> float:
    return float(np.log(a)) if a > 0 else -np.inf


def rate_row(r: dict[str, Any]) -> tuple[list[float], int]:
    """A thermal reaction's law row (RATE_ROW columns) and law flags (maxichem/_tables.py:69-106 holds the same row)."""
    if r["kind"] == KIND_TWO_BODY:
        k0, k1, ki = (r["A"], r["B"], r["C"]), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0)
    else:
        k0 = (r["A0"], r["B0"], r["C0"])
        k1 = (r["A1"], r["B1"], r["C1"]) if "A1" in r else (0.0, 0.0, 0.0)
        ki = (r["Ai"], r["Bi"], r["Ci"]) if r["kind"] in FALLOFF_KINDS else (1.0, 0.0, 0.0)
    if min(k0[0], k1[0], ki[0]) < 0:
        raise ValueError(f"negative Arrhenius prefactor in {r['eq']}: the kernel takes ln A")
    if "rev" in r:
        raise ValueError(
            f"reaction {r['eq']} has an explicit reverse law, which would be ignored: every reverse comes from NASA-9; "
            "remove its 'rev' entry"
        )
    row = [_log(k0[0]), k0[1], k0[2], _log(k1[0]), k1[1], k1[2], _log(ki[0]), ki[1], ki[2]]
    row += list(r.get("troe", (0.0, 1.0, 1.0, 0.0))) + [float(r.get("eff", 1.0))]
    fl = (
        FL_FALLOFF * (r["kind"] in FALLOFF_KINDS)
        | FL_SECOND_K0 * ("A1" in r)
        | FL_TROE * ("troe" in r)
        | FL_EFF * ("eff" in r)
        | FL_VM11 * (r.get("width") == "vm11")
        | FL_THREE_BODY * (r["kind"] in THREE_BODY_KINDS)
    )
    return row, fl


def min_degree_order(entries: set[tuple[int, int]], n: int) -> list[int]:
    """Greedy minimum-degree elimination order on the symmetrised pattern, ties to the lower index.

    The textbook algorithm (e.g. Duff, Erisman & Reid, Direct Methods for Sparse Matrices, ch. 7); the reference
    orders its LU the same way (maxichem/_tables.py:35-50), so both factor M0 in the same order.
    """
    nbr = [set() for _ in range(n)]
    for i, j in entries:
        if i != j:
            nbr[i].add(j)
            nbr[j].add(i)
    alive = set(range(n))
    order = []
    while alive:
        k = min(alive, key=lambda q: (len(nbr[q] & alive), q))
        order.append(k)
        clique = nbr[k] & alive
        for a in clique:
            nbr[a] |= clique - {a}
        alive.discard(k)
    return order


def _balance(sizes: list[int], lanes: int) -> list[list[int]]:
    """Items (by index) over lanes, longest first onto the lane with the least work so far."""
    load = [0] * lanes
    out: list[list[int]] = [[] for _ in range(lanes)]
    for it in sorted(range(len(sizes)), key=lambda q: (-sizes[q], q)):
        lane = min(range(lanes), key=lambda q: (load[q], q))
        out[lane].append(it)
        load[lane] += sizes[it]
    return out


def _pack_rounds(per_lane: list[list[int]], pad: int) -> np.ndarray:
    rounds = max(1, max(len(w) for w in per_lane))
    words = np.full((rounds, len(per_lane)), pad, np.uint32)
    for lane, w in enumerate(per_lane):
        words[: len(w), lane] = w
    return words


@dataclass
class Tables:
    """A network's tables for the HIP kernels and for the oracle (network order where a name says so)."""

    ns: int
    nr: int
    nrt: int
    nj: int
    nel: int
    k0: int
    nt: int
    perm: np.ndarray  # int32 [ns]: network species of elimination row k
    rx: np.ndarray  # int32 [nr, 6]: network species of each reaction's sides, -1 pads
    rc: np.ndarray  # float64 [nr, RATE_ROW]
    rf: np.ndarray  # int32 [nr]
    nasa: np.ndarray  # float64 [2, ns, 9], network order
    atoms: np.ndarray  # float64 [ns, nel], network order
    species: list[str]
    elements: list[str]
    has_troe: bool
    has_eff: bool
    # kernel streams
    rxe: np.ndarray = field(default=None)  # uint64 [nr]: 6 elimination-order species (8 bits, pad ns), dv offset
    rord: np.ndarray = field(default=None)  # int32 [nr]: the rate coefficients' order, reactions grouped by law
    ndv: int = 0
    diag_words: np.ndarray = field(default=None)  # uint32 [rounds, 64]
    lead_words: np.ndarray = field(default=None)  # uint32 [rounds, 64]
    chunk_words: list = field(default_factory=list)  # per chunk (lead columns, then trailing): uint32 [rounds, 64]
    nlch: int = 0  # the chunks of lead columns
    sv_words: np.ndarray = field(default=None)  # uint32 [rounds, 64]
    npart: int = 0
    row_pieces: np.ndarray = field(default=None)  # uint32 [ns]: first piece | count << 16
    jac_terms: int = 0
    sv_terms: int = 0
    live: int = 0

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.perm, self.rx, self.rc, self.rf, self.nasa, self.atoms):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    # ---- the kernel's LDS layout (doubles per cell) ----
    def layout(self, mixed: bool = False, rates_lds: bool = True) -> dict[str, int]:
        """Offsets (doubles) in a cell's LDS region: LEAD [K0][NS], n [NS + 1], dv [NDV + 1], then a
        union of what the rhs uses (v [NR + 1], pieces [NPART]) with what the Jacobian and the factorization use after
        it (the trailing diagonal [NT], staging [CHUNK][NT], column-major so that a lane's writes and reads of its row
        fall in distinct banks); G / RT [NS + 1] sits at its start while the rates are computed; a slot per lane
        (TRASH [WAVE]) where the streams write the sums of the words that end no entry. mixed: the mixed-precision form's exchange vectors [2][NS + 1] follow; then the rate coefficients
        KF, KR [NR] (kept for the call) unless rates_lds is False (they are then in global scratch). Everything the
        Jacobian streams address lies before the optional parts, so the streams hold for every layout."""
        nvec = self.k0 * self.ns  # the trailing rows' lead columns (A21) are held in registers, one row per lane
        dv = nvec + self.ns + 1
        union = dv + self.ndv + 1
        v = union
        part = v + self.nr + 1  # v[NR] = 0: the S v stream's padding term
        diag = union
        stage = diag + self.nt
        size = max(part + self.npart, stage + self.nt * CHUNK, union + self.ns + 1)
        trash = size  # a slot per lane for the stream words that end no entry (written, never read)
        size += WAVE
        xv = size  # the mixed-precision form's exchange vectors [2][NS + 1]
        if mixed:
            size += 2 * (self.ns + 1)
        kf = size
        kr = kf + self.nr
        if rates_lds:
            size = kr + self.nr
        size += size & 1  # keep the next cell's region 16-byte aligned
        return {"NVEC": nvec, "KF": kf, "KR": kr, "DV": dv, "UNION": union, "V": v, "PART": part,
                "DIAG": diag, "STAGE": stage, "TRASH": trash, "XV": xv, "SIZE": size}


def build(net: VulcanNetwork, photo: bool = False) -> Tables:
    """The tables of a network, with its photolysis branches after the thermal reactions when photo."""
    reac = list(net.reac) + (list(net.photo) if photo else [])
    S = np.asarray(net.S, np.float64)
    if photo:
        S = np.hstack([S, np.asarray(net.Sp, np.float64)])
    ns, nr = S.shape
    nrt = len(net.reac)
    if ns > MAX_SPECIES:
        raise ValueError(f"the kernels hold at most {MAX_SPECIES} species, one wavefront per cell; this has {ns}")
    rx = np.full((nr, 6), -1, np.int32)
    rc = np.zeros((nr, RATE_ROW), np.float64)
    rf = np.zeros(nr, np.int32)
    for j, r in enumerate(reac):
        re, pr = list(r["re"]), list(r["pr"])
        if max(len(re), len(pr)) > 3:
            raise ValueError(f"at most 3 species per side: {r['eq']}")
        rx[j, : len(re)] = re
        rx[j, 3 : 3 + len(pr)] = pr  # a photolysis branch's products too: kr = 0 makes its reverse side vanish
        if j >= nrt:
            rf[j] = FL_PHOTO | len(re) << FL_NF_SHIFT | len(pr) << FL_NR_SHIFT
            continue
        _check_m(r["eq"], r["kind"], f"reaction {j}")  # the reader's M rule, also for a hand-built reaction
        row, fl = rate_row(r)
        rc[j] = row
        rf[j] = fl | len(re) << FL_NF_SHIFT | len(pr) << FL_NR_SHIFT
    if np.abs(S).max(initial=0) > 2:
        raise ValueError("|S| <= 2 per species and reaction")
    live = [r for r in range(nr) if np.any(S[:, r] != 0)]
    alive = set(live)
    for r in range(nr):
        if r not in alive:
            rf[r] |= FL_INERT  # its net column is zero: no S v term and no Jacobian term
    # d v_r / d n_q terms: (reaction, side, occurrence) -> species q; each is one dv value
    dv_of: list[tuple[int, int, int]] = []  # (r, side, m)
    rdv = np.zeros(nr, np.int64)
    for r in range(nr):
        rdv[r] = len(dv_of)
        if r not in alive:
            continue
        for side in (0, 1) if r < nrt else (0,):
            for m in range(3):
                if rx[r, 3 * side + m] >= 0:
                    dv_of.append((r, side, m))
    ndv = len(dv_of)
    # entries of -Jn: (p, q) -> [(dv index, coefficient)], coefficient = -S[p, r] (forward) or +S[p, r] (reverse
    # side's derivative is stored as +kr prod n, entering d v / d n with a minus)
    entries: dict[tuple[int, int], list[tuple[int, int]]] = {}
    for d, (r, side, m) in enumerate(dv_of):
        q = int(rx[r, 3 * side + m])
        for p in np.flatnonzero(S[:, r]):
            c = -int(S[p, r]) if side == 0 else int(S[p, r])
            entries.setdefault((int(p), q), []).append((d, c))
    pattern = set(entries) | {(i, i) for i in range(ns)}
    perm = min_degree_order(pattern, ns)
    pos = np.empty(ns, np.int64)
    pos[perm] = np.arange(ns)
    nt = min(ns, WAVE)
    k0 = ns - nt
    # the unpivoted LU's filled pattern in the elimination order; the trailing rows' part as one bit mask per row
    # (bit j: column k0 + j), which the trailing factorization reads to leave out updates by structural zeros
    fill = np.zeros((ns, ns), bool)
    for i, j in pattern:
        fill[pos[i], pos[j]] = True
    for k in range(ns):
        rows = np.flatnonzero(fill[k + 1 :, k]) + k + 1
        cols = np.flatnonzero(fill[k, k + 1 :]) + k + 1
        fill[np.ix_(rows, cols)] = True
    trail_fill = [sum(1 << int(j) for j in np.flatnonzero(fill[k0 + t, k0:])) for t in range(nt)]
    tab = Tables(
        ns=ns, nr=nr, nrt=nrt, nj=nr - nrt, nel=len(net.elements), k0=k0, nt=nt,
        perm=np.asarray(perm, np.int32), rx=rx, rc=rc, rf=rf,
        nasa=np.stack([np.asarray(net.lo, np.float64), np.asarray(net.hi, np.float64)]),
        atoms=np.asarray(net.atoms, np.float64), species=list(net.species), elements=list(net.elements),
        has_troe=any("troe" in r for r in net.reac), has_eff=any("eff" in r for r in net.reac),
    )  # fmt: skip
    tab.ndv, tab.live = ndv, len(live)
    tab.trail_fill = trail_fill
    if ndv >= (1 << T_IDX_BITS) - 1 or nr >= (1 << SV_R_BITS):
        raise ValueError(f"the network is too large for the kernel's streams: {ndv} derivative terms, {nr} reactions")
    # reaction words: species in elimination order (pad ns), dv offset
    rxe = np.zeros(nr, np.uint64)
    for r in range(nr):
        w = 0
        for m in range(6):
            s = int(rx[r, m])
            w |= (int(pos[s]) if s >= 0 else ns) << (8 * m)
        w |= int(rdv[r]) << 48
        rxe[r] = w
    tab.rxe = rxe
    # the rate coefficients are computed in this order, 64 reactions a round: grouped by law (Troe falloff, falloff,
    # a second low-pressure term, plain Arrhenius, photolysis), a round takes the code of fewer laws
    law = np.where(rf & FL_PHOTO, 4, np.where(rf & FL_TROE, 0, np.where(rf & FL_FALLOFF, 1,
                                                                          np.where(rf & FL_SECOND_K0, 2, 3))))
    tab.rord = np.argsort(law, kind="stable").astype(np.int32)
    lay = tab.layout()
    reach = max(lay["STAGE"] + nt * CHUNK, lay["DIAG"] + nt)  # the furthest LDS offset a stream word addresses
    if reach >= (1 << T_DEST_BITS):
        raise ValueError(f"a cell's LDS region of {reach} doubles exceeds the stream's destination field")
    zero = ndv  # dv[ndv] holds 0.0: the padding words' term

    nlch = -(-k0 // CHUNK)  # chunks of the trailing rows' lead columns, before the trailing columns' chunks

    def dest(i: int, j: int) -> tuple[int, int]:
        """(stream, LDS offset): stream -2 the diagonal, -1 the rest of the lead rows, else a chunk of the trailing
        rows: chunk c < nlch holds lead columns c CHUNK .. c CHUNK + CHUNK - 1, chunk nlch + c trailing columns
        k0 + c CHUNK ... A chunk's entries land in the staging area, whence each lane takes its row into registers.

        The diagonal entries gather every derivative term of their species, so they are the long ones; they take
        a stream of their own, and the chunks' entries stay short."""
        if i == j:
            return -2, (i * ns + i) if i < k0 else lay["DIAG"] + i - k0
        if i < k0:
            return -1, i * ns + j
        t = i - k0
        if j < k0:
            return j // CHUNK, lay["STAGE"] + (j % CHUNK) * nt + t
        c = nlch + (j - k0) // CHUNK
        return c, lay["STAGE"] + ((j - k0) % CHUNK) * nt + t  # column-major: a column's rows are consecutive

    groups: dict[int, list[tuple[int, list[tuple[int, int]]]]] = {}
    for (p, q), terms in sorted(entries.items(), key=lambda e: (pos[e[0][0]], pos[e[0][1]])):
        g, off = dest(int(pos[p]), int(pos[q]))
        groups.setdefault(g, []).append((off, terms))
    tab.jac_terms = sum(len(t) for t in entries.values())

    def stream(items: list[tuple[int, list[tuple[int, int]]]]) -> np.ndarray:
        lanes = _balance([len(t) for _, t in items], WAVE)
        per_lane = []
        for its in lanes:
            w = []
            for it in sorted(its):
                off, terms = items[it]
                for e, (d, c) in enumerate(terms):
                    word = d | (c < 0) << T_NEG | (abs(c) == 2) << T_DBL
                    if e == len(terms) - 1:
                        word |= 1 << T_END | off << T_DEST
                    w.append(word)
            per_lane.append(w)
        return _pack_rounds(per_lane, zero)

    tab.diag_words = stream(groups.get(-2, []))
    tab.lead_words = stream(groups.get(-1, []))
    tab.chunk_words = [stream(groups.get(c, [])) for c in range(nlch + -(-nt // CHUNK))]
    tab.nlch = nlch
    # S v pieces: row i (elimination order) of S v from its reactions, cut into pieces of at most an even share
    rows = [[(r, int(S[perm[i], r])) for r in live if S[perm[i], r] != 0] for i in range(ns)]
    total = sum(len(t) for t in rows)
    share = max(1, -(-total // WAVE))
    pieces: list[list[tuple[int, int]]] = []
    row_pieces = np.zeros(ns, np.uint32)
    for i, t in enumerate(rows):
        first = len(pieces)
        for c in range(0, len(t), share):
            pieces.append(t[c : c + share])
        row_pieces[i] = first | (len(pieces) - first) << 16
    tab.npart = max(1, len(pieces))
    if tab.npart >= (1 << SV_PIECE_BITS):
        raise ValueError(f"{tab.npart} S v pieces exceed the stream's piece field")
    tab.row_pieces = row_pieces
    tab.sv_terms = total
    lanes = _balance([len(p) for p in pieces], WAVE)
    per_lane = []
    for its in lanes:
        w = []
        for it in sorted(its):
            for e, (r, c) in enumerate(pieces[it]):
                word = r | (c < 0) << SV_NEG | (abs(c) == 2) << SV_DBL
                if e == len(pieces[it]) - 1:
                    word |= 1 << SV_END | it << SV_PIECE
                w.append(word)
        per_lane.append(w)
    tab.sv_words = _pack_rounds(per_lane, nr)  # padding: reaction index nr, whose v slot holds 0
    return tab
