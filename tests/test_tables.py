"""The kernel tables on the CPU: the LDS layout's live regions stay apart, and the streams decode back to the
network's S v and -Jn (the reference checks its streams the same way, tests/test_tables_source.py)."""

import numpy as np
import pytest

from maxichem_amd import _tables, standin
from maxichem_amd._tables import (CHUNK, SV_DBL, SV_END, SV_NEG, SV_PIECE, SV_PIECE_BITS, SV_R_BITS, T_DEST,
                                  T_DEST_BITS, T_END)


def field(w, at: int, bits: int) -> int:
    """The bits [at, at + bits) of a stream word."""
    return (int(w) >> at) & ((1 << bits) - 1)


def _nets(read_network):
    return [(read_network(s), p) for s in ("synthetic", "synthetic_small", "synthetic_decay") for p in (False, True)
            ] + [(standin.load("NCHO"), True), (standin.load("SNCHO"), True)]  # fmt: skip


def test_layout_regions_that_live_together_do_not_overlap(read_network):
    for net, photo in _nets(read_network):
        tab = _tables.build(net, photo)
        L = tab.layout()
        regions = {
            "lead": (0, L["NVEC"]), "n": (L["NVEC"], L["NVEC"] + tab.ns + 1),
            "dv": (L["DV"], L["DV"] + tab.ndv + 1), "v": (L["V"], L["V"] + tab.nr + 1),
            "kf": (L["KF"], L["KF"] + tab.nr), "kr": (L["KR"], L["KR"] + tab.nr),
            "part": (L["PART"], L["PART"] + tab.npart), "diag": (L["DIAG"], L["DIAG"] + tab.nt),
            "stage": (L["STAGE"], L["STAGE"] + tab.nt * CHUNK),
        }  # fmt: skip
        # live at once: the rhs with the Jacobian's dv; the Jacobian's assembly; the factorization
        groups = [("lead", "n", "kf", "kr", "dv", "v", "part"),
                  ("lead", "n", "kf", "kr", "dv", "diag", "stage")]
        for g in groups:
            spans = sorted(regions[k] for k in g)
            for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
                assert a1 <= b0, (g, spans)
        assert max(b for _, b in regions.values()) <= L["SIZE"]
        assert L["UNION"] + tab.ns + 1 <= L["SIZE"]


def test_streams_decode_to_s_v_and_minus_jn(read_network):
    for net, photo in _nets(read_network):
        tab = _tables.build(net, photo)
        S = np.hstack([net.S, net.Sp]) if photo else np.asarray(net.S)
        perm, ns = tab.perm, tab.ns
        # S v: each piece is a signed sum of reactions; each row gathers its pieces
        pieces = {}
        for lane in range(64):
            acc = {}
            for w in tab.sv_words[:, lane]:
                r = field(w, 0, SV_R_BITS)
                c = (2 if (w >> SV_DBL) & 1 else 1) * (-1 if (w >> SV_NEG) & 1 else 1)
                if r < tab.nr:
                    acc[r] = acc.get(r, 0) + c
                if (w >> SV_END) & 1:
                    pieces[field(w, SV_PIECE, SV_PIECE_BITS)] = acc
                    acc = {}
        for i in range(ns):
            first, count = int(tab.row_pieces[i]) & 0xFFFF, int(tab.row_pieces[i]) >> 16
            row = {}
            for p in range(first, first + count):
                for r, c in pieces[p].items():
                    row[r] = row.get(r, 0) + c
            want = {r: int(S[perm[i], r]) for r in range(tab.nr) if S[perm[i], r] != 0}
            assert row == want
        # -Jn: every destination written once, from dv terms; the number of terms matches
        dests = []
        terms = 0
        for words in [tab.diag_words, tab.lead_words, *tab.chunk_words]:
            for lane in range(64):
                for w in words[:, lane]:
                    if (int(w) & 0xFFFF) != tab.ndv:
                        terms += 1
                    if (w >> T_END) & 1:
                        dests.append((id(words), field(w, T_DEST, T_DEST_BITS)))
        assert len(dests) == len(set(dests))
        assert terms == tab.jac_terms


def test_every_layout_keeps_the_offsets_the_streams_address(read_network):
    # the Jacobian streams carry LDS offsets built with the default layout: the optional parts (the mixed form's
    # vectors, the rates in LDS) must not move anything the streams or the rhs address
    from maxichem_amd import _tables, standin

    for net in (read_network("synthetic"), standin.load("NCHO")):
        tab = _tables.build(net, True)
        base = tab.layout()
        fixed = ("NVEC", "DV", "UNION", "V", "PART", "DIAG", "STAGE")
        for mixed in (False, True):
            for rates in (False, True):
                lay = tab.layout(mixed, rates)
                assert all(lay[k] == base[k] for k in fixed)
                assert lay["SIZE"] >= max(lay["XV"] + (2 * (tab.ns + 1) if mixed else 0),
                                          (lay["KR"] + tab.nr) if rates else 0)  # fmt: skip


def test_chunks_stage_the_trailing_rows_by_column(read_network):
    # the trailing rows' entries reach the staging area chunk by chunk: the first nlch chunks hold the lead columns
    # (kept in registers, a21), the others the trailing columns; each word's destination names its row and column
    for net, photo in _nets(read_network):
        tab = _tables.build(net, photo)
        L = tab.layout()
        S = np.hstack([net.S, net.Sp]) if photo else np.asarray(net.S)
        pos = np.empty(tab.ns, np.int64)
        pos[tab.perm] = np.arange(tab.ns)
        assert tab.nlch == -(-tab.k0 // CHUNK) and len(tab.chunk_words) == tab.nlch + -(-tab.nt // CHUNK)
        got = set()
        for c, words in enumerate(tab.chunk_words):
            for w in words.ravel():
                if (w >> T_END) & 1:
                    off = field(w, T_DEST, T_DEST_BITS) - L["STAGE"]
                    assert 0 <= off < CHUNK * tab.nt
                    slot, t = divmod(off, tab.nt)
                    col = c * CHUNK + slot if c < tab.nlch else tab.k0 + (c - tab.nlch) * CHUNK + slot
                    assert (col < tab.k0) == (c < tab.nlch) and col != tab.k0 + t
                    got.add((tab.k0 + t, col))
        # every off-diagonal entry of a trailing row that a reaction couples, in elimination order
        want = set()
        for r in range(S.shape[1]):
            rows = np.flatnonzero(S[:, r])
            if rows.size == 0:
                continue
            for q in (list(net.reac) + (list(net.photo) if photo else []))[r]["re"] + (
                    (list(net.reac) + (list(net.photo) if photo else []))[r]["pr"] if r < len(net.reac) else []):
                for p_ in rows:
                    i, j = int(pos[p_]), int(pos[q])
                    if i >= tab.k0 and i != j:
                        want.add((i, j))
        assert got == want


def test_tables_refuse_networks_beyond_the_kernels_limits(read_network):
    # one wavefront per cell holds at most 128 species (the lead part one register per lane), a reaction at most
    # three species a side, and a stoichiometric coefficient at most 2: each refused by name before any device work
    import copy
    import types

    net = read_network("synthetic_small")
    fields = {k: getattr(net, k) for k in ("reac", "photo", "elements", "atoms", "lo", "hi", "species", "S", "Sp")}
    big = types.SimpleNamespace(**dict(fields, S=np.zeros((_tables.MAX_SPECIES + 1, len(net.reac)))))
    with pytest.raises(ValueError, match=f"at most {_tables.MAX_SPECIES} species, one wavefront per cell"):
        _tables.build(big)
    reac = copy.deepcopy(list(net.reac))
    reac[0]["re"] = list(reac[0]["re"]) + [0] * (4 - len(reac[0]["re"]))
    with pytest.raises(ValueError, match="at most 3 species per side"):
        _tables.build(types.SimpleNamespace(**dict(fields, reac=reac)))
    S = np.array(net.S, dtype=np.float64)
    S[0, 0] = 3.0
    with pytest.raises(ValueError, match=r"\|S\| <= 2 per species and reaction"):
        _tables.build(types.SimpleNamespace(**dict(fields, S=S)))


def test_trailing_fill_is_the_symbolic_lu_of_minus_jn(read_network):
    # the trailing rows' filled pattern the factorization reads, against an elimination written here on the pattern of
    # -Jn (a species whose net stoichiometry a reaction changes, against each species the reaction's rate depends on)
    # plus the diagonal; and a dense numerical LU of a random matrix on that pattern is zero wherever it has no entry
    for net, photo in _nets(read_network):
        tab = _tables.build(net, photo)
        ns, k0 = tab.ns, tab.k0
        pos = np.argsort(tab.perm)
        entry = np.eye(ns, dtype=bool)
        for r in range(tab.nr):
            x = tab.rx[r]
            net_change = np.zeros(ns, int)
            for s_ in x[:3]:
                if s_ >= 0:
                    net_change[s_] -= 1
            for s_ in x[3:]:
                if s_ >= 0:
                    net_change[s_] += 1
            depends = {int(s_) for s_ in (x if r < tab.nrt else x[:3]) if s_ >= 0}
            for p_ in np.flatnonzero(net_change):
                for q in depends:
                    entry[pos[p_], pos[q]] = True
        want = entry.copy()
        for k in range(ns):  # fill by elimination, one pivot at a time
            below = want[k + 1 :, k].nonzero()[0] + k + 1
            right = want[k, k + 1 :].nonzero()[0] + k + 1
            want[np.ix_(below, right)] = True
        got = np.array([[(m >> j) & 1 for j in range(tab.nt)] for m in tab.trail_fill], bool)
        assert np.array_equal(got, want[k0:, k0:])
        rng = np.random.default_rng(1)
        a = np.where(entry, rng.uniform(0.5, 1.5, (ns, ns)), 0.0) + ns * np.eye(ns)
        for k in range(ns):
            a[k + 1 :, k] /= a[k, k]
            a[k + 1 :, k + 1 :] -= np.outer(a[k + 1 :, k], a[k, k + 1 :])
        assert not a[k0:, k0:][~got].any()
