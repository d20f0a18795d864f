"""Every rate law the reader takes, against closed forms evaluated here from the cases of
tests/golden/data/rate_laws.json (the reference checks the same laws in its tests/test_rate_laws.py and
tests/test_collision_limit.py).

Three checks:
- the oracle's rate coefficients (forward and reverse, the collision cap applied) against the closed forms, directly;
- the call on the oracle and on the GPU: each case is an isomerisation A <-> B beside spectators, so A relaxes to the
  pair's equilibrium at lambda = cap (kf + kr) times the spectators' densities, and after dt = 1 / lambda the
  distance to equilibrium has shrunk by exactly e;
- falloff rows at the ends of Pr, and the collision cap's closed forms on the reference's collision network.
"""

import json
import math
from pathlib import Path

import helpers
import numpy as np
import pytest
from helpers import BACKENDS

CASES = json.loads((Path(__file__).resolve().parent / "golden" / "data" / "rate_laws.json").read_text())
K_COLL, TRACE = CASES["k_coll"], CASES["trace"]
CONDITIONS = [tuple(c) for c in CASES["conditions"]]
RTOL = 1e-9  # the call's tolerance: its integration error on lambda dt = 1 stays far below RATE_TOLERANCE
RATE_TOLERANCE = 1e-6
COEFFICIENT_TOLERANCE = 1e-12  # the oracle's coefficients against the closed forms: rounding of exp and log
# VULCAN's section headings, by the case files' section keys
SECTION_HEADINGS = {"two": "# Two-body Reactions", "three": "# 3-body and Disscoiation Reactions",
                    "low": "# 3-body reactions without high-pressure rates", "special": "# special cases"}  # fmt: skip


# ---- the laws, written out (Troe 1983 and Venot et al. 2012 eq. C.2-C.5; Visscher & Moses 2011) ----


def modified_arrhenius(term, T):
    A, B, C = term
    return A * T**B * math.exp(-C / T)


def log10_fcent(troe, T):
    a, t3, t1, t2 = troe
    return math.log10((1.0 - a) * math.exp(-T / t3) + a * math.exp(-T / t1) + math.exp(-t2 / T))


def broadening(troe, vm11, T, pr):
    """F of the falloff k = F k0 / (1 + Pr): log10 F = log10 Fc / (1 + f^2), f from log10 Pr."""
    lfc = log10_fcent(troe, T)
    width = 0.75 - 1.27 * lfc
    if vm11:
        f = math.log10(pr) / width
    else:
        z = math.log10(pr) - 0.4 - 0.67 * lfc
        f = z / (width - 0.14 * z)
    return 10.0 ** (lfc / (1.0 + f * f))


def broadening_at_low_pressure(troe, T):
    """F as Pr -> 0 in Troe's form: f -> -1 / 0.14, so F = Fc^(0.14^2 / (1 + 0.14^2))."""
    return 10.0 ** (log10_fcent(troe, T) * 0.14**2 / (1.0 + 0.14**2))


def coefficient(law, T, nd):
    """k of a case's law at T and total density nd, before a third body's nd is folded in."""
    k = sum(modified_arrhenius(term, T) for term in law["low"])
    if law.get("high") is not None:
        pr = k * nd / modified_arrhenius(law["high"], T)
        k = k / (1.0 + pr) * (broadening(law["troe"], law.get("vm11", False), T, pr) if law.get("troe") else 1.0)
    return k * law.get("eff", 1.0)


# ---- a network of the cases, and each case's cell ----


def thermo_of(species, dH):
    """Constant-Cp NASA-9 ranges: every species at H/R = 0 except products at their dH; He monatomic."""
    out = {q: helpers.constant_cp(0.0, 2.5 if q == "He" else 3.5) for q in species}
    out.update({q: helpers.constant_cp(h) for q, h in dH.items()})
    return out


def build(cases, folder, lines_of):
    by_section = {}
    for i, case in enumerate(cases):
        by_section.setdefault(case["section"], []).append(f"{2 * i + 1} " + "\n".join(lines_of(case)))
    body = "\n\n".join(SECTION_HEADINGS[s] + "\n" + "\n".join(rows) for s, rows in by_section.items())
    text = "# one isomerisation per rate law\n\n" + body + "\n\n# reverse stops\n"
    species = {"He", "H2"} | {q for c in cases for q in (c["from"], c["to"], *c.get("with", ()), *c.get("bath", {}))}
    dH = {c["to"]: c.get("dH", -1000.0) for c in cases}
    thermo = thermo_of(sorted(species), dH)
    return helpers.write_network(folder, text, thermo), thermo


def start_state(case, species):
    """The cell's mole fractions: the reactant at 0.1 in He 0.5 and H2 0.4, or traced at TRACE in its bath."""
    x = np.zeros(len(species))
    if case.get("traced"):
        bath = case["bath"]
        for q, f in bath.items():
            x[species.index(q)] = f * (1.0 - TRACE) / sum(bath.values())
        x[species.index(case["from"])] = TRACE
    else:
        x[species.index("He")], x[species.index("H2")], x[species.index(case["from"])] = 0.5, 0.4, 0.1
    return x


def folded(case, T, P, thermo):
    """The case's forward and reverse coefficients as the call uses them (a third body's nd folded in, the
    collision cap applied; kr = kf exp(dG/RT) (k_B T / P0)^(products - reactants), molecules without M) and the
    isomer pair's equilibrium constant K (infinite for the special association, whose reverse is negligible)."""
    nd = float(helpers.number_density(T, P))
    k = coefficient(case["k"], T, nd)
    third = case["section"] != "two"
    g = {q: helpers.gibbs_rt(a, T) for q, a in thermo.items()}
    association = case["section"] == "special"
    partners = case["with"] if association else []  # spectators appear on both sides otherwise
    dg = g[case["to"]] - g[case["from"]] - sum(g[q] for q in partners)
    lkt = math.log(helpers.BOLTZMANN * T / helpers.STANDARD_PRESSURE)
    ratio = math.exp(dg - len(partners) * lkt)  # kr / kf
    if association:  # the reverse against the forward at the end of a call, traced in its bath
        assert ratio / nd / (1.0 - TRACE) * 2.0 * math.e < 1e-12
    molecules = 1 + len(case["with"]) + int(third)
    cap = 1.0
    if molecules >= 2:
        cap = min(1.0, K_COLL / (max(k, k * ratio) * nd ** (molecules - 2)))
    assert (cap < 1.0) == case.get("capped", False), case["name"]
    kf = cap * k * (nd if third else 1.0)
    return kf, kf * ratio, (math.inf if association else 1.0 / ratio)


def spectators(case, x, species, T, P):
    nd = float(helpers.number_density(T, P))
    return math.prod(nd * x[species.index(q)] for q in case["with"])


def all_cases(troe: bool):
    return [c for c in CASES["laws"] if troe or not c["k"].get("troe")]


@pytest.mark.parametrize("troe", [True, False], ids=["troe", "plain"])
def test_oracle_coefficients_are_the_closed_forms(tmp_path, troe):
    import oracle

    from maxichem_amd import _tables

    cases = all_cases(troe)
    net, thermo = build(cases, tmp_path, lambda c: c["lines"])
    orc = oracle.Oracle(_tables.build(net, photo=False))
    rows = {r["eq"]: i for i, r in enumerate(net.reac)}
    for case in cases:
        q = next(i for eq, i in rows.items() if eq.split("->")[0].split()[0] == case["from"])
        for T, P in CONDITIONS:
            kf, kr = orc.rates(T, P, np.zeros(0))
            want_f, want_r, _ = folded(case, T, P, thermo)
            assert kf[q] == pytest.approx(want_f, rel=COEFFICIENT_TOLERANCE), (case["name"], T, P)
            assert kr[q] == pytest.approx(want_r, rel=COEFFICIENT_TOLERANCE), (case["name"], T, P)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("troe", [True, False], ids=["troe", "plain"])
def test_each_law_relaxes_its_isomer_by_e_in_one_time_constant(tmp_path, kind, troe):
    cases = all_cases(troe)
    net, thermo = build(cases, tmp_path, lambda c: c["lines"])
    cells = [(case, T, P) for case in cases for T, P in CONDITIONS]
    x0 = np.array([start_state(case, net.species) for case, _, _ in cells])
    lam, x_eq = np.empty(len(cells)), np.empty(len(cells))
    for b, (case, T, P) in enumerate(cells):
        kf, kr, K = folded(case, T, P, thermo)
        lam[b] = (kf + (kr if math.isfinite(K) else 0.0)) * spectators(case, x0[b], net.species, T, P)
        x_eq[b] = x0[b, net.species.index(case["from"])] / (1.0 + K)
    dt = 1.0 / lam
    run = helpers.backend(kind, net)
    out = run(x0, [T for _, T, _ in cells], [P for _, _, P in cells], dt, 1e-3 * dt, rtol=RTOL)
    run.close()
    assert out["done"].all()
    for b, (case, T, P) in enumerate(cells):
        i = net.species.index(case["from"])
        shrink = (out["x"][b, i] - x_eq[b]) / (x0[b, i] - x_eq[b])
        assert abs(math.log(shrink) + 1.0) < RATE_TOLERANCE, (case["name"], T, P, shrink)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("troe", [True, False], ids=["troe", "plain"])
def test_falloff_at_the_ends_of_pr(tmp_path, kind, troe):
    lim = CASES["pr_limits"]
    rows = [dict(r, section="three", k={"low": []}, dH=lim["dH"]) for r in lim["rows"]]
    troe_line = [f"Troe {lim['troe']}"] if troe else []
    net, _ = build(rows, tmp_path, lambda c: [c["line"], *troe_line])
    cells = [(row, T, P) for row in rows for T, P in CONDITIONS]
    x0 = np.zeros((len(cells), len(net.species)))
    lam = np.zeros(len(cells))
    for b, (row, T, P) in enumerate(cells):
        x0[b, net.species.index(row["from"])], x0[b, net.species.index("He")] = 0.1, 0.9
        if row["k0"] is not None:
            nd = float(helpers.number_density(T, P))
            f = broadening_at_low_pressure(lim["troe"], T) if troe else 1.0
            lam[b] = modified_arrhenius(row["k0"], T) * f * nd
    dt = np.where(lam > 0.0, 1.0 / np.where(lam > 0.0, lam, 1.0), 1.0)
    run = helpers.backend(kind, net)
    out = run(x0, [T for _, T, _ in cells], [P for _, _, P in cells], dt, 1e-3 * dt, rtol=RTOL)
    run.close()
    assert out["done"].all()
    for b, (row, _, _) in enumerate(cells):
        i = net.species.index(row["from"])
        # at constant P the isomer's loss is made up by its product: x = x0 exp(-lambda t)
        assert out["x"][b, i] == pytest.approx(0.1 * math.exp(-lam[b] * dt[b]), rel=RATE_TOLERANCE), row["name"]


# ---- the collision cap on the reference's collision network, against closed forms ----

K_FIRST, J_COLL, T_COLL = 1e9, 1e9, 3000.0
P_LOW, P_HIGH = 1e3, 1e7


def capped_pairs(net, P):
    """(kf, kr) of each capped reaction of tests/golden/data/synthetic_collision_network.txt at T_COLL and P: the
    larger direction of each, in bimolecular form, held at K_COLL and the other following by K."""
    nd = P / (helpers.BOLTZMANN * T_COLL)
    lkt = math.log(helpers.BOLTZMANN * T_COLL / helpers.STANDARD_PRESSURE)

    def g(q):  # G/RT of a constant-Cp species of the network at T_COLL
        cp, h, b2 = net.lo[net.species.index(q)][[2, 7, 8]]
        return cp + h / T_COLL - (cp * math.log(T_COLL) + b2)

    dissociation = math.exp(2 * g("H") - g("H2") + lkt)  # kr / kf of H2 + M -> H + H + M
    association = math.exp(g("O2") - 2 * g("O") - lkt)  # kr / kf of O + O + M -> O2 + M
    capped = K_COLL * nd
    if P == P_LOW:  # the forward direction is the larger one
        h2, o2 = (capped, capped * dissociation), (capped / association, capped)
    else:
        h2, o2 = (K_COLL / dissociation, K_COLL), (K_COLL, K_COLL * association)
    return {"H2 + M -> H + H + M": h2, "O + O + M -> O2 + M": o2, "N_1 + M -> N_2 + M": (capped, capped),
            "N_3 -> N_4": (K_FIRST, K_FIRST), "H2O -> H + OH": (J_COLL, 0.0)}  # fmt: skip


@pytest.mark.parametrize("P", [P_LOW, P_HIGH])
def test_capped_coefficients_match_closed_forms(read_network, P):
    import oracle

    from maxichem_amd import _tables

    net = read_network("synthetic_collision")
    orc = oracle.Oracle(_tables.build(net, photo=True))
    kf, kr = orc.rates(T_COLL, P, np.full(len(net.photo), J_COLL))
    equations = [r["eq"] for r in net.reac] + [r["eq"] for r in net.photo]
    pairs = capped_pairs(net, P)
    assert set(pairs) <= set(equations)
    for eq, (want_f, want_r) in pairs.items():
        q = equations.index(eq)
        assert kf[q] == pytest.approx(want_f, rel=1e-13, abs=0.0), eq
        assert kr[q] == pytest.approx(want_r, rel=1e-13, abs=0.0), eq


# the tangents of the capped coefficients (the tangent kernel's rates_tangent) against the closed forms' derivatives:
# d ln k / dT at fixed P and d ln k / d ln P, from the same closed forms as above (a capped forward direction K_COLL nd
# has -1/T and 1; the other follows by the NASA-9 ratio's d/dT; a direction capped at K_COLL has 0)
DERIVATIVE_TOLERANCE = 1e-11  # of 1/T along T, of 1 along ln P: exp and log rounding


def capped_log_derivatives(net, P):
    """{equation: (d ln kf/dT, d ln kr/dT, d ln kf/d ln P, d ln kr/d ln P)} at T_COLL and P, fixed P along T."""
    t = T_COLL

    def dg(q):  # d(G/RT)/dT = -(H/RT)/T of a constant-Cp species: H/RT = cp + h / T
        cp, h = net.lo[net.species.index(q)][[2, 7]]
        return -(cp + h / t) / t

    d_diss = 2 * dg("H") - dg("H2") + 1 / t  # d ln(kr / kf)/dT of H2 + M -> H + H + M
    d_assoc = dg("O2") - 2 * dg("O") - 1 / t  # of O + O + M -> O2 + M
    if P == P_LOW:  # the forward directions capped at K_COLL nd: -1/T along T, 1 along ln P
        h2, o2 = (-1 / t, d_diss - 1 / t, 1.0, 1.0), (-1 / t - d_assoc, -1 / t, 1.0, 1.0)
    else:  # the reverse of H2 + M and the forward of O + O + M capped at K_COLL
        h2, o2 = (-d_diss, 0.0, 0.0, 0.0), (0.0, d_assoc, 0.0, 0.0)
    return {"H2 + M -> H + H + M": h2, "O + O + M -> O2 + M": o2, "N_1 + M -> N_2 + M": (-1 / t, -1 / t, 1.0, 1.0),
            "N_3 -> N_4": (0.0, 0.0, 0.0, 0.0)}  # fmt: skip


@pytest.mark.gpu
def test_capped_coefficients_tangents_match_closed_forms(read_network, gpu):
    import torch

    from maxichem_amd import _hip, _kernels, _tables

    net = read_network("synthetic_collision")
    tab = _tables.build(net, photo=True)
    probe = (Path(__file__).resolve().parent / "rate_tangent_probe.hip").read_text()
    src = _kernels.source(tab, 1, 1e-6, extra=("jvp.hip",)) + "\n" + probe
    module = _hip.Module(_hip.compile_source(src, "rate_tangent_probe.hip"), torch.cuda.current_device())
    try:
        fn = module.function("mxc_rate_tangent_probe")
        host = _kernels.host_tables(tab)
        tables = [torch.from_numpy(host[k]).cuda() for k in _kernels.NET_FIELDS]
        netp = torch.tensor([a.data_ptr() for a in tables], dtype=torch.int64, device="cuda")
        # four cells: (P_LOW, P_HIGH) x (along T at fixed P, along ln P)
        cells = [(P, along) for P in (P_LOW, P_HIGH) for along in (0, 1)]
        T = torch.full((4,), T_COLL, dtype=torch.float64, device="cuda")
        P = torch.tensor([c[0] for c in cells], dtype=torch.float64, device="cuda")
        nd = P / (helpers.BOLTZMANN * T_COLL)
        Td = torch.tensor([1.0 if a == 0 else 0.0 for _, a in cells], dtype=torch.float64, device="cuda")
        ndd = torch.where(Td > 0, -nd / T_COLL, nd)  # d nd along T at fixed P, or along ln P
        J = torch.full((4, tab.nj), J_COLL, dtype=torch.float64, device="cuda")
        out = [torch.full((4, tab.nr), float("nan"), dtype=torch.float64, device="cuda") for _ in range(4)]
        p = _hip.ptr
        lds = 8 * (tab.layout()["SIZE"] + _kernels._jvp_region(tab))
        _hip.launch(fn, 4, 64, lds, torch.cuda.current_stream().cuda_stream,
                    [p(netp), p(T), p(P), p(Td), p(ndd), p(J), *[p(o) for o in out]])  # fmt: skip
        torch.cuda.synchronize()
        kf, kr, dkf, dkr = (o.cpu().numpy() for o in out)
    finally:
        module.unload()
    equations = [r["eq"] for r in net.reac] + [r["eq"] for r in net.photo]
    worst = 0.0
    for i, (Pc, along) in enumerate(cells):
        scale = 1.0 / T_COLL if along == 0 else 1.0
        for eq, (want_f, want_r) in capped_pairs(net, Pc).items():
            q = equations.index(eq)
            assert kf[i, q] == pytest.approx(want_f, rel=1e-13, abs=0.0), eq
            assert kr[i, q] == pytest.approx(want_r, rel=1e-13, abs=0.0), eq
        for eq, d in capped_log_derivatives(net, Pc).items():
            q = equations.index(eq)
            for value, want in ((dkf[i, q] / kf[i, q], d[2 * along]), (dkr[i, q] / kr[i, q], d[2 * along + 1])):
                worst = max(worst, abs(value - want) / scale)
                assert abs(value - want) <= DERIVATIVE_TOLERANCE * scale, (eq, Pc, along, value, want)
    print(f"capped coefficients' tangents against the closed forms: largest error {worst:.2e} (of 1/T or 1)")
