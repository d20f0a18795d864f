"""The GCM side of a call (the reference's tests/test_coupling_composition.py and test_coupling_chemistry.py):

- without a GPU: the oracle pinned to the reference's hand-worked mixture and element totals and to exact fractions;
  molar masses, the start gas's element ratios, the set-up refusals, require_finished's messages;
- on the GPU: the fused kernels on the same hand cases and refusals, against the oracle at 65,536 cells in float64
  and float32, the totals' bits on repeated runs, empty and ragged shapes, and a GCM step around a solver call
  (mass to mole, the call, the stop on unfinished cells, mole to mass) keeping each cell's mass and elements.
"""

from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle.coupling as ref
from maxichem_amd._constants import COUNTS
from maxichem_amd.coupling.chemistry import PA_TO_DYN_CM2, UnfinishedCellsError, require_finished
from maxichem_amd.coupling.composition import (
    GAS_CONSTANT,
    START_PER_H,
    element_moles,
    mass_to_mole,
    molar_masses,
    mole_to_mass,
    signed_mole_fractions,
    start_mole_fractions,
    tracer_totals,
)

# atoms of each species over H, He, C, N, O, S
ATOMS = {
    "H2": (2, 0, 0, 0, 0, 0), "He": (0, 1, 0, 0, 0, 0), "CO": (0, 0, 1, 0, 1, 0), "H2O": (2, 0, 0, 0, 1, 0),
    "O2": (0, 0, 0, 0, 2, 0), "CH4": (4, 0, 1, 0, 0, 0), "N2": (0, 0, 0, 2, 0, 0), "NH3": (3, 0, 0, 1, 0, 0),
    "H2S": (2, 0, 0, 0, 0, 1), "SO2": (0, 0, 0, 0, 2, 1),
}  # fmt: skip
ELEMENTS = ("H", "He", "C", "N", "O", "S")


def table(species):
    """The atoms table [S, E] of these species over the elements they hold, and those elements."""
    full = np.array([ATOMS[s] for s in species], dtype=np.float64)
    keep = [j for j in range(len(ELEMENTS)) if full[:, j].any()]
    return full[:, keep], [ELEMENTS[j] for j in keep]


# the reference's hand-worked cases (tests/test_coupling_composition.py:36-99): H2, He, H2O, O2
HAND_MW = np.array([2.016e-3, 4.002602e-3, 18.015e-3, 31.998e-3])
HAND_Q = np.array([[0.5179290046119381, 0.19280726094845174, 0.2892637344396102, 0.0]])
HAND_X = np.array([[0.8, 0.15, 0.05, 0.0]])
TOTALS_Q = np.array([[0.5, 0.3, 0.2, 0.0], [0.0, 0.0, 0.0, 1.0]])
TOTALS_CELL_MASS = np.array([2.0, 3.0])
TOTALS_HAND = np.array([5.0, 1036.47093, 149.902488, 209.715439])  # kg, then mol H, He, O
PER_KG_HAND = np.array([[518.235465, 74.9512442, 11.1018596], [0.0, 0.0, 62.5039065]])


def test_molar_masses_and_gas_constant():
    atoms, elements = table(["H2", "He", "H2O", "O2"])
    assert elements == ["H", "He", "O"]
    np.testing.assert_allclose(molar_masses(atoms, elements).numpy(), HAND_MW, rtol=1e-15)
    assert float(Fraction("1.380649e-23") * Fraction("6.02214076e23")) == GAS_CONSTANT


def test_oracle_pinned_to_the_hand_cases():
    atoms, _ = table(["H2", "He", "H2O", "O2"])
    np.testing.assert_allclose(ref.mass_to_mole(HAND_Q, HAND_MW), HAND_X, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ref.tracer_totals(TOTALS_Q, TOTALS_CELL_MASS, HAND_MW, atoms), TOTALS_HAND, rtol=1e-8)
    np.testing.assert_allclose(ref.element_moles(TOTALS_Q, HAND_MW, atoms), PER_KG_HAND, rtol=1e-8)
    q = ref.mole_to_mass(HAND_X, HAND_MW, np.array([1.0]))
    np.testing.assert_allclose(q, HAND_Q, rtol=1e-15, atol=0)


def test_oracle_against_exact_fractions():
    rng = np.random.default_rng(4)
    q = rng.uniform(0.0, 1.0, (3, 5))
    mw = rng.uniform(1e-3, 5e-2, 5)
    mass = rng.uniform(0.5, 2.0, 3)
    x = ref.mass_to_mole(q, mw)
    back = ref.mole_to_mass(x, mw, mass)
    for c in range(3):
        n = [Fraction(a) / Fraction(b) for a, b in zip(q[c], mw)]
        exact = [v / sum(n) for v in n]
        np.testing.assert_allclose(x[c], [float(v) for v in exact], rtol=4e-16)
        xm = [Fraction(a) * Fraction(b) for a, b in zip(x[c], mw)]
        exact_q = [v * Fraction(mass[c]) / sum(xm) for v in xm]
        np.testing.assert_allclose(back[c], [float(v) for v in exact_q], rtol=4e-16)


# the solar rows (dex, H = 12) and the chosen He/H, and the relative rounding of a five-digit ratio
SOLAR = {"O": 10 ** (8.7826 - 12), "C": 10 ** (8.4434 - 12), "N": 10 ** (7.9130 - 12), "S": 10 ** (7.1492 - 12),
         "He": 0.09692}  # fmt: skip
FIVE_DIGITS = 4e-5


@pytest.mark.parametrize("species", [["H2", "He", "CO", "H2O", "CH4", "O2"], ["H2", "He", "CO", "H2O", "N2", "NH3", "CH4"],
                                     ["H2", "He", "CO", "H2O", "N2", "NH3", "H2S", "SO2"]])  # fmt: skip
def test_start_gas_holds_the_element_ratios(species):
    atoms, elements = table(species)
    x = start_mole_fractions(species, elements)
    assert x.dtype == torch.float64 and x.device.type == "cpu"
    assert float(x.sum()) == pytest.approx(1.0, abs=1e-15)
    per_atom = x.numpy() @ atoms
    h = per_atom[elements.index("H")]
    for k, e in enumerate(elements):
        if e != "H":
            assert per_atom[k] / h == pytest.approx(START_PER_H[e], rel=1e-14)
            assert per_atom[k] / h == pytest.approx(SOLAR[e], rel=FIVE_DIGITS)


def test_set_up_refusals():
    species = ["H2", "He", "CO", "H2O", "NH3"]  # nitrogen, but no N2 to start it in
    with pytest.raises(ValueError, match=r"N2 \(for N\)"):
        start_mole_fractions(species, table(species)[1])
    with pytest.raises(ValueError, match=r"elements \['Ti'\] with no start carrier"):
        start_mole_fractions(["H2", "He", "CO", "H2O", "TiO"], ["H", "He", "C", "O", "Ti"])
    with pytest.raises(ValueError, match=r"no atomic mass for \['Ti'\]"):
        molar_masses(torch.ones((2, 2)), ["H", "Ti"])
    with pytest.raises(ValueError, match="atoms must be finite and not negative"):
        molar_masses(torch.tensor([[2.0, float("inf")]]), ["H", "He"])
    with pytest.raises(ValueError, match=r"atoms must be \[species, 2\]"):
        molar_masses(torch.ones((2, 3)), ["H", "He"])
    with pytest.raises(ValueError, match=r"species rows \[1\] have no atoms"):
        molar_masses(torch.tensor([[2.0, 0.0], [0.0, 0.0]]), ["H", "He"])


def test_conversions_run_on_the_gpu_only():
    mw = torch.tensor(HAND_MW)
    with pytest.raises(ValueError, match="must be on one GPU"):
        mass_to_mole(torch.tensor(HAND_Q), mw)


# ---- require_finished (the reference's tests/test_coupling_chemistry.py) ----

T3 = torch.tensor([1500.0, 280.0, 1200.0], dtype=torch.float64)
P3 = torch.tensor([1.0e5, 3.0e2, 2.0e6], dtype=torch.float64)
CAUSES = {
    "cold": ("below the solver's 300-3000 K domain (refused at 300 K)", "look at the deck"),
    "hot": ("above the solver's 300-3000 K domain (refused at 3000 K)", "look at the deck"),
    "bad_input": ("bad input", "inputs to the call are wrong"),
    "still_failed": ("still failing the guard after the recovery", "a shorter --dt"),
}
NOT_CAUSES = {"nonfinite", "unfinished", "drift", "recomputed", "recovered", "nonfinite_h"}


def counts(**rose):
    return {k: rose.get(k, 0) for k in COUNTS}


def test_every_count_is_judged():
    assert set(COUNTS) == set(CAUSES) | NOT_CAUSES and not set(CAUSES) & NOT_CAUSES


@pytest.mark.parametrize("count", COUNTS)
def test_stop_names_the_counted_cause_and_only_its_fix(count):
    with pytest.raises(UnfinishedCellsError) as err:
        require_finished(torch.tensor([True, False, True]), T3, P3, counts(**{count: 1}), "at cycle 4")
    text = str(err.value)
    assert "1 chemistry cell(s) not done at cycle 4; the first, cell 1: T 280.0 K, P 300 Pa" in text
    assert f"The call's counts: { {count: 1} }" in text
    if count not in CAUSES:
        assert "no count of the solver explains it" in text
        return
    cause, fix = CAUSES[count]
    assert cause in text and fix in text
    for other, (other_cause, other_fix) in CAUSES.items():
        if other != count:
            assert other_cause not in text and (other_fix == fix or other_fix not in text)


def test_still_failed_names_what_the_recovery_could_not_mend():
    c = counts(unfinished=1, nonfinite=2, recomputed=3, still_failed=3)
    with pytest.raises(UnfinishedCellsError, match=r"in the call: 1 out of steps, 2 non-finite, 0 element drift"):
        require_finished(torch.tensor([True, False, True]), T3, P3, c, "at cycle 4")


def test_done_cells_and_shapes():
    require_finished(torch.ones(3, dtype=torch.bool), T3, P3, counts(cold=1), "at cycle 5")
    with pytest.raises(UnfinishedCellsError, match="no count of the solver explains it"):
        require_finished(torch.tensor([False, True, True]), T3, P3, counts(), "at cycle 1")
    with pytest.raises(UnfinishedCellsError, match="2 chemistry cell"):
        require_finished(torch.tensor([False, True, False]), T3, P3, counts(), "at cycle 1")
    with pytest.raises(ValueError, match="must all be"):
        require_finished(torch.ones(2, dtype=torch.bool), T3, P3, counts(), "at cycle 1")


# ---- on the GPU ----


def g(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dtype).cuda()


@pytest.mark.gpu
def test_hand_mixture_and_elements_kept_through_a_reaction(gpu):
    atoms_np, elements = table(["H2", "He", "H2O", "O2"])
    mw, atoms = g(molar_masses(atoms_np, elements)), g(atoms_np)
    x = mass_to_mole(g(HAND_Q), mw)
    np.testing.assert_allclose(x.cpu().numpy(), HAND_X, rtol=0, atol=1e-15)
    gen = torch.Generator().manual_seed(3)
    q = (torch.rand((5, 4), generator=gen, dtype=torch.float64) * torch.tensor([1.0, 0.3, 0.1, 0.2])).cuda()
    mass = q.sum(1)
    x = mass_to_mole(q, mw)
    assert torch.allclose(mole_to_mass(x, mw, mass), q, rtol=1e-14, atol=0)
    extent = 0.5 * torch.minimum(x[:, 0] / 2, x[:, 3])  # 2 H2 + O2 -> 2 H2O between the calls
    n = x.clone()
    n[:, 0] -= 2 * extent
    n[:, 3] -= extent
    n[:, 2] += 2 * extent
    q_out = mole_to_mass(n / n.sum(1, keepdim=True), mw, mass)
    assert torch.allclose(q_out.sum(1), mass, rtol=1e-14, atol=0)
    assert torch.allclose(element_moles(q_out, mw, atoms), element_moles(q, mw, atoms), rtol=1e-13, atol=0)
    cm = torch.full((5,), 2.0, dtype=torch.float64, device="cuda")
    totals = tracer_totals(q_out, cm, mw, atoms)
    assert torch.allclose(totals, tracer_totals(q, cm, mw, atoms), rtol=1e-13, atol=0)
    assert float(totals[0]) == pytest.approx(2.0 * float(mass.sum()), rel=1e-14)


@pytest.mark.gpu
def test_element_totals_by_hand(gpu):
    atoms_np, elements = table(["H2", "He", "H2O", "O2"])
    mw, atoms = g(molar_masses(atoms_np, elements)), g(atoms_np)
    q = g(TOTALS_Q)
    np.testing.assert_allclose(tracer_totals(q, g(TOTALS_CELL_MASS), mw, atoms).cpu().numpy(), TOTALS_HAND, rtol=1e-8)
    np.testing.assert_allclose(element_moles(q, mw, atoms).cpu().numpy(), PER_KG_HAND, rtol=1e-8)


@pytest.mark.gpu
def test_conversions_refuse_unusable_input(gpu):
    mw = g([2.016e-3, 4.002602e-3])
    one = g([1.0])
    cases = [
        ("negatives", lambda: mass_to_mole(g([[0.9, -1e-9]]), mw)),
        ("non-finite", lambda: mass_to_mole(g([[0.9, float("nan")]]), mw)),
        ("cell 1 holds no tracer mass", lambda: mass_to_mole(g([[0.9, 0.1], [0.0, 0.0]]), mw)),
        ("non-finite", lambda: mole_to_mass(g([[float("inf"), 0.1]]), mw, one)),
        ("cell 0 has mole fractions of no mass", lambda: mole_to_mass(g(np.zeros((1, 2))), mw, one)),
        (r"mass must be \[1\]", lambda: mole_to_mass(g(np.ones((1, 2))), mw, g([1.0, 1.0]))),
        (r"q must be \[cells, 2\]", lambda: mass_to_mole(g(np.ones((1, 3))), mw)),
        (r"atoms must be \[2, elements\]", lambda: element_moles(g(np.ones((1, 2))), mw, g(np.ones((3, 1))))),
        ("atoms holds non-finite", lambda: element_moles(g(np.ones((1, 2))), mw, g(np.full((2, 1), np.nan)))),
        (r"cell_mass must be \[1\]", lambda: tracer_totals(g(np.ones((1, 2))), g([1.0, 1.0]), mw, g(np.ones((2, 1))))),
        ("cell_mass must be finite and not negative", lambda: tracer_totals(g(np.ones((1, 2))), g([-1.0]), mw,
                                                                            g(np.ones((2, 1))))),
        ("share one precision", lambda: element_moles(g(np.ones((1, 2))), mw, g(np.ones((2, 1)), torch.float32))),
        ("q and mw must share one precision", lambda: mass_to_mole(g(np.ones((1, 2)), torch.float32), mw)),
        ("q and mw must share one precision", lambda: signed_mole_fractions(g(np.ones((1, 2)), torch.float32), mw)),
        ("x, mw and mass must share one precision", lambda: mole_to_mass(g(np.ones((1, 2)), torch.float32), mw, one)),
        ("x, mw and mass must share one precision", lambda: mole_to_mass(g(np.ones((1, 2))), mw, one.float())),
        (r"mw must be \[S\], got shape \(2, 1\)", lambda: mole_to_mass(g([[0.5, 0.5], [0.9, 0.1]]), mw[:, None],
                                                                     g([1.0, 1.0]))),
        ("mw must be finite and positive", lambda: mass_to_mole(g([[0.5, 0.5]]), g([2.016e-3, 0.0]))),
        (r"mass must be \[2\], got shape \(2, 1\)", lambda: mole_to_mass(g([[0.5, 0.5], [0.9, 0.1]]), mw,
                                                                        g(np.ones((2, 1))))),
        ("mass must be finite and not negative", lambda: mole_to_mass(g([[0.5, 0.5], [0.9, 0.1]]), mw, g([1.0, -0.5]))),
        ("x holds negatives", lambda: mole_to_mass(g([[1.1, -0.1]]), mw, one)),
    ]  # fmt: skip
    for message, call in cases:
        with pytest.raises(ValueError, match=message):
            call()
    # the output form keeps the transport's negatives: n = q / mw, x = n / sum(n)
    signed = signed_mole_fractions(g([[1.008e-3, -4.002602e-4]]), mw)
    np.testing.assert_allclose(signed.cpu().numpy(), [[0.5 / 0.4, -0.1 / 0.4]], rtol=1e-14, atol=0)


def random_state(cells, species, seed):
    """Mass mixing ratios spanning 30 decades, per-cell tracer masses near 1, molar masses and atoms of a network."""
    from maxichem_amd import standin

    net = standin.load(species)
    rng = np.random.default_rng(seed)
    q = 10.0 ** rng.uniform(-30, 0, (cells, len(net.species)))
    q *= rng.uniform(0.5, 1.5, cells)[:, None] / q.sum(1, keepdims=True)
    mw = molar_masses(net.atoms, net.elements).numpy()
    return q, mw, np.asarray(net.atoms, np.float64), rng.uniform(1e3, 1e6, cells)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_kernels_match_the_oracle_at_full_size(gpu, dtype):
    q, mw, atoms, cm = random_state(65536, "NCHO", 1)
    tol = 1e-14 if dtype == torch.float64 else 3e-6
    qd, mwd, ad, cmd = g(q, dtype), g(mw, dtype), g(atoms, dtype), g(cm, dtype)
    qc, mwc = qd.double().cpu().numpy(), mwd.double().cpu().numpy()  # the rounded inputs, for float32
    x = mass_to_mole(qd, mwd)
    want = ref.mass_to_mole(qc, mwc)
    np.testing.assert_allclose(x.double().cpu().numpy(), want, rtol=tol, atol=1e-300)
    mass = g(q.sum(1), dtype)
    back = mole_to_mass(x, mwd, mass)
    np.testing.assert_allclose(back.double().cpu().numpy(), ref.mole_to_mass(x.double().cpu().numpy(), mwc,
                               mass.double().cpu().numpy()), rtol=tol, atol=1e-300)  # fmt: skip
    np.testing.assert_allclose(back.double().sum(1).cpu().numpy(), mass.double().cpu().numpy(), rtol=4 * tol)
    e = element_moles(qd, mwd, ad)
    np.testing.assert_allclose(e.double().cpu().numpy(), ref.element_moles(qc, mwc, ad.double().cpu().numpy()),
                               rtol=10 * tol, atol=0)  # fmt: skip
    t1 = tracer_totals(qd, cmd, mwd, ad)
    t2 = tracer_totals(qd, cmd, mwd, ad)
    assert torch.equal(t1, t2)  # a fixed summation order: the same bits
    np.testing.assert_allclose(t1.double().cpu().numpy(), ref.tracer_totals(qc, cmd.double().cpu().numpy(), mwc,
                               ad.double().cpu().numpy()), rtol=100 * tol)  # fmt: skip


@pytest.mark.gpu
@pytest.mark.parametrize("cells,species", [(0, 3), (1, 1), (7, 64), (9, 65), (13, 256)])
def test_empty_and_ragged_shapes(gpu, cells, species):
    rng = np.random.default_rng(cells + species)
    q = rng.uniform(0.1, 1.0, (cells, species))
    mw = rng.uniform(1e-3, 5e-2, species)
    atoms = rng.integers(0, 3, (species, 3)).astype(np.float64)
    atoms[:, 0] += 1.0
    x = mass_to_mole(g(q), g(mw))
    assert x.shape == (cells, species)
    np.testing.assert_allclose(x.cpu().numpy(), ref.mass_to_mole(q, mw) if cells else q, rtol=1e-14)
    np.testing.assert_allclose(mole_to_mass(x, g(mw), g(q.sum(1))).cpu().numpy(), q, rtol=1e-13)
    np.testing.assert_allclose(element_moles(g(q), g(mw), g(atoms)).cpu().numpy(), ref.element_moles(q, mw, atoms),
                               rtol=1e-13)  # fmt: skip
    cm = rng.uniform(1.0, 2.0, cells)
    want = ref.tracer_totals(q, cm, mw, atoms) if cells else np.zeros(4)
    np.testing.assert_allclose(tracer_totals(g(q), g(cm), g(mw), g(atoms)).cpu().numpy(), want, rtol=1e-13)


@pytest.mark.gpu
def test_a_gcm_step_keeps_each_cells_mass_and_elements(gpu):
    """Tracers to mole fractions, one chemistry call (pressure in Pa converted at the edge), the stop on unfinished
    cells, back to tracers with each cell's mass: the mass to rounding, the elements to the guard's drift bound."""
    from maxichem_amd import standin, workloads
    from maxichem_amd.kincuda import CudaSolver

    net = standin.load("NCHO")
    cells = 4096
    d = workloads.config1(net.species, len(net.photo), cells=cells, seed=9)
    mw = g(molar_masses(net.atoms, net.elements))
    atoms = g(net.atoms)
    x0 = g(d["x"])
    mass = g(np.random.default_rng(9).uniform(0.5, 1.5, cells))
    q0 = mole_to_mass(x0, mw, mass)
    with CudaSolver(net, "cuda", photo=True) as solver:
        x = mass_to_mole(q0, mw)
        T, P_pa = g(d["T"]), g(d["P"]) / PA_TO_DYN_CM2
        r = solver(x, T, P_pa * PA_TO_DYN_CM2, g(d["dt"]), g(d["h"]), J=g(d["J"]))
        require_finished(r.done, T, P_pa, solver.fallback, "at cycle 1")
        q1 = mole_to_mass(r.x, mw, mass)
    np.testing.assert_allclose(q1.sum(1).cpu().numpy(), mass.cpu().numpy(), rtol=1e-14)
    e0, e1 = (element_moles(q, mw, atoms) / mass[:, None] for q in (q0, q1))
    # the call keeps element shares to the guard's drift allowance; per kg the shares' scale follows the mean
    # molar mass, so compare shares
    s0, s1 = (e / e.sum(1, keepdim=True) for e in (e0, e1))
    assert float(((s1 - s0).abs() / s0.clamp_min(1e-300)).max()) < 1e-5
