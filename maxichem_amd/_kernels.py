"""The HIP kernels' source for one network: the generated header (sizes, LDS layout, stream rounds, settings) and the
kernel files, compiled at run time with hiprtc for gfx950.

The reference generates network.h from its tables the same way (maxichem/_kernel_header.py:164-232) for NVRTC; the
macros here are this build's own layout.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from . import _constants as C
from ._tables import (
    CHUNK,
    SV_DBL,
    SV_END,
    SV_NEG,
    SV_PIECE,
    SV_PIECE_BITS,
    SV_R_BITS,
    T_DBL,
    T_DEST,
    T_DEST_BITS,
    T_END,
    T_IDX_BITS,
    T_NEG,
    Tables,
)

KERNEL_DIR = Path(__file__).resolve().parent / "kernels"
KERNEL_FILES = ("common.hip", "guard.hip", "call.hip", "recover.hip")
ARCH = "gfx950"
# hiprtc options: gfx950 code; packed FP32 instructions are kept out of all code (a known hazard on these parts)
OPTIONS = (
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=fast",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
)
REC_WAVES = 4  # wavefronts of a recovery workgroup: the first steps the cell, all of them share its dense LU
KERNEL_NAMES = ("mxc_call", "mxc_recover")
# Each kernel's parameters in order (kernels/*.hip); an export lists them and the C++ loader refuses another list, as
# the reference's KERNEL_ARGUMENTS (maxichem/_constants.py:148-164) does for its kernels.
KERNEL_ARGUMENTS = {
    "mxc_call": ("net", "x0", "Ta", "Pa", "dta", "hio", "xout", "doneo", "noko", "status", "redo", "list", "count",
                 "tries", "order", "Jin", "B", "rtol", "atol", "max_steps", "xprop", "rscratch", "fp64"),
    "mxc_recover": ("net", "x0", "Ta", "Pa", "dta", "hin", "xout", "hio", "doneo", "noko", "redo", "list", "count",
                    "tries", "Jin", "rtol", "atol", "max_steps", "rscratch"),
}  # fmt: skip
# The kernels' Net struct (kernels/common.hip): one device pointer per table, in this order.
NET_FIELDS = ("perm", "rc", "rf", "nasa", "rxe", "svw", "rowp", "diagw", "leadw", "chw", "atoms", "rord")

# The method's settings the kernels read as constexprs of the same names.
_SETTINGS = (
    "H_MIN", "CALL_END", "STEP_SAFETY", "ERR_FLOOR", "FAC_NONFINITE", "FACACC", "FACREJ", "FACMAX",
    "FAC_AFTER_REJECT", "SECMIN", "SECANT_MIN_ORDER", "SECANT_MAX_ORDER", "SECANT_TARGET", "PIVTAU",
    "ELEMENT_PIVOT_MIN", "REC_RETRY_SPLIT", "GMAX", "NASA_SPLIT_T",
)  # fmt: skip
_FLAGS = ("FL_FALLOFF", "FL_SECOND_K0", "FL_THREE_BODY", "FL_TROE", "FL_EFF", "FL_VM11", "FL_PHOTO", "FL_INERT",
          "FL_NF_SHIFT", "FL_NR_SHIFT")  # fmt: skip


DEFAULT_BATCH = 16


def words_count(tab: Tables) -> int:
    """The stream words a cell reads every try (S v, then the Jacobian's diagonal, lead and chunk streams)."""
    rounds = tab.sv_words.shape[0] + tab.diag_words.shape[0] + tab.lead_words.shape[0]
    return C.WAVE * (rounds + sum(w.shape[0] for w in tab.chunk_words))


CU_LDS = 160 * 1024  # LDS per compute unit (gfx950)
SIMDS = 4  # SIMDs per compute unit


def words_in_lds(tab: Tables, cpw: int, mixed: bool = False, wps: int = 1, rates_global: bool = False) -> bool:
    """Whether the call kernel keeps the stream words in its workgroup's LDS (read with the LDS's latency instead of
    the caches'; measured 2.6 % faster on NCHO): the plain form only, when the words fit beside the workgroup's cells
    without lowering the cells a compute unit holds."""
    if mixed:
        return False
    cells = 8 * tab.layout(mixed, not rates_global)["SIZE"] * cpw
    words = cells + 4 * words_count(tab)

    def per_cu(nbytes: int) -> int:
        return min(SIMDS * wps, CU_LDS // nbytes * cpw)

    return words <= CU_LDS and per_cu(words) >= per_cu(cells)


def header(tab: Tables, cpw: int, drift: float, batch: int = DEFAULT_BATCH, mixed: bool = False, wps: int = 1,
           rates_global: bool = False, words_lds: bool = False) -> str:
    """The network header: every size, offset and setting the kernel files read. cpw: cells (wavefronts) per
    workgroup of the call kernel; batch: the rounds of a lane stream whose words are loaded together; mixed: the
    mixed-precision form (FP32 trailing factors, FP64 refinement); wps: the call kernel's register budget in
    wavefronts per SIMD (2: at most 256 registers, the excess spilled); rates_global: the rate coefficients in a
    global scratch slot per wavefront instead of LDS; words_lds: the call kernel reads the stream words from a copy
    in its workgroup's LDS, after the cells' regions."""
    lay = tab.layout(mixed, not rates_global)
    nch = len(tab.chunk_words)
    rounds = [w.shape[0] for w in tab.chunk_words]
    offsets = [sum(rounds[:c]) for c in range(nch)]
    ints = {
        "WAVE": C.WAVE, "NS": tab.ns, "NR": tab.nr, "NRT": tab.nrt, "NJ": tab.nj, "K0": tab.k0, "NT": tab.nt,
        "NEL": tab.nel, "RCW": C.RATE_ROW, "RROUNDS": -(-tab.nr // C.WAVE), "NDV": tab.ndv, "NPART": tab.npart,
        "CPW": cpw, "STREAM_BATCH": batch, "REC_WAVES": REC_WAVES, "CHUNK": CHUNK,
        "NCH": nch, "NLCH": tab.nlch,
        "SV_ROUNDS": tab.sv_words.shape[0], "DIAG_ROUNDS": tab.diag_words.shape[0],
        "LEAD_ROUNDS": tab.lead_words.shape[0], "REC_RETRIES": C.REC_RETRIES,
        "T_NEG": T_NEG, "T_DBL": T_DBL, "T_END": T_END, "T_DEST": T_DEST, "SV_NEG": SV_NEG, "SV_DBL": SV_DBL,
        "SV_END": SV_END, "SV_PIECE": SV_PIECE,
        **{f"L_{k}": v for k, v in lay.items()},
        **{f"COUNT_{name.upper()}": i for i, name in enumerate(C.COUNT_SLOTS)},
        **{name: getattr(C, name) for name in _FLAGS},
    }  # fmt: skip
    doubles = {
        **{name: getattr(C, name) for name in _SETTINGS},
        "TMIN": C.T_MIN, "TMAX": C.T_MAX, "XSUMTOL": C.X_SUM_TOL, "DRIFT": float(drift),
        "ETOT": float(tab.atoms.sum()), "KCOLL": C.K_COLL, "KBOLTZ": C.KB, "P0BAR": C.P0,
    }  # fmt: skip
    # the tangent kernel's region per cell after the call's (kernels/jvp.hip): rate tangents [2 NR] (in the
    # wavefront's global slot after its rates instead, with the rates in global scratch), then the hyper-dual n
    # [4][NS + 1] (G / RT and its tangent there while the rates are computed), v [4][NR + 1], pieces [4][NPART]
    jn = 0 if rates_global else 2 * tab.nr
    jv = jn + 4 * (tab.ns + 1)
    jp = jv + 4 * (tab.nr + 1)
    jsize = jp + 4 * tab.npart
    ints.update({"J_KD": 0, "J_N": jn, "J_G": jn, "J_V": jv, "J_PART": jp, "J_SIZE": jsize + (jsize & 1)})
    lines = [f"// network header: {tab.ns} species, {tab.nr} reactions, tables {tab.digest()}"]
    lines += [f"constexpr int {k} = {v};" for k, v in ints.items()]
    lines += [f"constexpr double {k} = {v!r};" for k, v in doubles.items()]
    lines += [
        f"constexpr unsigned T_IDX_MASK = {(1 << T_IDX_BITS) - 1}u;",
        f"constexpr unsigned SV_R_MASK = {(1 << SV_R_BITS) - 1}u;",
        f"constexpr unsigned T_DEST_MASK = {(1 << T_DEST_BITS) - 1}u;",
        f"constexpr unsigned SV_PIECE_MASK = {(1 << SV_PIECE_BITS) - 1}u;",
        f"__device__ constexpr int CH_ROUNDS[{nch}] = {{{', '.join(map(str, rounds))}}};",
        f"__device__ constexpr int CH_OFF[{nch}] = {{{', '.join(map(str, offsets))}}};",
        # the trailing rows' filled pattern: bit j of TRAIL_FILL[t] is column K0 + j of row K0 + t
        f"__device__ constexpr unsigned long long TRAIL_FILL[{tab.nt}] = "
        f"{{{', '.join(f'{m}ull' for m in tab.trail_fill)}}};",
        f"#define HAS_TROE {int(tab.has_troe)}",
        f"#define MIXED {int(mixed)}",
        f"#define CALL_WAVES_PER_SIMD {int(wps)}",
        f"#define RATES_GLOBAL {int(rates_global)}",
        f"#define WORDS_LDS {int(words_lds)}",
        f"constexpr int WORDS_TOTAL = {words_count(tab)};",
        f"#define HAS_EFF {int(tab.has_eff)}",
        "#define DNAN __longlong_as_double(0x7ff8000000000000ll)",
    ]
    return "\n".join(lines) + "\n"


def _jvp_region(tab: Tables, rates_global: bool = False) -> int:
    jsize = (0 if rates_global else 2 * tab.nr) + 4 * (tab.ns + 1) + 4 * (tab.nr + 1) + 4 * tab.npart
    return jsize + (jsize & 1)


def jvp_lds_bytes(tab: Tables, cpw: int, rates_global: bool = False) -> int:
    """Dynamic LDS of a tangent workgroup (kernels/jvp.hip): each cell's call region and its tangent region (the
    rates and their tangents in global scratch when rates_global)."""
    return 8 * (tab.layout(False, not rates_global)["SIZE"] + _jvp_region(tab, rates_global)) * cpw


def jvp_recover_lds_bytes(tab: Tables, rates_global: bool = False) -> int:
    """Dynamic LDS of a tangent recovery workgroup (mxc_jvp_recover): the recovery's, with the tangent region after
    the cell's call region."""
    return lds_bytes(tab, 1, rates_global=rates_global)[1] + 8 * _jvp_region(tab, rates_global)


# doubles of a tangent wavefront's (or tangent recovery workgroup's) global slot with the rates in global scratch:
# the rates kf, kr, then their tangents
def jvp_slot(tab: Tables) -> int:
    return 4 * tab.nr


def lds_bytes(tab: Tables, cpw: int, mixed: bool = False, rates_global: bool = False, wps: int = 1,
              extra: tuple[str, ...] = ()) -> tuple[int, int]:
    """Dynamic LDS of a call workgroup (cpw cells, then the stream words where source() puts them there) and of a
    recovery workgroup (one cell, its dense W, the reciprocal pivots, a solve's vector, its factors' row order, and
    the word by which its first wavefront calls the others to a factorization)."""
    size = tab.layout(mixed, not rates_global)["SIZE"]
    words = 4 * words_count(tab) if not extra and words_in_lds(tab, cpw, mixed, wps, rates_global) else 0
    return 8 * size * cpw + words, 8 * (size + tab.ns * tab.ns + 2 * tab.ns) + 4 * tab.ns + 4


def host_tables(tab: Tables) -> dict[str, np.ndarray]:
    """The tables the kernels read through their Net struct, as contiguous host arrays (species in elimination
    order where the kernels index them so)."""
    t = {
        "perm": tab.perm.astype(np.int32),
        "rc": tab.rc.astype(np.float64),
        "rf": tab.rf.astype(np.int32),
        "nasa": tab.nasa[:, tab.perm, :].astype(np.float64),
        "rxe": tab.rxe.view(np.int64),
        "svw": tab.sv_words.ravel().view(np.int32),
        "rowp": tab.row_pieces.view(np.int32),
        "diagw": tab.diag_words.ravel().view(np.int32),
        "leadw": tab.lead_words.ravel().view(np.int32),
        "chw": np.concatenate([w.ravel() for w in tab.chunk_words]).view(np.int32),
        "atoms": tab.atoms.astype(np.float64),
        "rord": tab.rord.astype(np.int32),
    }
    return {k: np.ascontiguousarray(v) for k, v in t.items()}


MLP_FILE = "mlp.hip"


def mlp_source() -> str:
    """The learned warm start's kernels (kernels/mlp.hip): not specialised to a network."""
    return (KERNEL_DIR / MLP_FILE).read_text()


def source(tab: Tables, cpw: int, drift: float, batch: int = DEFAULT_BATCH, mixed: bool = False, wps: int = 1,
           rates_global: bool = False, extra: tuple[str, ...] = ()) -> str:
    """The whole translation unit hiprtc compiles: the header, then the kernel files in order (mixed.hip after
    common.hip in the mixed-precision form), then the extra kernel files (the tangent's jvp.hip, whose module keeps
    the stream words in global memory)."""
    words = not extra and words_in_lds(tab, cpw, mixed, wps, rates_global)
    parts = [header(tab, cpw, drift, batch, mixed, wps, rates_global, words)]
    for name in KERNEL_FILES[:1] + (("mixed.hip",) if mixed else ()) + KERNEL_FILES[1:] + tuple(extra):
        parts.append(f"// ---- {name} ----\n" + (KERNEL_DIR / name).read_text())
    return "\n".join(parts)
