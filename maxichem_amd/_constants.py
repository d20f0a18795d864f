"""Every setting of the chemistry step, with its unit and reason.

The values are those of the reference solver's settings module (maxichem/_constants.py), so a caller that switches
solvers keeps its defaults; the kernels read them through the generated network header (_header.py), so each value is
written once, here.
"""

# ---- What a call accepts (maxichem/_constants.py:12-25) ----
T_MIN = 300.0  # K: colder cells are refused (counted "cold", not stepped)
T_MAX = 3000.0  # K: hotter cells are refused (counted "hot"); the NASA-9 fits are trusted to here
X_SUM_TOL = 1e-10  # |sum(x) - 1| above this is bad input: the state is molecules per starting molecule
DRIFT = 1e-6  # default element-share drift the guard leaves to rounding, relative
REC_STEPS = 10  # the recovery's tries, in units of the call's max_steps
DEFAULT_RTOL = 1e-3
DEFAULT_MAX_STEPS = 4000  # tries per call before a cell counts as unfinished
ATOL = 1e-25  # absolute tolerance, mole fraction
MAX_GIBBS_JUMP = 1.0e-5  # largest |G/RT(low) - G/RT(high)| the reader accepts at NASA_SPLIT_T

# ---- The method (maxichem/_constants.py:40-72; docs/method.md "Step control", "Guard and recovery") ----
H_MIN = 1e-12  # s: the smallest first try
CALL_END = 1.0 - 1e-14  # a call ends at this share of dt
STEP_SAFETY = 0.9  # safety factor on err^(-1/3)
ERR_FLOOR = 1e-10  # err floor inside the step factor
FAC_NONFINITE = 0.1  # step factor after a non-finite err
FACACC = 0.2  # smallest step factor after an accepted try
FACREJ = 0.2  # smallest step factor after a rejected try
FACMAX = 20.0  # largest step factor
FAC_AFTER_REJECT = 0.5  # largest step factor after a rejected try
SECMIN = 1e-4  # smallest secant shrink
SECANT_MIN_ORDER = 0.05  # below this fitted order the secant shrink is SECMIN
SECANT_MAX_ORDER = 3.0  # Rodas3's error order: the largest order the secant fit may use (the kernels take cbrt)
SECANT_TARGET = 0.5  # the err a secant shrink aims at
PIVTAU = 0.1  # recovery: rows exchange only where |W_kk| < PIVTAU * column maximum
ELEMENT_PIVOT_MIN = 1e-10  # guard: an element whose pivot falls below this share of its start is left out
REC_RETRIES = 3  # recovery: passes after the first for a cell that only drifted
REC_RETRY_SPLIT = 0.5  # recovery: retry k starts with at most dt * REC_RETRY_SPLIT^k
GMAX = 1e-3  # guard: the share correction changes a species by at most this, relative

# ---- Physical constants (maxichem/_constants.py:75-78) ----
KB = 1.380649e-16  # erg K^-1
P0 = 1.0e6  # dyn cm^-2: the NASA-9 standard pressure (1 bar)
NASA_SPLIT_T = 1000.0  # K: the low NASA-9 range applies below this
K_COLL = 1e-9  # cm^3 s^-1: no thermal direction exceeds this as a bimolecular-equivalent rate

# ---- Rate-law codes (maxichem/_constants.py:92-98): a reaction dict's "kind", as VULCAN numbers the sections ----
KIND_TWO_BODY = 2
KIND_FALLOFF = 3
KIND_LOW_PRESSURE = 5
KIND_SPECIAL = 6
FALLOFF_KINDS = (KIND_FALLOFF, KIND_SPECIAL)
THREE_BODY_KINDS = (KIND_FALLOFF, KIND_LOW_PRESSURE, KIND_SPECIAL)

# Flag bits of a reaction's law word (the kernels and the oracle read the same word).
FL_FALLOFF = 1
FL_SECOND_K0 = 2
FL_THREE_BODY = 4
FL_TROE = 8
FL_EFF = 16
FL_VM11 = 32
FL_PHOTO = 64  # a photolysis branch: kf = J, kr = 0
FL_INERT = 128  # a reaction whose net stoichiometry is zero (A + B -> A + B): it moves nothing
FL_NF_SHIFT = 8  # two bits: molecules on the forward side
FL_NR_SHIFT = 10  # two bits: molecules on the reverse side
RATE_ROW = 14  # ln A0, B0, C0, ln A1, B1, C1, ln Ainf, Binf, Cinf, Troe a, T3, T1, T2, efficiency

# ---- Counts (maxichem/_constants.py:100-113) ----
COUNTS = (
    "bad_input",
    "cold",
    "hot",
    "nonfinite",
    "unfinished",
    "drift",
    "recomputed",
    "recovered",
    "still_failed",
    "nonfinite_h",
)
COUNT_SLOTS = ("queue", *COUNTS, "recovery_queue")  # the device's count array: the two queues around COUNTS

# Per-cell status written by the call kernel's screen.
REFUSED_COLD = 1
REFUSED_BAD_INPUT = 2
REFUSED_HOT = 3
# Bits of the guard's verdict on a stepped cell.
GUARD_NONFINITE = 4
GUARD_UNFINISHED = 8
GUARD_DRIFT = 16

WAVE = 64  # lanes of a CDNA4 wavefront: one wavefront steps one cell
