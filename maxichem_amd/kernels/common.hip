// Shared device arithmetic of the chemistry step: one 64-lane wavefront steps one cell.
//
// A cell's state vectors are held two values per lane: the lead part (elimination row `lane`, valid for lane < K0)
// and the trailing part (row K0 + lane, valid for lane < NT). The trailing block of M0 = I / (gamma h) - Jn sits in
// registers, lane t owning row K0 + t in c[NT], and the same row's lead columns (A21) in a21[K0]; the lead rows sit
// dense in the cell's LDS region (LEAD [K0][NS]). The network header (generated per network) defines the sizes, the
// LDS layout (L_*), the stream rounds and the method's settings.

typedef unsigned int u32;
typedef unsigned long long u64;

// Rodas3 (Sandu et al. 1997): gamma = 1/2 and the stage couplings C21 = 4, C43 = -8/3.
constexpr double GAMMA = 0.5;
constexpr double RODAS_C21 = 4.0;
constexpr double RODAS_C43 = 8.0 / 3.0;
// The Troe factor's published coefficients (Venot et al. 2012, eq. C.2-C.5).
constexpr double TROE_C0 = 0.4;
constexpr double TROE_C1 = 0.67;
constexpr double TROE_N0 = 0.75;
constexpr double TROE_N1 = 1.27;
constexpr double TROE_D = 0.14;
constexpr int REFUSED_COLD = 1;
constexpr int REFUSED_BAD_INPUT = 2;
constexpr int REFUSED_HOT = 3;
constexpr int BROADCAST_WINDOW = 32;  // v_readlane broadcasts of a pivot row in flight ahead of their updates
constexpr int SCHUR_BATCH = 16;  // a lead row's entries read ahead of their Schur updates
constexpr int SMALL_LEAD = 8;  // at most this many lead rows: their factorization and solves are unrolled
typedef double LeadRow[K0 > 0 ? K0 : 1];  // a trailing row's lead columns, one row per lane in registers

// Orders the LDS accesses of the wavefront's lanes: the wave executes in lockstep and LDS serves one wave's accesses
// in issue order, so only the compiler must not move loads and stores across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane l's value of v, for a wave-uniform l (v_readlane, no LDS).
__device__ __forceinline__ double readlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The LLVM intrinsic behind v_writelane_b32 (clang has no builtin for it): old with lane `lane` set to src, a
// wave-uniform value.
__device__ int writelane_i32(int src, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// v with lane l's value replaced by x, a wave-uniform value (v_writelane, both 32-bit halves).
__device__ __forceinline__ double writelane(double v, double x, int l) {
  const long long b = __double_as_longlong(v), xb = __double_as_longlong(x);
  const int lo = writelane_i32((int)(xb & 0xffffffffll), l, (int)(b & 0xffffffffll));
  const int hi = writelane_i32((int)(xb >> 32), l, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// v from another lane by a DPP pattern (both 32-bit halves), all lanes active.
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Reductions over the 64 lanes in a fixed order, every lane ending with the same bits: within each row of 16 lanes
// by DPP (pairs, quads, half-rows, rows: each step adds two values in either order, which rounds alike), then the
// four row results in one wave-uniform sum.
constexpr int DPP_XOR1 = 0xb1, DPP_XOR2 = 0x4e, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp<DPP_XOR1>(v);
  v += dpp<DPP_XOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  v += dpp<DPP_MIRROR>(v);
  return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

// N sums at once, each as wave_sum (the same bits): every DPP step is taken for all N values before the next, so
// the N reductions overlap instead of waiting out each other's latency.
template <int N>
__device__ __forceinline__ void wave_sums(double (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) v[i] += dpp<DPP_XOR1>(v[i]);
#pragma unroll
  for (int i = 0; i < N; i++) v[i] += dpp<DPP_XOR2>(v[i]);
#pragma unroll
  for (int i = 0; i < N; i++) v[i] += dpp<DPP_HALF_MIRROR>(v[i]);
#pragma unroll
  for (int i = 0; i < N; i++) v[i] += dpp<DPP_MIRROR>(v[i]);
#pragma unroll
  for (int i = 0; i < N; i++) v[i] = (readlane(v[i], 0) + readlane(v[i], 16)) + (readlane(v[i], 32) + readlane(v[i], 48));
}

__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp<DPP_XOR1>(v));
  v = fmax(v, dpp<DPP_XOR2>(v));
  v = fmax(v, dpp<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp<DPP_MIRROR>(v));
  return fmax(fmax(readlane(v, 0), readlane(v, 16)), fmax(readlane(v, 32), readlane(v, 48)));
}
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(0, v, DPP_XOR1, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(0, v, DPP_XOR2, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(0, v, DPP_HALF_MIRROR, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(0, v, DPP_MIRROR, 0xf, 0xf, false));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// The lane index as an opaque value, taken inside each step of an unrolled loop so that the step's lane masks are
// computed there and not all ahead of the loop, where they would fill the scalar registers.
__device__ __forceinline__ int opaque_lane(int lane) {
  asm volatile("" : "+v"(lane));
  return lane;
}

// The lane index as an opaque value that the compiler must take after x is computed: a step's lane mask then
// cannot be computed ahead of the step whose value x it follows.
__device__ __forceinline__ int lane_after(int lane, double x) {
  asm volatile("" : "+v"(lane) : "v"(x));
  return lane;
}

// The table pointer p, made opaque to the compiler where it is used, so that the loads and decodes from the
// per-network tables are not hoisted out of the try loop and held in registers across it.
template <typename P>
__device__ __forceinline__ P opaque(P p) {
  asm volatile("" : "+s"(p));
  return p;
}

// a and b materialised here, in this order with the other volatile statements (an empty asm that reads and writes
// them): it keeps a step's values from being computed early or late, which would stretch their live ranges.
__device__ __forceinline__ void pin(double& a, double& b) { asm volatile("" : "+v"(a), "+v"(b)); }
__device__ __forceinline__ void pin(double& a) { asm volatile("" : "+v"(a)); }
// a's next use waits for b (no instruction; an ordering the compiler keeps)
__device__ __forceinline__ void tie(double& a, double b) { asm volatile("" : "+v"(a) : "v"(b)); }
__device__ __forceinline__ void tie(float& a, float b) { asm volatile("" : "+v"(a) : "v"(b)); }

__device__ __forceinline__ bool wave_all(bool p) { return __ballot(!p) == 0ull; }

// A stream word's coefficient: +-1 or +-2 from its sign and x2 bits.
__device__ __forceinline__ double coefficient(u32 w, int neg, int dbl) {
  const double c = ((w >> dbl) & 1u) ? 2.0 : 1.0;
  return ((w >> neg) & 1u) ? -c : c;
}

// A stream word's term: d times the word's coefficient, exactly as fma(coefficient, d, .) takes it: x2 is an exponent
// step, and the sign (the word's bit 31) is XORed into the double's sign bit. Four instructions, where the
// coefficient's compares and selects take eight and wait states for their condition masks.
static_assert(T_NEG == 31 && SV_NEG == 31, "the streams' sign bit is the word's sign bit");
template <int DBL>
__device__ __forceinline__ double term(u32 w, double d) {
  d = __builtin_amdgcn_ldexp(d, (int)((w >> DBL) & 1u));
  return __longlong_as_double(__double_as_longlong(d) ^ ((long long)(w & 0x80000000u) << 32));
}

// A table pointer in the global address space. Loads through it are global loads, counted by the vector-memory
// counter alone, which the compiler may issue ahead of LDS accesses and wait for only where their values are used.
// Through a plain (flat) pointer, which may address LDS, each load was ordered with the wave's LDS traffic and waited
// for at once: a round of reaction words cost a memory latency.
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;

// The network's tables in device memory, one pointer for every kernel (a function reads the fields it needs through
// opaque(net), so they are loaded where used rather than held in registers across the call).
struct Net {
  gptr<int> perm;      // [NS] network species of elimination row k
  gptr<double> rc;     // [NR][RCW] rate-law rows
  gptr<int> rf;        // [NR] law flags
  gptr<double> nasa;   // [2][NS][9] NASA-9 ranges, elimination order
  gptr<u64> rxe;       // [NR] six species (elimination order, pad NS) and the reaction's first dv slot
  gptr<u32> svw;       // S v stream [SV_ROUNDS][WAVE]
  gptr<u32> rowp;      // [NS] each row's first S v piece | count << 16
  gptr<u32> diagw;     // -Jn streams: diagonal, lead, trailing chunks
  gptr<u32> leadw;
  gptr<u32> chw;
  gptr<double> atoms;  // [NS][NEL] network order
  gptr<int> rord;      // [NR] the order the rate coefficients are computed in (reactions grouped by law)
};
// The Net struct in the constant address space: its fields come by scalar loads, uniform across the wave.
using NetP = const __attribute__((address_space(4))) Net*;

// ---- rate coefficients (the reference's law, collision_limit, gibbs_rt and rate_pair: common.cuh:135-242) ----

// a0, a1, a2: q[0..2], the first Arrhenius term's, which the caller may have loaded ahead
__device__ __forceinline__ double law(gptr<double> __restrict__ q, int fl, double T, double lT, double iT, double nd,
                                      double a0, double a1, double a2) {
  double k = exp(a0 + a1 * lT - a2 * iT);
  if (fl & FL_SECOND_K0) k += exp(q[3] + q[4] * lT - q[5] * iT);
  if (fl & FL_FALLOFF) {
    const double pr = k * nd / exp(q[6] + q[7] * lT - q[8] * iT);
    k = k > 0.0 ? k / (1.0 + pr) : 0.0;  // k0 = 0: no rate, also at Pr = 0 / 0; Pr = inf: the high-pressure 0
#if HAS_TROE
    if ((fl & FL_TROE) && k > 0.0) {
      const double fc = log10((1.0 - q[9]) * exp(-T / q[10]) + q[9] * exp(-T / q[11]) + exp(-q[12] / T));
      const double x = (fl & FL_VM11) ? log10(pr) : log10(pr) - TROE_C0 - TROE_C1 * fc;
      // Pr = 0: Troe's ratio tends to -1 / d, the VM11 one to -inf (F = 1)
      const double y = (fl & FL_VM11) ? x / (TROE_N0 - TROE_N1 * fc)
                                      : (isfinite(x) ? x / (TROE_N0 - TROE_N1 * fc - TROE_D * x) : -1.0 / TROE_D);
      k *= exp10(fc / (1.0 + y * y));
    }
#endif
  }
#if HAS_EFF
  if (fl & FL_EFF) k *= q[13];
#endif
  return k;
}
__device__ __forceinline__ double law(gptr<double> __restrict__ q, int fl, double T, double lT, double iT, double nd) {
  return law(q, fl, T, lT, iT, nd, q[0], q[1], q[2]);
}

// A direction's bimolecular-equivalent rate: M-folded k times nd^(molecules - 2); one molecule without M: none.
__device__ __forceinline__ double bimolecular(int m, bool three_body, double k, double nd) {
  return m >= 2 ? k * (m == 3 ? nd : 1.0) : (three_body ? k / nd : 0.0);
}

// G / RT of one species from its NASA-9 range a at T.
__device__ __forceinline__ double gibbs_rt(gptr<double> __restrict__ a, double T, double lT, double iT) {
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double h = -a[0] * iT * iT + a[1] * lT * iT + a[2] + a[3] * T / 2 + a[4] * T2 / 3 + a[5] * T3 / 4 +
                   a[6] * T4 / 5 + a[7] * iT;
  const double s = -a[0] * iT * iT / 2 - a[1] * iT + a[2] * lT + a[3] * T + a[4] * T2 / 2 + a[5] * T3 / 3 +
                   a[6] * T4 / 4 + a[8];
  return h - s;
}

__device__ __forceinline__ int species_of(u64 w, int m) { return (int)((w >> (8 * m)) & 0xffull); }

// The rate coefficients of every reaction at the cell's T and nd into rk (kf [NR] then kr [NR]; LDS, or the wave's
// slot of global scratch), kept there for the call.
// G / RT (elimination order) goes to the scratch at L_UNION first.
// lmc: the cell's log rate multipliers [NR] or null; each reaction's two coefficients are scaled by exp(lm) after the
// collision limit, a photolysis branch's kf = J exp(lm) (the reference's tangent kernels' rates, ad_common.cuh:136-168)
__device__ __forceinline__ void cell_rates(double* ws, double* rk, double T, double nd, NetP net,
                                           const double* Jc, int lane, const double* lmc = nullptr) {
  net = opaque(net);
  const gptr<double> rc = net->rc;
  const gptr<int> rf = net->rf;
  const gptr<u64> rxe = net->rxe;
  double* g = ws + L_UNION;
  double* kf = rk;
  double* kr = rk + NR;
  const double lT = log(T), iT = 1.0 / T, lKT = log(KBOLTZ * T / P0BAR);
  const gptr<double> cf = net->nasa + (T < NASA_SPLIT_T ? 0 : NS * 9);
  for (int i = lane; i < NS; i += WAVE) g[i] = gibbs_rt(cf + 9 * i, T, lT, iT);
  if (lane == 0) g[NS] = 0.0;
  wave_sync();
  const gptr<int> rord = net->rord;
  // the rounds' table words are loaded a round ahead (the reaction's index two rounds ahead), so that the chain of
  // dependent loads rord -> rf, rxe, rc is not waited on in each round; a lane past NR reads reaction rord[NR - 1]
  auto at = [](int q) { return q < NR ? q : NR - 1; };
  int rn = rord[at(lane)], rnn = rord[at(lane + WAVE)];
  int fn = rf[rn];
  u64 wn = rxe[rn];
  double an0 = rc[RCW * rn], an1 = rc[RCW * rn + 1], an2 = rc[RCW * rn + 2];
#pragma unroll 1
  for (int q = lane; q < NR; q += WAVE) {
    const int r = rn, fl = fn;  // a round's reactions mostly of one law, so it runs the code of few laws
    const u64 w = wn;
    const double a0 = an0, a1 = an1, a2 = an2;
    rn = rnn;
    rnn = rord[at(q + 2 * WAVE)];
    fn = rf[rn];
    wn = rxe[rn];
    an0 = rc[RCW * rn];
    an1 = rc[RCW * rn + 1];
    an2 = rc[RCW * rn + 2];
    double f = 0.0, b = 0.0;
    if (fl & FL_PHOTO) {
      f = Jc[r - NRT];
    } else {
      const double k = law(rc + RCW * r, fl, T, lT, iT, nd, a0, a1, a2);
      const double dg = g[species_of(w, 3)] + g[species_of(w, 4)] + g[species_of(w, 5)] - g[species_of(w, 0)] -
                        g[species_of(w, 1)] - g[species_of(w, 2)];
      const int mf = (fl >> FL_NF_SHIFT) & 3, mr = (fl >> FL_NR_SHIFT) & 3;
      const bool tb = (fl & FL_THREE_BODY) != 0;
      const double c3 = tb ? nd : 1.0;
      f = k * c3;
      b = (k * exp(dg + (double)(mr - mf) * lKT)) * c3;
      const double qb = fmax(bimolecular(mf, tb, f, nd), bimolecular(mr, tb, b, nd));
      if (qb > KCOLL) {  // the collision limit: both directions scaled, their ratio kept
        const double c = KCOLL / qb;
        f *= c;
        b *= c;
      }
    }
    if (lmc) {
      const double e = exp(lmc[r]);
      f *= e;
      b *= e;
    }
    kf[r] = f;
    kr[r] = b;
  }
  wave_sync();
}

// ---- right-hand side and Jacobian ----

// A term stream (the Jacobian's and S v's): each lane sums whole entries from src terms and writes each at its
// destination in dst. The words of STREAM_BATCH rounds are loaded together, ahead of their terms, so that their
// latency is paid once per batch: whole batches in a loop, then the last rounds as a batch of their own (the rounds
// are known when the kernel is built, so no padding terms are taken). Every word stores the running sum, at its
// entry's destination when it ends one and else in the lane's trash slot, and the sum restarts where an entry ended:
// no branch and no exec-mask change per word. (Loading each batch, or a stream's first batch, one step ahead was
// measured slower: it holds 36 more registers.)
template <unsigned IDX_MASK, int DBL, int END, int DEST, unsigned DEST_MASK, typename WP>
__device__ __forceinline__ void stream_sum(WP __restrict__ words, int rounds, const double* src, double* dst,
                                           double* trash, int lane) {
  double acc = 0.0;
  auto batch = [&](int q0, int n) {  // rounds q0 .. q0 + n - 1, n <= STREAM_BATCH
    u32 w[STREAM_BATCH];
    double d[STREAM_BATCH];
#pragma unroll
    for (int i = 0; i < STREAM_BATCH; i++)
      if (i < n) w[i] = words[(q0 + i) * WAVE + lane];
#pragma unroll
    for (int i = 0; i < STREAM_BATCH; i++)
      if (i < n) d[i] = src[w[i] & IDX_MASK];
#pragma unroll
    for (int i = 0; i < STREAM_BATCH; i++) {
      if (i >= n) break;
      acc += term<DBL>(w[i], d[i]);
      const bool end = (w[i] >> END) & 1u;
      *(end ? dst + ((w[i] >> DEST) & DEST_MASK) : trash) = acc;
      acc = end ? 0.0 : acc;
    }
  };
  const int full = rounds / STREAM_BATCH * STREAM_BATCH;
#pragma unroll 1
  for (int q0 = 0; q0 < full; q0 += STREAM_BATCH) batch(q0, STREAM_BATCH);
  if (rounds > full) batch(full, rounds - full);
}

// One Jacobian stream into the cell's LDS: entries of -Jn at their offsets.
template <typename WP>
__device__ __forceinline__ void jac_stream(WP __restrict__ words, int rounds, const double* DV, double* ws, int lane) {
  stream_sum<T_IDX_MASK, T_DBL, T_END, T_DEST, T_DEST_MASK>(words, rounds, DV, ws, ws + L_TRASH + lane, lane);
}

// The stream words in the workgroup's LDS (WORDS_LDS: the call kernel copies them there at its start, after its
// cells' regions), in the order S v, the Jacobian's diagonal, lead and chunk streams: read with the LDS's latency
// instead of the caches'.
constexpr int W_SV = 0, W_DIAG = SV_ROUNDS * WAVE, W_LEAD = W_DIAG + DIAG_ROUNDS * WAVE;
constexpr int W_CH = W_LEAD + LEAD_ROUNDS * WAVE;
__device__ __forceinline__ const u32* lds_words(const double* ws) {  // ws: this wavefront's cell region
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  return (const u32*)(ws + (CPW - wv) * L_SIZE);
}

// F = (n_tot / nd) S v(n) at the state Y (lead / trailing parts), u0 = S v / nd; with JAC also the partial
// derivatives dv/dn of every live reaction into DV. Returns n_tot = sum Y.
template <bool JAC, bool LW = false>
__device__ __forceinline__ double rhs(double Yl, double Yt, double& Fl, double& Ft, double& u0l, double& u0t,
                                      NetP net, double nd, double* ws, const double* rk, int lane) {
  const double ntot = wave_sum((lane < K0 ? Yl : 0.0) + (lane < NT ? Yt : 0.0));
  const double xf = nd / ntot, ixf = ntot / nd;
  double* n = ws + L_NVEC;
  if (lane < K0) n[lane] = Yl * xf;
  if (lane < NT) n[K0 + lane] = Yt * xf;
  if (lane == 0) n[NS] = 1.0;  // the padding species of a shorter side
  wave_sync();
  net = opaque(net);
  const gptr<u64> rxe = net->rxe;
  const gptr<int> rf = net->rf;
  const gptr<u32> svw = net->svw;
  const gptr<u32> rowp = net->rowp;
  const double* kf = rk;
  const double* kr = rk + NR;
  double* V = ws + L_V;
  double* DV = ws + L_DV;
  // every round's reaction word (and law flags) loaded first: the loads are in bounds for every lane (the last round's
  // spare lanes read reaction NR - 1), so none waits on the branch that ends a lane's rounds
  u64 wr[RROUNDS];
  int fr[JAC ? RROUNDS : 1];
  const u32 rpl = rowp[lane < K0 ? lane : 0], rpt = rowp[K0 + (lane < NT ? lane : 0)];  // each row's S v pieces
#pragma unroll
  for (int q = 0; q < RROUNDS; q++) {
    const int r = q * WAVE + lane < NR ? q * WAVE + lane : NR - 1;
    wr[q] = rxe[r];
    if (JAC) fr[q] = rf[r];
  }
#pragma unroll
  for (int q = 0; q < RROUNDS; q++) {
    const int r = q * WAVE + lane;
    if (r >= NR) break;
    const u64 w = wr[q];
    const double n0 = n[species_of(w, 0)], n1 = n[species_of(w, 1)], n2 = n[species_of(w, 2)];
    const double n3 = n[species_of(w, 3)], n4 = n[species_of(w, 4)], n5 = n[species_of(w, 5)];
    const double f = kf[r], b = kr[r];
    V[r] = f * n0 * n1 * n2 - b * n3 * n4 * n5;
    if (JAC) {
      const int fl = fr[JAC ? q : 0];
      if (!(fl & FL_INERT)) {
        int o = (int)(w >> 48);
        const int mf = (fl >> FL_NF_SHIFT) & 3, mr = (fl >> FL_NR_SHIFT) & 3;
        if (mf >= 1) DV[o] = f * n1 * n2;
        if (mf >= 2) DV[o + 1] = f * n0 * n2;
        if (mf >= 3) DV[o + 2] = f * n0 * n1;
        if (!(fl & FL_PHOTO)) {
          o += mf;
          if (mr >= 1) DV[o] = b * n4 * n5;
          if (mr >= 2) DV[o + 1] = b * n3 * n5;
          if (mr >= 3) DV[o + 2] = b * n3 * n4;
        }
      }
    }
  }
  if (lane == 0) V[NR] = 0.0;  // the S v stream's padding term
  wave_sync();
  double* part = ws + L_PART;
  if constexpr (LW)
    stream_sum<SV_R_MASK, SV_DBL, SV_END, SV_PIECE, SV_PIECE_MASK>(lds_words(ws) + W_SV, SV_ROUNDS, V, part,
                                                                  ws + L_TRASH + lane, lane);
  else
    stream_sum<SV_R_MASK, SV_DBL, SV_END, SV_PIECE, SV_PIECE_MASK>(svw, SV_ROUNDS, V, part, ws + L_TRASH + lane,
                                                                  lane);
  wave_sync();
  if (lane < K0) {
    double s = 0.0;
    for (int e = 0; e < (int)(rpl >> 16); e++) s += part[(rpl & 0xffffu) + e];
    Fl = s * ixf;
    u0l = s / nd;
  }
  if (lane < NT) {
    const u32 rp = rpt;
    double s = 0.0;
    for (int e = 0; e < (int)(rp >> 16); e++) s += part[(rp & 0xffffu) + e];
    Ft = s * ixf;
    u0t = s / nd;
  }
  wave_sync();
  return ntot;
}

// -Jn into LEAD (LDS), the trailing rows' lead columns a21 and the trailing block c (registers), from the dv terms
// that rhs<true> left in DV; then u = w = S v / nd - Jn N / n_tot, the constant-pressure part of the exact Jacobian
// (lead / trailing parts).
template <bool LW = false, typename T>
__device__ __forceinline__ void jacobian(T (&c)[NT], LeadRow& a21, double Nl, double Nt, double u0l, double u0t,
                                         double ntot, double& ul, double& ut, NetP net, double* ws, int lane) {
  double* DV = ws + L_DV;
  for (int s = lane; s < L_NVEC; s += WAVE) ws[s] = 0.0;  // LEAD
  if (lane < NT) ws[L_DIAG + lane] = 0.0;                   // a species in no reaction has no diagonal term
  if (lane == 0) DV[NDV] = 0.0;                             // the padding words' term
  // the state N in the exchange vector (n's place: free once the dv terms are in DV)
  double* xv = ws + L_NVEC;
  if (lane < K0) xv[lane] = Nl;
  if (lane < NT) xv[K0 + lane] = Nt;
  wave_sync();
  if constexpr (LW) {
    jac_stream(lds_words(ws) + W_DIAG, DIAG_ROUNDS, DV, ws, lane);
    jac_stream(lds_words(ws) + W_LEAD, LEAD_ROUNDS, DV, ws, lane);
  } else {
    jac_stream(opaque(net)->diagw, DIAG_ROUNDS, DV, ws, lane);
    jac_stream(opaque(net)->leadw, LEAD_ROUNDS, DV, ws, lane);
  }
  wave_sync();
  // -Jn N in FP64, from the entries before the trailing block takes them in T: the lead rows through the exchange
  // vector, then the trailing rows chunk by chunk (their lead columns, then the trailing ones), N_j from lane j
  double ml = 0.0, mt = 0.0;
  if constexpr (K0 > 0 && K0 <= SMALL_LEAD) {
    // few lead rows: each row times N as this lane's share (its own N entries), summed over the lanes; the trailing
    // rows' lead columns times N with N_k by v_readlane
    double sum[K0 > 0 ? K0 : 1];
#pragma unroll
    for (int i = 0; i < K0; i++)
      sum[i] = (lane < K0 ? ws[i * NS + lane] * Nl : 0.0) + (lane < NT ? ws[i * NS + K0 + lane] * Nt : 0.0);
    wave_sums(sum);
#pragma unroll
    for (int i = 0; i < K0; i++) ml = lane == i ? sum[i] : ml;
  } else if (K0 > 0) {
#pragma unroll 1
    for (int j0 = 0; j0 < NS; j0 += STREAM_BATCH) {  // the lead rows, a batch of columns at a time
      double a[STREAM_BATCH], nv[STREAM_BATCH];
#pragma unroll
      for (int i = 0; i < STREAM_BATCH; i++) {
        const int j = j0 + i;
        a[i] = lane < K0 && j < NS ? ws[lane * NS + j] : 0.0;
        nv[i] = j < NS ? xv[j] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < STREAM_BATCH; i++) ml = fma(a[i], nv[i], ml);
    }
  }
  const double cd = lane < NT ? ws[L_DIAG + lane] : 0.0;
  double* stage = ws + L_STAGE;
  // chunks 0 .. NLCH - 1 hold the lead columns (K0 > 0 means NT == WAVE: every lane has a trailing row), the
  // others the trailing columns from t0
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int t0 = (ch - NLCH) * CHUNK;
    if (lane < NT) {
#pragma unroll
      for (int j = 0; j < CHUNK; j++) stage[j * NT + lane] = 0.0;
      // the chunk streams hold no diagonal entry: a trailing row's own diagonal goes in by its lane
      if (ch >= NLCH && lane / CHUNK == ch - NLCH) stage[(lane % CHUNK) * NT + lane] = cd;
    }
    wave_sync();
    if constexpr (LW)
      jac_stream(lds_words(ws) + W_CH + CH_OFF[ch] * WAVE, CH_ROUNDS[ch], DV, ws, lane);
    else
      jac_stream(opaque(net)->chw + CH_OFF[ch] * WAVE, CH_ROUNDS[ch], DV, ws, lane);
    wave_sync();
    // N_j for -Jn N as wave-uniform values: with few lead rows from the exchange vector (same-address LDS reads,
    // kept in their chunk: hoisted ahead, they would hold a register pair each), else by v_readlane (with the lead
    // columns in registers the LDS reads' registers make the kernel spill)
    double nj[CHUNK];
    if constexpr (K0 <= SMALL_LEAD) {
#pragma unroll
      for (int j = 0; j < CHUNK; j++)
        nj[j] = ch < NLCH ? xv[ch * CHUNK + j < K0 ? ch * CHUNK + j : 0] : xv[K0 + (t0 + j < NT ? t0 + j : 0)];
      asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int j = 0; j < CHUNK; j++) {
      if (ch < NLCH) {
        if (ch * CHUNK + j < K0) {
          const double v = stage[j * NT + lane];
          a21[ch * CHUNK + j] = v;
          mt = fma(v, K0 <= SMALL_LEAD ? nj[j] : readlane(Nl, ch * CHUNK + j), mt);
        }
      } else if (t0 + j < NT) {
        const double v = lane < NT ? stage[j * NT + lane] : 0.0;
        c[t0 + j] = (T)v;
        mt = fma(v, K0 <= SMALL_LEAD ? nj[j] : readlane(Nt, t0 + j), mt);
      }
    }
  }
  ul = u0l + ml / ntot;
  ut = lane < NT ? u0t + mt / ntot : 0.0;
  wave_sync();
}

// ---- the unpivoted LU of M0 in the elimination order, and its solves ----

// Adds 1 / (gamma h) to M0's diagonal: the lead rows in LDS, the trailing block in registers.
template <typename T>
__device__ __forceinline__ void add_diagonal(T (&c)[NT], double igh, double* ws, int lane) {
  if (lane < K0) ws[lane * NS + lane] += igh;
#pragma unroll
  for (int j = 0; j < NT; j++) c[j] += j == lane ? (T)igh : (T)0;
  wave_sync();
}

// One lead row's Schur update of the trailing block: c_j -= a U_kj, a = this lane's L entry in column k. The row's
// trailing entries come one per lane from a single LDS read and reach every lane by v_readlane (a same-address LDS
// read per entry would wait out the LDS latency once per pair of columns), at most BROADCAST_WINDOW ahead of their
// updates.
template <typename T>
__device__ __forceinline__ void schur_row(T (&c)[NT], double a, const double* ws, int k, int lane) {
  // the row's trailing entries as wave-uniform values from same-address LDS reads (one instruction per pair of
  // entries, where a v_readlane broadcast takes two per entry), SCHUR_BATCH of them issued ahead of their updates
  const double* u = ws + k * NS + K0;
#pragma unroll
  for (int j0 = 0; j0 < NT; j0 += SCHUR_BATCH) {
    double b[SCHUR_BATCH];
#pragma unroll
    for (int i = 0; i < SCHUR_BATCH; i++) b[i] = j0 + i < NT ? u[j0 + i] : 0.0;
#pragma unroll
    for (int i = 0; i < SCHUR_BATCH; i++)
      if (j0 + i < NT) c[j0 + i] = (T)fma(-a, b[i], (double)c[j0 + i]);
    asm volatile("" ::: "memory");  // the next batch's reads stay after this one: hoisted, they hold registers
  }
}

// The Schur update of the trailing block by the lead rows, one lead row at a time. This lane's multipliers go
// through the staging area (free during the factorization) CHUNK at a time, so that the loop over the rows stays
// rolled: unrolled, the compiler would load every row's broadcast values at once.
template <typename T>
__device__ __forceinline__ void schur(T (&c)[NT], const LeadRow& a21, double* ws, int lane) {
  double* stage = ws + L_STAGE;
#pragma unroll
  for (int k0 = 0; k0 < K0; k0 += CHUNK) {
#pragma unroll
    for (int j = 0; j < CHUNK; j++)
      if (k0 + j < K0) stage[j * NT + lane] = a21[k0 + j];
    wave_sync();
#pragma unroll 1
    for (int k = k0; k < (k0 + CHUNK < K0 ? k0 + CHUNK : K0); k++) schur_row(c, stage[(k - k0) * NT + lane], ws, k, lane);
    wave_sync();
  }
}

// LU of the lead columns (right-looking: the lead rows dense in LDS, this lane's trailing row in a21) and the Schur
// update of the trailing block (in T). dl: 1 / U_ii of lead row lane.
template <typename T>
__device__ __forceinline__ void factor_lead(T (&c)[NT], LeadRow& a21, double& dl, double* ws, int lane) {
  dl = 1.0;
  if constexpr (K0 <= SMALL_LEAD) {
    // few lead rows: the steps unrolled
#pragma unroll
    for (int k = 0; k < K0; k++) {
      const double dk = 1.0 / ws[k * NS + k];
      if (lane == k) dl = dk;
      if (lane > k && lane < K0) ws[lane * NS + k] *= dk;
      a21[k] *= dk;
      wave_sync();
#pragma unroll
      for (int i = k + 1; i < K0; i++) {
        const double l = ws[i * NS + k];
#pragma unroll
        for (int j0 = k + 1; j0 < NS; j0 += WAVE) {
          const int j = j0 + lane;
          if (j < NS) ws[i * NS + j] = fma(-l, ws[k * NS + j], ws[i * NS + j]);
        }
      }
#pragma unroll
      for (int j = k + 1; j < K0; j++) a21[j] = fma(-a21[k], ws[k * NS + j], a21[j]);
      wave_sync();
    }
  } else {
    // the steps unrolled for the registers of a21, the lead rows' updates rolled
#pragma unroll
    for (int k = 0; k < K0; k++) {
      const double dk = 1.0 / ws[k * NS + k];
      if (lane == k) dl = dk;
      if (lane > k && lane < K0) ws[lane * NS + k] *= dk;
      a21[k] *= dk;
      wave_sync();
#pragma unroll 1
      for (int i = k + 1; i < K0; i++) {
        const double l = ws[i * NS + k];
        for (int j = k + 1 + lane; j < NS; j += WAVE) ws[i * NS + j] = fma(-l, ws[k * NS + j], ws[i * NS + j]);
      }
#pragma unroll
      for (int j = k + 1; j < K0; j++) a21[j] = fma(-a21[k], ws[k * NS + j], a21[j]);
      wave_sync();
    }
  }
  schur(c, a21, ws, lane);
}

// The trailing block's LU in registers (FP64). dt: 1 / U_ii of trailing row lane.
__device__ __forceinline__ void factor_trailing(double (&c)[NT], double& dt, double* ws, int lane) {
  // the trailing block: at pivot k every lane takes row k's tail from lane k by v_readlane, as wave-uniform values
  // in scalar registers (no LDS round trip, no wait)
  // 1 / U_kk goes through LDS (the union region is free here): a select per step into a register would have the
  // compiler hold every step's reciprocal to the end
  double* dinv = ws + L_UNION;
#pragma unroll
  for (int k = 0; k < NT; k++) {
    const double dk = 1.0 / readlane(c[k], k);
    if (lane == 0) dinv[k] = dk;
    // column k is scaled by 1 / U_kk off the diagonal: below the pivot it becomes the multiplier L_ik, above it
    // U_ik / U_kk, the unit upper factor the backward solve uses (U = U' diag(U_kk))
    const int ln = lane_after(lane, dk);  // the step's lane masks made here, not hoisted and spilled for every step
    if (ln != k) c[k] *= dk;
    const double l = ln > k ? c[k] : 0.0;
#pragma unroll
    for (int j = k + 1; j < NT; j++) {
      // at most BROADCAST_WINDOW of row k's broadcasts ahead of their updates: hoisted further, they would hold up
      // to 126 scalar registers and spill them to vector lanes
      if (j >= k + 1 + BROADCAST_WINDOW) tie(c[j], c[j - BROADCAST_WINDOW]);
      // U_kj is a structural zero of the filled pattern (exactly 0.0): no broadcast, no update
      if ((TRAIL_FILL[k] >> j) & 1ull) c[j] = fma(-l, readlane(c[j], k), c[j]);
    }
  }
  wave_sync();
  dt = lane < NT ? dinv[lane] : 1.0;
  wave_sync();
}

// LU of the lead columns (right-looking, dense in LDS), the Schur update of the trailing block, then the trailing
// block's LU in registers. dl: 1 / U_ii of lead row lane; dt: of trailing row lane.
__device__ __forceinline__ void factor(double (&c)[NT], LeadRow& a21, double& dl, double& dt, double* ws, int lane) {
  factor_lead(c, a21, dl, ws, lane);
  factor_trailing(c, dt, ws, lane);
}

// y <- M0^-1 y for NV vectors (lead / trailing parts), with factor's factors, in three parts: the lead columns of L,
// the trailing block, the lead rows of U.
template <int NV>
__device__ __forceinline__ void solve_lead_forward(double (&yl)[NV], double (&yt)[NV], const LeadRow& a21,
                                                   const double* ws, int lane) {
#pragma unroll
  for (int k = 0; k < K0; k++) {  // unit L, lead columns (the trailing rows' entries in a21)
    const double lk = (lane > k && lane < K0) ? ws[lane * NS + k] : 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const double yk = readlane(yl[v], k);
      yl[v] = fma(-lk, yk, yl[v]);
      yt[v] = fma(-a21[k], yk, yt[v]);
    }
  }
}

template <int NV>
__device__ __forceinline__ void solve_trailing(double (&yt)[NV], const double (&c)[NT], double dt, int lane) {
  // the trailing block's solves: at step k every lane takes lane k's final value by v_readlane and is updated with
  // its factor entry in column k, unmasked; the value read, lane k's final one, is written into lane k of the record
  // by v_writelane (before the update spoils lane k; the lanes already recorded are spoilt freely), so a step has no
  // compare or select, and the chain from step to step is a readlane and an fma. pin keeps each step's record in its
  // step: left free, the compiler sinks the records to the end and holds every step's value until then
  double rec[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) rec[v] = 0.0;  // lanes past NT keep 0
  // (a step's writelanes come after its fmas: a written scalar must wait a cycle before v_writelane reads it, and a
  // written vector register before v_readlane reads it, so this order needs no wait instructions)
  double yk[NV];
#pragma unroll
  for (int k = 0; k < NT; k++) {  // unit L, trailing block
#pragma unroll
    for (int v = 0; v < NV; v++) yk[v] = readlane(yt[v], k);
#pragma unroll
    for (int v = 0; v < NV; v++)
      if (k + 1 < NT) yt[v] = fma(-c[k], yk[v], yt[v]);
#pragma unroll
    for (int v = 0; v < NV; v++) rec[v] = writelane(rec[v], yk[v], k);
#pragma unroll
    for (int v = 0; v < NV; v++) pin(rec[v], yt[v]);
  }
#pragma unroll
  for (int v = 0; v < NV; v++) yt[v] = 0.0;
#pragma unroll
  for (int k = NT - 1; k >= 0; k--) {  // unit U' = U diag(1 / U_kk), trailing block
#pragma unroll
    for (int v = 0; v < NV; v++) yk[v] = readlane(rec[v], k);
#pragma unroll
    for (int v = 0; v < NV; v++)
      if (k > 0) rec[v] = fma(-c[k], yk[v], rec[v]);
#pragma unroll
    for (int v = 0; v < NV; v++) yt[v] = writelane(yt[v], yk[v], k);
#pragma unroll
    for (int v = 0; v < NV; v++) pin(rec[v], yt[v]);
  }
#pragma unroll
  for (int v = 0; v < NV; v++) yt[v] *= dt;
}

template <int NV>
__device__ __forceinline__ void solve_lead_backward(double (&yl)[NV], const double (&yt)[NV], double dl, double* ws,
                                                    int lane) {
  if constexpr (K0 > 0 && K0 <= SMALL_LEAD) {
    // U of the lead rows, trailing columns: lane t's share of each lead row, summed over the lanes
    double sum[K0 > 0 ? K0 * NV : 1];
#pragma unroll
    for (int i = 0; i < K0; i++) {
      const double a = lane < NT ? ws[i * NS + K0 + lane] : 0.0;
#pragma unroll
      for (int v = 0; v < NV; v++) sum[i * NV + v] = a * yt[v];
    }
    wave_sums(sum);
#pragma unroll
    for (int i = 0; i < K0; i++) {
#pragma unroll
      for (int v = 0; v < NV; v++) yl[v] = lane == i ? yl[v] - sum[i * NV + v] : yl[v];
    }
#pragma unroll
    for (int k = K0 - 1; k >= 0; k--) {  // U, lead columns
      const double uk = lane < k ? ws[lane * NS + k] : 0.0;
#pragma unroll
      for (int v = 0; v < NV; v++) yl[v] = fma(-uk, readlane(yl[v] * dl, k), yl[v]);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) yl[v] *= dl;
  } else if (K0 > 0) {
    // U of the lead rows, trailing columns: the trailing solution through LDS (the union region is free between
    // the right-hand sides), a batch of columns at a time
    double* xv = ws + L_UNION;
    wave_sync();
#pragma unroll
    for (int v = 0; v < NV; v++)
      if (lane < NT) xv[v * NT + lane] = yt[v];
    wave_sync();
    if (lane < K0) {
#pragma unroll 1
      for (int t0 = 0; t0 < NT; t0 += STREAM_BATCH) {
        double a[STREAM_BATCH];
#pragma unroll
        for (int i = 0; i < STREAM_BATCH; i++) a[i] = t0 + i < NT ? ws[lane * NS + K0 + t0 + i] : 0.0;
#pragma unroll
        for (int v = 0; v < NV; v++) {
#pragma unroll
          for (int i = 0; i < STREAM_BATCH; i++)
            if (t0 + i < NT) yl[v] = fma(-a[i], xv[v * NT + t0 + i], yl[v]);
        }
      }
    }
    wave_sync();
#pragma unroll 1
    for (int k = K0 - 1; k >= 0; k--) {  // U, lead columns
      const double uk = lane < k ? ws[lane * NS + k] : 0.0;
#pragma unroll
      for (int v = 0; v < NV; v++) yl[v] = fma(-uk, readlane(yl[v] * dl, k), yl[v]);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) yl[v] *= dl;
  }
}

template <int NV>
__device__ __forceinline__ void solve(double (&yl)[NV], double (&yt)[NV], const double (&c)[NT], const LeadRow& a21,
                                      double dl, double dt, double* ws, int lane) {
  solve_lead_forward(yl, yt, a21, ws, lane);
  solve_trailing(yt, c, dt, lane);
  solve_lead_backward(yl, yt, dl, ws, lane);
}

__device__ __forceinline__ double vsum(double yl, double yt, int lane) {
  return wave_sum((lane < K0 ? yl : 0.0) + (lane < NT ? yt : 0.0));
}

// ---- step control (the reference's step_factor, next_try and try_error: common.cuh:87-133, 568-589) ----

struct Step {
  double h, t, nok, pen, pha;
  bool prej;
};

__device__ __forceinline__ double step_factor(double en, bool ok, bool prej, double pen, double pha, double ha) {
  const bool efin = isfinite(en);
  double fac = efin ? STEP_SAFETY / cbrt(fmax(en, ERR_FLOOR)) : FAC_NONFINITE;
  fac = fmin(fmax(fac, ok ? FACACC : FACREJ), FACMAX);
  if (!ok && efin && prej && en > 1.0 && pen > 1.0 && pha > ha) {  // secant: err ~ h^pe through the two tries
    const double pe = log(pen / en) / log(pha / ha);
    fac = pe > SECANT_MIN_ORDER ? fmax(pow(SECANT_TARGET / en, 1.0 / fmin(pe, SECANT_MAX_ORDER)), SECMIN) : SECMIN;
  }
  if (!ok) fac = fmin(fac, FAC_AFTER_REJECT);
  return fac;
}

// The try's error (RMS; while no step of the call is accepted, the larger of RMS and max-norm) and its new state y.
__device__ __forceinline__ double try_error(double Nl, double Nt, double K1l, double K1t, double K3l, double K3t,
                                            double K4l, double K4t, double& yl, double& yt, bool first, double rtol,
                                            double atol, int lane, bool& ok) {
  yl = Nl + 2.0 * K1l + K3l + K4l;
  yt = Nt + 2.0 * K1t + K3t + K4t;
  const double ql = K4l / (atol + rtol * fmax(fabs(Nl), fabs(yl)));
  const double qt = K4t / (atol + rtol * fmax(fabs(Nt), fabs(yt)));
  const bool hl = lane < K0, ht = lane < NT;
  const double e2 = wave_sum((hl ? ql * ql : 0.0) + (ht ? qt * qt : 0.0));
  const double em = wave_max(fmax(hl ? fabs(ql) : 0.0, ht ? fabs(qt) : 0.0));
  const bool nn = wave_all((!hl || yl >= -atol) && (!ht || yt >= -atol));
  const double er = sqrt(e2 / NS);
  const double en = (first && isfinite(er)) ? fmax(er, em) : er;
  ok = (en <= 1.0) && nn && isfinite(en);
  return en;
}

// x = N / sum N at the cell's output row (network order); returns the sum.
__device__ __forceinline__ void write_state(double Nl, double Nt, double* __restrict__ out, NetP net,
                                            int lane) {
  const gptr<int> perm = opaque(net)->perm;
  const double s = vsum(Nl, Nt, lane);
  if (lane < K0) out[perm[lane]] = Nl / s;
  if (lane < NT) out[perm[K0 + lane]] = Nt / s;
}
