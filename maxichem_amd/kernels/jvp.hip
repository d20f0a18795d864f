// The tangent (JVP) of a call: the call's own tries, bit for bit (the call kernel's device functions in the call
// kernel's order), and along them the tangent of the state through every accepted try, then through the guard's
// normalisation and share correction. Under the frozen contract the accepted step sizes are held fixed (a rejected
// try drops out); under the adaptive one every try is carried with the tangent of its step size, through the error
// estimate and the step factor, and dt and h have tangents too. The inputs' tangents: x, T, P, J, the log rate
// multipliers lm and the rate table's Arrhenius columns, and dt and h (the reference's tangent kernel,
// maxichem/kernels/jvp.cuh, and its contracts in docs/method.md, "Derivatives").
//
// An accepted try of step h from N solves W K_i = b_i with W = I / (gamma h) - dF/dN at N. Its tangent, with h fixed:
//   W Kdot_i = bdot_i + Jdot K_i,
// where Jdot K_i, the Jacobian's derivative along the tangent applied to K_i, is the eps-del part of
// F(N + eps Ndot + del K_i; rates + eps ratesdot), and bdot_1 = Fdot(N), bdot_2 = Fdot(N) + 4 Kdot_1 / h,
// bdot_3 = Fdot(Y3)[Ydot3] + (Kdot_1 - Kdot_2) / h, bdot_4 = Fdot(Y4)[Ydot4] + (Kdot_1 - Kdot_2 - 8/3 Kdot_3) / h.
// The solves reuse the try's factors of M0 and its Sherman-Morrison vector. The state's tangent is then
// Ndot + 2 Kdot_1 + Kdot_3 + Kdot_4 where the new state is not clamped at 0 (and 0 where it is). With h moving
// (the adaptive contract) each stage adds hdot / (gamma h^2) K_i, and each K / h term of bdot its -hdot K / h^2.
//
// F is evaluated in hyper-dual numbers (value, eps, del, eps-del parts) reaction by reaction: n = nd Y / sum Y, every
// rate a product of its coefficient and up to three n, S v by the rhs's lane streams, F = (sum Y / nd) S v.

// ---- dual numbers (value, first derivative along the tangent) for the rate coefficients ----

struct Dual {
  double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, fma(a.d, b.v, a.v * b.d)}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual dexp(Dual a) {
  const double e = exp(a.v);
  return {e, e * a.d};
}
__device__ __forceinline__ Dual dlog10(Dual a) { return {log10(a.v), a.d / (a.v * 2.302585092994045684)}; }
__device__ __forceinline__ Dual dexp10(Dual a) {
  const double e = exp10(a.v);
  return {e, e * a.d * 2.302585092994045684};
}

// The rate law of row q at T (with its tangent) and nd (with its tangent): the call's law() in dual numbers. qd: the
// tangents of the row's nine Arrhenius columns (ln A, B, C of k0, of the second k0 and of k_inf), or null; the Troe
// parameters and the efficiency have none.
__device__ __forceinline__ Dual law_dual(gptr<double> __restrict__ q, const double* __restrict__ qd, int fl, Dual T,
                                         Dual lT, Dual iT, Dual nd) {
  auto a = [&](int j) { return Dual{q[j], qd ? qd[j] : 0.0}; };
  Dual k = dexp(a(0) + a(1) * lT - a(2) * iT);
  if (fl & FL_SECOND_K0) k = k + dexp(a(3) + a(4) * lT - a(5) * iT);
  if (fl & FL_FALLOFF) {
    const Dual kinf = dexp(a(6) + a(7) * lT - a(8) * iT);
    const Dual pr = k * nd / kinf;
    k = k.v > 0.0 ? k / (Dual{1.0, 0.0} + pr) : Dual{0.0, 0.0};
#if HAS_TROE
    if ((fl & FL_TROE) && k.v > 0.0) {
      const Dual one{1.0, 0.0};
      const Dual fc = dlog10((Dual{1.0 - q[9], 0.0}) * dexp(Dual{0.0, 0.0} - T / Dual{q[10], 0.0}) +
                             Dual{q[9], 0.0} * dexp(Dual{0.0, 0.0} - T / Dual{q[11], 0.0}) +
                             dexp(Dual{-q[12], 0.0} / T));
      Dual y;
      if (fl & FL_VM11) {
        y = dlog10(pr) / (Dual{TROE_N0, 0.0} - TROE_N1 * fc);
      } else if (isfinite(log10(pr.v))) {
        const Dual x = dlog10(pr) - Dual{TROE_C0, 0.0} - TROE_C1 * fc;
        y = x / (Dual{TROE_N0, 0.0} - TROE_N1 * fc - TROE_D * x);
      } else {
        y = Dual{-1.0 / TROE_D, 0.0};
      }
      k = k * dexp10(fc / (one + y * y));
    }
#endif
  }
#if HAS_EFF
  if (fl & FL_EFF) k = q[13] * k;
#endif
  return k;
}

// The tangents of every reaction's coefficients (kfdot, krdot into rkd [2 NR]) from the tangents of T, nd, J (Jdc),
// the log multipliers (lmc and their tangents lmdc, each null or [NR]) and the rate table's Arrhenius columns (rcd
// [NR][9] or null). The coefficients follow the call's cell_rates: G / RT from NASA-9, whose T-derivative is
// -H / (R T^2); the collision limit KCOLL / q of the larger direction q takes q's tangent (half of each direction's at a
// tie of the two, and half of it when q is at the limit); then both directions times exp(lm).
__device__ __forceinline__ void rates_tangent(double* ts, double* rkd, double T, double Td, double nd, double ndd,
                                              NetP net, const double* Jc, const double* Jdc, const double* lmc,
                                              const double* lmdc, const double* rcd, int lane) {
  net = opaque(net);
  const gptr<double> rc = net->rc;
  const gptr<int> rf = net->rf;
  const gptr<u64> rxe = net->rxe;
  double* g = ts + J_G;    // G / RT
  double* gd = g + NS + 1;  // its tangent
  const double lT = log(T), iT = 1.0 / T, lKT = log(KBOLTZ * T / P0BAR);
  const gptr<double> cf = net->nasa + (T < NASA_SPLIT_T ? 0 : NS * 9);
  for (int i = lane; i < NS; i += WAVE) {
    const gptr<double> a = cf + 9 * i;
    g[i] = gibbs_rt(a, T, lT, iT);
    const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    const double h = -a[0] * iT * iT + a[1] * lT * iT + a[2] + a[3] * T / 2 + a[4] * T2 / 3 + a[5] * T3 / 4 +
                     a[6] * T4 / 5 + a[7] * iT;  // H / RT
    gd[i] = -h * iT * Td;
  }
  if (lane == 0) g[NS] = gd[NS] = 0.0;
  wave_sync();
  const Dual Tq{T, Td}, lTq{lT, Td * iT}, iTq{iT, -Td * iT * iT}, ndq{nd, ndd};
  const double lKTd = Td * iT;
  for (int r = lane; r < NR; r += WAVE) {
    const int fl = rf[r];
    Dual f{0.0, 0.0}, b{0.0, 0.0};
    if (fl & FL_PHOTO) {
      f = Dual{Jc[r - NRT], Jdc ? Jdc[r - NRT] : 0.0};
    } else {
      const u64 w = rxe[r];
      const Dual k = law_dual(rc + RCW * r, rcd ? rcd + 9 * r : nullptr, fl, Tq, lTq, iTq, ndq);
      const int s0 = species_of(w, 0), s1 = species_of(w, 1), s2 = species_of(w, 2);
      const int s3 = species_of(w, 3), s4 = species_of(w, 4), s5 = species_of(w, 5);
      const int mf = (fl >> FL_NF_SHIFT) & 3, mr = (fl >> FL_NR_SHIFT) & 3;
      const Dual dg{g[s3] + g[s4] + g[s5] - g[s0] - g[s1] - g[s2], gd[s3] + gd[s4] + gd[s5] - gd[s0] - gd[s1] - gd[s2]};
      const bool tb = (fl & FL_THREE_BODY) != 0;
      const Dual c3 = tb ? ndq : Dual{1.0, 0.0};
      f = k * c3;
      b = (k * dexp(dg + (double)(mr - mf) * Dual{lKT, lKTd})) * c3;
      auto bim = [&](int m, Dual kk) {
        return m >= 2 ? kk * (m == 3 ? ndq : Dual{1.0, 0.0}) : (tb ? kk / ndq : Dual{0.0, 0.0});
      };
      const Dual qf = bim(mf, f), qr = bim(mr, b);
      Dual qb = qf.v > qr.v ? qf : (qr.v > qf.v ? qr : Dual{qf.v, 0.5 * (qf.d + qr.d)});
      if (qb.v >= KCOLL) {  // at the limit itself the factor is 1 and takes half of its tangent
        if (qb.v == KCOLL) qb.d *= 0.5;
        const Dual c = Dual{KCOLL, 0.0} / qb;
        f = f * c;
        b = b * c;
      }
    }
    if (lmc || lmdc) {  // a tangent of lm without lm: at lm = 0
      const double e = lmc ? exp(lmc[r]) : 1.0;
      const Dual m{e, lmdc ? e * lmdc[r] : 0.0};
      f = f * m;
      b = b * m;
    }
    rkd[r] = f.d;
    rkd[NR + r] = b.d;
  }
  wave_sync();
}

// ---- F in hyper-dual numbers ----

// A hyper-dual number: value, eps part, del part, eps-del part.
struct Hyper {
  double v, e, d, ed;
};
__device__ __forceinline__ Hyper hmul(Hyper a, Hyper b) {
  return {a.v * b.v, fma(a.e, b.v, a.v * b.e), fma(a.d, b.v, a.v * b.d),
          fma(a.ed, b.v, fma(a.e, b.d, fma(a.d, b.e, a.v * b.ed)))};
}

// The eps part (Fe) of F(Y + eps Ye; rates + eps ratesdot, nd + eps ndd) and, with HD, the eps-del part (Fed) of
// F(Y + eps Ye + del K; ...): lead / trailing parts per lane. Uses the tangent region ts (n, v and pieces, four parts
// each) and the call's rate coefficients rk and their tangents rkd.
template <bool HD>
__device__ __forceinline__ void rhs_tangent(double Yl, double Yt, double Yel, double Yet, double Kl, double Kt, double& Fel,
                            double& Fet, double& Fedl, double& Fedt, NetP net, double nd, double ndd,
                            double* ts, const double* rk, const double* rkd, int lane) {
  const bool hl = lane < K0, ht = lane < NT;
  double sums[3] = {(hl ? Yl : 0.0) + (ht ? Yt : 0.0), (hl ? Yel : 0.0) + (ht ? Yet : 0.0),
                    HD ? (hl ? Kl : 0.0) + (ht ? Kt : 0.0) : 0.0};
  wave_sums(sums);
  const double nt = sums[0], nte = sums[1], ntd = sums[2];
  // x = nd / ntot in hyper-duals: nd has an eps part, 1 / ntot all four
  const double in = 1.0 / nt;
  const Hyper inv{in, -nte * in * in, -ntd * in * in, 2.0 * nte * ntd * in * in * in};
  const Hyper xf = hmul(Hyper{nd, ndd, 0.0, 0.0}, inv);
  double* n = ts + J_N;  // [4][NS + 1]
  constexpr int SN = NS + 1;
  auto put = [&](int i, double y, double ye, double kd) {
    const Hyper v = hmul(xf, Hyper{y, ye, kd, 0.0});
    n[i] = v.v;
    n[SN + i] = v.e;
    n[2 * SN + i] = v.d;
    n[3 * SN + i] = v.ed;
  };
  if (hl) put(lane, Yl, Yel, HD ? Kl : 0.0);
  if (ht) put(K0 + lane, Yt, Yet, HD ? Kt : 0.0);
  if (lane == 0) {
    n[NS] = 1.0;
    n[SN + NS] = n[2 * SN + NS] = n[3 * SN + NS] = 0.0;
  }
  wave_sync();
  net = opaque(net);
  const gptr<u64> rxe = net->rxe;
  const gptr<u32> svw = net->svw;
  const gptr<u32> rowp = net->rowp;
  double* V = ts + J_V;  // [4][NR + 1]
  constexpr int SV = NR + 1;
  auto nh = [&](int i) { return Hyper{n[i], n[SN + i], HD ? n[2 * SN + i] : 0.0, HD ? n[3 * SN + i] : 0.0}; };
  // every round's reaction word loaded first (in bounds: the last round's spare lanes read reaction NR - 1), so the
  // rounds wait for one memory latency, not one each
  u64 wr[RROUNDS];
#pragma unroll
  for (int q = 0; q < RROUNDS; q++) wr[q] = rxe[q * WAVE + lane < NR ? q * WAVE + lane : NR - 1];
#pragma unroll
  for (int q = 0; q < RROUNDS; q++) {
    const int r = q * WAVE + lane;
    if (r >= NR) break;
    const u64 w = wr[q];
    const Hyper pf = hmul(hmul(nh(species_of(w, 0)), nh(species_of(w, 1))), nh(species_of(w, 2)));
    const Hyper pr = hmul(hmul(nh(species_of(w, 3)), nh(species_of(w, 4))), nh(species_of(w, 5)));
    const Hyper vf = hmul(Hyper{rk[r], rkd[r], 0.0, 0.0}, pf);
    const Hyper vr = hmul(Hyper{rk[NR + r], rkd[NR + r], 0.0, 0.0}, pr);
    V[r] = vf.v - vr.v;
    V[SV + r] = vf.e - vr.e;
    V[2 * SV + r] = vf.d - vr.d;
    V[3 * SV + r] = vf.ed - vr.ed;
  }
  if (lane == 0) V[NR] = V[SV + NR] = V[2 * SV + NR] = V[3 * SV + NR] = 0.0;
  wave_sync();
  // S v by the rhs's lane streams, four parts at once
  double* part = ts + J_PART;  // [4][NPART]
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // the words of TB rounds loaded together, ahead of their terms (one memory latency per batch; a batch of 16 would
  // make the tangent kernel spill)
  constexpr int TB = 8;
#pragma unroll 1
  for (int q0 = 0; q0 < SV_ROUNDS; q0 += TB) {
    u32 wb[TB];
#pragma unroll
    for (int i = 0; i < TB; i++) wb[i] = q0 + i < SV_ROUNDS ? svw[(q0 + i) * WAVE + lane] : (u32)NR;
#pragma unroll
    for (int i = 0; i < TB; i++) {
      const u32 w = wb[i];
      const int r = (int)(w & SV_R_MASK);
      const double c = coefficient(w, SV_NEG, SV_DBL);
#pragma unroll
      for (int p = 0; p < 4; p++) acc[p] = fma(c, V[p * SV + r], acc[p]);
      if ((w >> SV_END) & 1u) {
        const int piece = (int)((w >> SV_PIECE) & SV_PIECE_MASK);
#pragma unroll
        for (int p = 0; p < 4; p++) {
          part[p * NPART + piece] = acc[p];
          acc[p] = 0.0;
        }
      }
    }
  }
  wave_sync();
  // F = (ntot / nd) S v in hyper-duals
  const double ind = 1.0 / nd;
  const Hyper ixf = hmul(Hyper{nt, nte, ntd, 0.0}, Hyper{ind, -ndd * ind * ind, 0.0, 0.0});
  auto row = [&](int k, double& fe, double& fed) {
    const u32 rp = rowp[k];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = 0; e < (int)(rp >> 16); e++) {
#pragma unroll
      for (int p = 0; p < 4; p++) s[p] += part[p * NPART + (rp & 0xffffu) + e];
    }
    const Hyper F = hmul(ixf, Hyper{s[0], s[1], s[2], s[3]});
    fe = F.e;
    fed = F.ed;
  };
  Fel = Fet = Fedl = Fedt = 0.0;
  if (hl) row(lane, Fel, Fedl);
  if (ht) row(K0 + lane, Fet, Fedt);
  wave_sync();
}

// W^-1 y (lead / trailing parts) with a try's factors: M0^-1 y, then the Sherman-Morrison term of w 1^T.
__device__ __forceinline__ void w_solve(double& yl, double& yt, const double (&c)[NT], const LeadRow& a21,
                                        const StageKeep& kp, double dtr, double* ws, int lane) {
  double vl[1] = {yl}, vt[1] = {yt};
  solve<1>(vl, vt, c, a21, kp.dl, dtr, ws, lane);
  const double sk = vsum(vl[0], vt[0], lane) * kp.c0;
  yl = vl[0] + sk * kp.Ul;
  yt = vt[0] + sk * kp.Ut;
}

// The tangent of the guard's result x' (network order, into out) from the stepped state N and its tangent, and the
// input x0 and its tangent x0d, for a cell that passed the guard: x = N / sum N, then x' = x g / sum x g with
// g = clamp(1 + sum_e lambda_e a_e) and (A^T X A) lambda = e0 s1 / s0 - e1 (live elements as the guard takes them).
__device__ __forceinline__ void guard_tangent(double Nl, double Nt, double Ndl, double Ndt, const double* __restrict__ x0,
                              const double* __restrict__ x0d, double* __restrict__ out, NetP net, int lane) {
  net = opaque(net);
  const gptr<int> perm = net->perm;
  const gptr<double> atoms = net->atoms;
  const bool hl = lane < K0, ht = lane < NT;
  const int il = hl ? perm[lane] : 0, it = ht ? perm[K0 + lane] : 0;
  double Al[NEL], At[NEL];
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    Al[e] = hl ? atoms[il * NEL + e] : 0.0;
    At[e] = ht ? atoms[it * NEL + e] : 0.0;
  }
  double ss[2] = {(hl ? Nl : 0.0) + (ht ? Nt : 0.0), (hl ? Ndl : 0.0) + (ht ? Ndt : 0.0)};
  wave_sums(ss);
  const double s = ss[0], sd = ss[1];
  const double xl = hl ? Nl / s : 0.0, xt = ht ? Nt / s : 0.0;
  const double xdl = hl ? (Ndl - xl * sd) / s : 0.0, xdt = ht ? (Ndt - xt * sd) / s : 0.0;
  const double al = hl ? x0[il] : 0.0, at = ht ? x0[it] : 0.0;
  const double adl = hl ? x0d[il] : 0.0, adt = ht ? x0d[it] : 0.0;
  // e0, e1 and their tangents
  double e[4 * NEL];
#pragma unroll
  for (int k = 0; k < NEL; k++) {
    e[k] = fma(Al[k], al, At[k] * at);
    e[NEL + k] = fma(Al[k], xl, At[k] * xt);
    e[2 * NEL + k] = fma(Al[k], adl, At[k] * adt);
    e[3 * NEL + k] = fma(Al[k], xdl, At[k] * xdt);
  }
  wave_sums(e);
  double s0 = 0.0, s1 = 0.0, s0d = 0.0, s1d = 0.0;
#pragma unroll
  for (int k = 0; k < NEL; k++) {
    s0 += e[k];
    s1 += e[NEL + k];
    s0d += e[2 * NEL + k];
    s1d += e[3 * NEL + k];
  }
  // A^T X A and its tangent
  double m[2 * NSYM];
#pragma unroll
  for (int a = 0; a < NEL; a++) {
#pragma unroll
    for (int b = a; b < NEL; b++) {
      m[sym_index(a, b)] = fma(Al[a] * Al[b], xl, (At[a] * At[b]) * xt);
      m[NSYM + sym_index(a, b)] = fma(Al[a] * Al[b], xdl, (At[a] * At[b]) * xdt);
    }
  }
  wave_sums(m);
  const double r = s1 / s0, rd = (s1d - r * s0d) / s0;
  double y[NEL], yd[NEL], d0[NEL], lam[NEL], lamd[NEL], l[NEL * NEL];
  bool live[NEL];
#pragma unroll
  for (int k = 0; k < NEL; k++) {
    y[k] = fma(e[k], r, -e[NEL + k]);
    yd[k] = e[2 * NEL + k] * r + e[k] * rd - e[3 * NEL + k];
    d0[k] = m[sym_index(k, k)];
  }
  // the guard's elimination on M (its multipliers kept), then both right-hand sides
  double mm[NSYM];
#pragma unroll
  for (int k = 0; k < NSYM; k++) mm[k] = m[k];
#pragma unroll
  for (int k = 0; k < NEL; k++) {
    live[k] = mm[sym_index(k, k)] > ELEMENT_PIVOT_MIN * d0[k];
#pragma unroll
    for (int j = k + 1; j < NEL; j++) {
      l[k * NEL + j] = live[k] ? mm[sym_index(k, j)] / mm[sym_index(k, k)] : 0.0;
#pragma unroll
      for (int f = j; f < NEL; f++) mm[sym_index(j, f)] = fma(-l[k * NEL + j], mm[sym_index(k, f)], mm[sym_index(j, f)]);
    }
  }
  auto solve_sym = [&](double (&b)[NEL], double (&z)[NEL]) {
#pragma unroll
    for (int k = 0; k < NEL; k++) {
#pragma unroll
      for (int j = k + 1; j < NEL; j++) b[j] = fma(-l[k * NEL + j], b[k], b[j]);
    }
#pragma unroll
    for (int k = NEL - 1; k >= 0; k--) {
      double v = b[k];
#pragma unroll
      for (int f = k + 1; f < NEL; f++) v = fma(-mm[sym_index(k, f)], z[f], v);
      z[k] = live[k] ? v / mm[sym_index(k, k)] : 0.0;
    }
  };
  solve_sym(y, lam);
  // M lamd = yd - Mdot lam, over the live elements
  double t[NEL];
#pragma unroll
  for (int a = 0; a < NEL; a++) {
    double v = yd[a];
#pragma unroll
    for (int b = 0; b < NEL; b++) v -= m[NSYM + (a <= b ? sym_index(a, b) : sym_index(b, a))] * lam[b];
    t[a] = v;
  }
  solve_sym(t, lamd);
  double gl = 1.0, gt = 1.0, gdl = 0.0, gdt = 0.0;
#pragma unroll
  for (int k = 0; k < NEL; k++) {
    if (Al[k] != 0.0) {
      gl = fma(lam[k], Al[k], gl);
      gdl = fma(lamd[k], Al[k], gdl);
    }
    if (At[k] != 0.0) {
      gt = fma(lam[k], At[k], gt);
      gdt = fma(lamd[k], At[k], gdt);
    }
  }
  // a clamp passes its tangent at the bound, and holds it beyond
  if (gl < 1.0 - GMAX || gl > 1.0 + GMAX) gdl = 0.0;
  if (gt < 1.0 - GMAX || gt > 1.0 + GMAX) gdt = 0.0;
  gl = fmin(fmax(gl, 1.0 - GMAX), 1.0 + GMAX);
  gt = fmin(fmax(gt, 1.0 - GMAX), 1.0 + GMAX);
  const double yl = xl * gl, yt = xt * gt, ydl = fma(xdl, gl, xl * gdl), ydt = fma(xdt, gt, xt * gdt);
  double S[2] = {(hl ? yl : 0.0) + (ht ? yt : 0.0), (hl ? ydl : 0.0) + (ht ? ydt : 0.0)};
  wave_sums(S);
  if (hl) out[il] = (ydl - yl / S[0] * S[1]) / S[0];
  if (ht) out[it] = (ydt - yt / S[0] * S[1]) / S[0];
}

// The tangent of a try's error (the call's try_error) along the state's tangent Ndot, the new state's Ydot and K4's
// K4dot: q = K4 / (atol + rtol max(|N|, |Y|)) per species, the RMS of q, and while no step of the call is accepted
// the larger of RMS and max |q|. Ties follow the reference's conventions (docs/method.md, "Derivatives"): half to
// each side of a tie of min or max, a maximum over species shared evenly among those attaining it, d|a| = 0 at 0, and a
// clamp passes at its bound. Wave-uniform.
__device__ __forceinline__ double error_tangent(double Nl, double Nt, double Yl, double Yt, double K4l, double K4t,
                                                double Ndl, double Ndt, double Ydl, double Ydt, double K4dl, double K4dt,
                                                bool first, double en, double rtol, double atol, int lane) {
  const bool hl = lane < K0, ht = lane < NT;
  auto part = [&](double N, double Y, double K4, double Nd, double Yd, double K4d, double& q, double& qd) {
    const double aN = fabs(N), aY = fabs(Y);
    const double w = atol + rtol * fmax(aN, aY);
    // d|a| = 0 at a = 0, and a tie of the max takes half of each side's tangent
    const double dN = N > 0.0 ? Nd : (N < 0.0 ? -Nd : 0.0), dY = Y > 0.0 ? Yd : (Y < 0.0 ? -Yd : 0.0);
    const double wd = rtol * (aN > aY ? dN : (aN < aY ? dY : 0.5 * (dN + dY)));
    q = K4 / w;
    qd = (K4d - q * wd) / w;
  };
  double ql, qld, qt, qtd;
  part(Nl, Yl, K4l, Ndl, Ydl, K4dl, ql, qld);
  part(Nt, Yt, K4t, Ndt, Ydt, K4dt, qt, qtd);
  if (!isfinite(en)) return 0.0;
  double S[2] = {(hl ? ql * ql : 0.0) + (ht ? qt * qt : 0.0), (hl ? ql * qld : 0.0) + (ht ? qt * qtd : 0.0)};
  wave_sums(S);
  const double er = sqrt(S[0] / NS);
  const double erd = er > 0.0 ? S[1] / (NS * er) : 0.0;
  if (!(first && isfinite(er))) return erd;
  const double a = hl ? fabs(ql) : 0.0, b = ht ? fabs(qt) : 0.0;
  const double da = hl ? (ql > 0.0 ? qld : (ql < 0.0 ? -qld : 0.0)) : 0.0;
  const double db = ht ? (qt > 0.0 ? qtd : (qt < 0.0 ? -qtd : 0.0)) : 0.0;
  const double em = wave_max(fmax(a, b));
  if (em < er) return erd;
  // the max over the species shared evenly among those attaining it
  double M[2] = {(a == em ? da : 0.0) + (b == em ? db : 0.0), (a == em ? 1.0 : 0.0) + (b == em ? 1.0 : 0.0)};
  wave_sums(M);
  const double emd = M[0] / M[1];
  return em > er ? emd : 0.5 * (emd + erd);
}

// The tangent of the call's step factor (step_factor, the same branches) along the error's tangent end, the previous
// try's error and step tangents pend and phad, and the step's tangent had: the law 0.9 err^(-1/3), held beyond its
// clamps, and the secant shrink after two rejections through its fitted order.
__device__ __forceinline__ double step_factor_tangent(double en, double end, bool ok, bool prej, double pen,
                                                      double pend, double pha, double phad, double ha, double had) {
  const bool efin = isfinite(en);
  double fac = efin ? STEP_SAFETY / cbrt(fmax(en, ERR_FLOOR)) : FAC_NONFINITE;
  double facd = (efin && en >= ERR_FLOOR) ? -fac * end / (3.0 * en) : 0.0;
  const double lo = ok ? FACACC : FACREJ;
  if (fac < lo || fac > FACMAX) facd = 0.0;
  fac = fmin(fmax(fac, lo), FACMAX);
  if (!ok && efin && prej && en > 1.0 && pen > 1.0 && pha > ha) {
    const double L1 = log(pen / en), L2 = log(pha / ha);
    const double pe = L1 / L2;
    if (pe > SECANT_MIN_ORDER) {
      const double ped = ((pend / pen - end / en) * L2 - L1 * (phad / pha - had / ha)) / (L2 * L2);
      const double o = fmin(pe, SECANT_MAX_ORDER), od = pe <= SECANT_MAX_ORDER ? ped : 0.0;
      const double r = SECANT_TARGET / en, a = 1.0 / o;
      const double v = pow(r, a);
      facd = v >= SECMIN ? v * (-od / (o * o) * log(r) - a * end / en) : 0.0;
      fac = fmax(v, SECMIN);
    } else {
      fac = SECMIN;
      facd = 0.0;
    }
  }
  if (!ok && fac > FAC_AFTER_REJECT) facd = 0.0;
  return facd;
}

// The call and its tangent: the call kernel's loop (no proposals, FP64) with the tangent carried through each
// accepted try (adapt 0: the frozen contract), or through every try with the step control's tangents (adapt 1: the
// adaptive contract, with the tangents of dt and h in dtd, hd, and of the result's h into hdo), then the guard with
// its tangent. D directions in one launch: the persistent queue hands out (direction, cell) pairs, each wavefront
// repeating its cell's call (the same bits for every direction) with that direction's tangent, so a small batch
// still fills the GPU. Direction d's tangents: x0d [D][B][NS], Td, Pd, dtd, hd [D][B], Jd [D][B][NJ], lmd [D][B][NR],
// rcd [D][NR][9] (each may be null: zero). Outputs: xd [D][B][NS], hdo [D][B], the tangents of the guard's result
// (network order) and of h; the call's results, counts and list come from direction 0's pairs (the other pairs write
// the same bits to xout). A cell the guard fails gets NaN, and mxc_jvp_recover then recomputes it with its tangent.
extern "C" __global__ void __launch_bounds__(WAVE* CPW)
    mxc_jvp(NetP __restrict__ net, const double* __restrict__ x0, const double* __restrict__ Ta,
            const double* __restrict__ Pa, const double* __restrict__ dta, const double* __restrict__ hin_a,
            double* __restrict__ hout, double* __restrict__ xout, int* __restrict__ doneo, double* __restrict__ noko,
            int* __restrict__ status,
            int* __restrict__ redo, int* __restrict__ list, int* __restrict__ count, double* __restrict__ tries,
            const double* __restrict__ Jin, int B, double rtol, double atol, int max_steps,
            const double* __restrict__ x0d, const double* __restrict__ Td, const double* __restrict__ Pd,
            const double* __restrict__ Jd, double* __restrict__ xd, int adapt, const double* __restrict__ dtd,
            const double* __restrict__ hd, double* __restrict__ hdo, const double* __restrict__ lm,
            const double* __restrict__ lmd, const double* __restrict__ rcd, int D, double* __restrict__ rsc) {
  extern __shared__ double lds[];
  const int lane = (int)(threadIdx.x & (WAVE - 1)), wv = (int)(threadIdx.x / WAVE);
  double* ws = lds + wv * (L_SIZE + J_SIZE);
  double* ts = ws + L_SIZE;
#if RATES_GLOBAL
  double* rk = rsc + (size_t)((int)blockIdx.x * CPW + wv) * (4 * NR);  // this wavefront's slot: the rates,
  double* rkd = rk + 2 * NR;                                            // then their tangents
#else
  double* rk = ws + L_KF;
  double* rkd = ts + J_KD;
#endif
  const int waves = (int)gridDim.x * CPW;
  int q = (int)blockIdx.x * CPW + wv;
  double c[NT];
  LeadRow a21;
  const int pairs = B * D;
  while (q < pairs) {
    const int dir = q / B, cell = q - dir * B;
    const bool lead = dir == 0;  // direction 0's pair writes the call's results and counts
    const size_t dc = (size_t)dir * B + cell;  // the pair's row in the per-direction arrays
    const double Tc = Ta[cell], dt = dta[cell], nd = Pa[cell] / (KBOLTZ * Tc);
    const double* xr = x0 + (size_t)cell * NS;
    const double* xdr = x0d + dc * NS;
    const gptr<int> perm = opaque(net)->perm;
    double Nl = lane < K0 ? xr[perm[lane]] : 0.0, Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
    double Ndl = lane < K0 ? xdr[perm[lane]] : 0.0, Ndt = lane < NT ? xdr[perm[K0 + lane]] : 0.0;
    // the call kernel's screen
    double p = (lane < K0 ? (Nl >= 0.0 ? Nl : DNAN) : 0.0) + (lane < NT ? (Nt >= 0.0 ? Nt : DNAN) : 0.0);
    double pm = fmax(lane < K0 ? Nl : 0.0, lane < NT ? Nt : 0.0);
    const double* Jc = NJ ? Jin + (size_t)cell * NJ : nullptr;
    for (int b = lane; b < NJ; b += WAVE) p += Jc[b] >= 0.0 ? 0.0 * Jc[b] : DNAN;
    const double* lmc = lm ? lm + (size_t)cell * NR : nullptr;
    if (lmc)  // a non-finite log multiplier makes the cell bad input
      for (int r = lane; r < NR; r += WAVE) p += 0.0 * lmc[r];
    const double ps = wave_sum(p), pmax = wave_max(pm);
    const bool fine = fabs(ps - 1.0) <= XSUMTOL && pmax > 0.0 && isfinite(Tc) && isfinite(nd) && nd > 0.0 &&
                      isfinite(dt) && dt >= 0.0;
    const int st = __builtin_amdgcn_readfirstlane(
        Tc < TMIN ? REFUSED_COLD : (Tc > TMAX ? REFUSED_HOT : (fine ? 0 : REFUSED_BAD_INPUT)));
    if (lane == 0 && lead) status[cell] = st;
    // nd = P / (k_B T): its tangent
    const double Tdc = Td ? Td[dc] : 0.0, Pdc = Pd ? Pd[dc] : 0.0;
    const double ndd = nd * (Pdc / Pa[cell] - Tdc / Tc);
    if (st == 0) {
      cell_rates(ws, rk, Tc, nd, net, Jc, lane, lmc);
      rates_tangent(ts, rkd, Tc, Tdc, nd, ndd, net, Jc, (NJ && Jd) ? Jd + dc * NJ : nullptr, lmc,
                    lmd ? lmd + dc * NR : nullptr, rcd ? rcd + (size_t)dir * NR * 9 : nullptr, lane);
    }
    const double hin = hin_a[cell];
    const double h0 = isfinite(hin) ? hin : dt;
    if (lane == 0 && lead && !isfinite(hin)) atomicAdd(count + COUNT_NONFINITE_H, 1);
    Step s = {fmax(fmin(h0, dt), H_MIN), 0.0, 0.0, 0.0, 0.0, false};
    // the adaptive contract's tangents of the step control (wave-uniform): of s.h, s.t, s.pen, s.pha; the first step
    // min(h, dt) clamped at H_MIN takes the tangent of the side it takes, half of each at a tie
    const bool ad = adapt != 0;
    const double dtdc = (ad && dtd) ? dtd[dc] : 0.0;
    double hdot = 0.0, tdot = 0.0, pendot = 0.0, phadot = 0.0;
    if (ad) {
      const double h0d = isfinite(hin) ? (hd ? hd[dc] : 0.0) : dtdc;
      const double md = h0 < dt ? h0d : (h0 > dt ? dtdc : 0.5 * (h0d + dtdc)), m = fmin(h0, dt);
      hdot = m > H_MIN ? md : (m < H_MIN ? 0.0 : 0.5 * md);  // the first step max(m, H_MIN)
    }
    int it = (dt > 0.0 && st == 0) ? 0 : max_steps;
    int ntry = 0;
    while (it < max_steps && s.t < dt * CALL_END) {
      const double ha = fmin(s.h, dt - s.t);
      const double rem = dt - s.t;
      const double had = ad ? (s.h < rem ? hdot : (s.h > rem ? dtdc - tdot : 0.5 * (hdot + dtdc - tdot))) : 0.0;
      double K1l, K1t, K3l, K3t, K4l, K4t;
      StageKeep kp;
      double dtr;
      stages(c, a21, Nl, Nt, ha, rtol, atol, net, nd, ws, rk, lane, K1l, K1t, K3l, K3t, K4l, K4t, &kp, &dtr);
      bool ok = false;
      double Yl, Yt;
      const bool first = s.t == 0.0;
      const double en = try_error(Nl, Nt, K1l, K1t, K3l, K3t, K4l, K4t, Yl, Yt, first, rtol, atol, lane, ok);
      double end = 0.0;  // the error's tangent (adaptive contract)
      if (ok || ad) {
        // the tangent of the try (every try under the adaptive contract: a rejected one sets the next step), with
        // its factors still in place
        const double ih = 1.0 / ha;
        double Fel, Fet, Gl, Gt, Ql, Qt;  // Fdot(N); Jdot K_i (eps-del parts)
        double K1dl, K1dt, K2dl, K2dt, K3dl, K3dt, K4dl, K4dt;
        rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K1l, K1t, Fel, Fet, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
        K1dl = Fel + Gl;
        K1dt = Fet + Gt;
        // adaptive: W = I / (gamma h) - J moves with h, W Kdot = bdot + Jdot K + hdot / (gamma h^2) K, and the stages'
        // K / h terms give -hdot / h^2 K
        const double sg = had * ih * ih / GAMMA, hq = had * ih * ih;
        if (ad) {
          K1dl = fma(sg, K1l, K1dl);
          K1dt = fma(sg, K1t, K1dt);
        }
        w_solve(K1dl, K1dt, c, a21, kp, dtr, ws, lane);
        rhs_tangent<true>(Nl, Nt, Ndl, Ndt, kp.K2l, kp.K2t, Ql, Qt, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
        K2dl = fma(RODAS_C21 * ih, K1dl, Fel) + Gl;
        K2dt = fma(RODAS_C21 * ih, K1dt, Fet) + Gt;
        if (ad) {
          K2dl += sg * kp.K2l - hq * RODAS_C21 * K1l;
          K2dt += sg * kp.K2t - hq * RODAS_C21 * K1t;
        }
        w_solve(K2dl, K2dt, c, a21, kp, dtr, ws, lane);
        // stage 3: F at Y3 = N + 2 K1 along Ydot3 = Ndot + 2 Kdot1; Jdot K3 at N
        double Y3el = Ndl + 2.0 * K1dl, Y3et = Ndt + 2.0 * K1dt, dummyl, dummyt;
        rhs_tangent<false>(Nl + 2.0 * K1l, Nt + 2.0 * K1t, Y3el, Y3et, 0.0, 0.0, Ql, Qt, dummyl, dummyt, net, nd,
                           ndd, ts, rk, rkd, lane);
        rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K3l, K3t, dummyl, dummyt, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
        K3dl = fma(ih, K1dl - K2dl, Ql) + Gl;
        K3dt = fma(ih, K1dt - K2dt, Qt) + Gt;
        if (ad) {
          K3dl += sg * K3l - hq * (K1l - kp.K2l);
          K3dt += sg * K3t - hq * (K1t - kp.K2t);
        }
        w_solve(K3dl, K3dt, c, a21, kp, dtr, ws, lane);
        // stage 4: F at Y4 = N + 2 K1 + K3 along Ndot + 2 Kdot1 + Kdot3; Jdot K4 at N
        rhs_tangent<false>(Nl + 2.0 * K1l + K3l, Nt + 2.0 * K1t + K3t, Y3el + K3dl, Y3et + K3dt, 0.0, 0.0, Ql, Qt,
                           dummyl, dummyt, net, nd, ndd, ts, rk, rkd, lane);
        rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K4l, K4t, dummyl, dummyt, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
        K4dl = fma(ih, fma(-RODAS_C43, K3dl, K1dl - K2dl), Ql) + Gl;
        K4dt = fma(ih, fma(-RODAS_C43, K3dt, K1dt - K2dt), Qt) + Gt;
        if (ad) {
          K4dl += sg * K4l - hq * fma(-RODAS_C43, K3l, K1l - kp.K2l);
          K4dt += sg * K4t - hq * fma(-RODAS_C43, K3t, K1t - kp.K2t);
        }
        w_solve(K4dl, K4dt, c, a21, kp, dtr, ws, lane);
        const double ydl = Ndl + 2.0 * K1dl + K3dl + K4dl, ydt = Ndt + 2.0 * K1dt + K3dt + K4dt;
        if (ad) end = error_tangent(Nl, Nt, Yl, Yt, K4l, K4t, Ndl, Ndt, ydl, ydt, K4dl, K4dt, first, en, rtol, atol, lane);
        if (ok) {
          // the new state's tangent: zero where the state is clamped at 0
          Ndl = lane < K0 && Yl >= 0.0 ? ydl : 0.0;
          Ndt = lane < NT && Yt >= 0.0 ? ydt : 0.0;
        }
      }
      if (ok) {
        Nl = fmax(Yl, 0.0);
        Nt = fmax(Yt, 0.0);
        s.t += ha;
        s.nok += 1.0;
        tdot += had;
      }
      const double fac = step_factor(en, ok, s.prej, s.pen, s.pha, ha);
      if (ad) {
        hdot = fma(had, fac, ha * step_factor_tangent(en, end, ok, s.prej, s.pen, pendot, s.pha, phadot, ha, had));
        pendot = end;
        phadot = had;
      }
      s.h = ha * fac;
      s.prej = !ok;
      s.pen = en;
      s.pha = ha;
      it++;
      ntry++;
    }
    const bool fin = st == 0 && s.t >= dt * CALL_END;
    double* xo = xout + (size_t)cell * NS;
    double* xdo = xd + dc * NS;
    int code = 0;
    if (st == 0)
      code = __builtin_amdgcn_readfirstlane(guard_wave(Nl, Nt, xr, xo, fin, s.nok, atol, net, lane));
    else
      write_state(Nl, Nt, xo, net, lane);
    if (st == 0 && code == 0) {
      guard_tangent(Nl, Nt, Ndl, Ndt, xr, xdr, xdo, net, lane);
    } else {
      for (int i = lane; i < NS; i += WAVE) xdo[i] = DNAN;
    }
    if (lane == 0 && hdo) hdo[dc] = !ad ? 0.0 : (st == 0 && code == 0 ? hdot : DNAN);
    if (lane == 0 && lead) {
      hout[cell] = s.h;
      doneo[cell] = fin ? 1 : 0;
      noko[cell] = s.nok;
      if (tries) tries[cell] = (double)ntry;
      if (st != 0) {
        atomicAdd(count + (st == REFUSED_COLD ? COUNT_COLD : st == REFUSED_HOT ? COUNT_HOT : COUNT_BAD_INPUT), 1);
      } else if (code != 0) {
        if (code & GUARD_NONFINITE) atomicAdd(count + COUNT_NONFINITE, 1);
        if (code & GUARD_UNFINISHED) atomicAdd(count + COUNT_UNFINISHED, 1);
        if (code & GUARD_DRIFT) atomicAdd(count + COUNT_DRIFT, 1);
        list[atomicAdd(count + COUNT_RECOMPUTED, 1)] = cell;
      }
      redo[cell] = code != 0 ? 1 : 0;
    }
    int nx = 0;
    if (lane == 0) nx = waves + atomicAdd(count + COUNT_QUEUE, 1);
    q = __builtin_amdgcn_readfirstlane(nx);
  }
}

// y <- W^-1 y for a lead / trailing vector pair, with the recovery's dense factors (through the LDS vector yv).
__device__ __forceinline__ void dense_solve_pair(const double* W, const int* pr, const double* dinv, double* yv,
                                                 double& vl, double& vt, int lane) {
  to_lds(yv, vl, vt, lane);
  dense_solve(W, pr, dinv, yv, lane);
  from_lds(yv, vl, vt, lane);
}

// The recovery of the cells the tangent kernel listed (mxc_recover's passes, bit for bit: the same device functions
// in the same order, with the multipliers lm), with the tangent carried through the pass whose result is the cell's:
// each pass starts again from the inputs' tangents; each try's stage tangents solve the dense W of the try
// (W Kdot = bdot + Jdot K, plus the step's terms under the adaptive contract) on its threshold-pivoted factors. A
// pass's first step min(h, dt split^k) takes the tangent of its smaller side (half of each at a tie), then that of
// max(., H_MIN). The result's tangent goes through the guard's, as the tangent kernel's; a cell that still fails the
// guard (flag 2) gets NaN. Workgroups of REC_WAVES wavefronts share each dense LU as mxc_recover's.
extern "C" __global__ void __launch_bounds__(WAVE* REC_WAVES)
    mxc_jvp_recover(NetP __restrict__ net, const double* __restrict__ x0, const double* __restrict__ Ta,
                    const double* __restrict__ Pa, const double* __restrict__ dta, const double* __restrict__ hin,
                    double* __restrict__ xout, double* __restrict__ hio, int* __restrict__ doneo,
                    double* __restrict__ noko, int* __restrict__ redo, const int* __restrict__ list,
                    int* __restrict__ count, double* __restrict__ tries, const double* __restrict__ Jin, double rtol,
                    double atol, int max_steps, const double* __restrict__ lm, const double* __restrict__ x0d,
                    const double* __restrict__ Td, const double* __restrict__ Pd, const double* __restrict__ Jd,
                    double* __restrict__ xd, int adapt, const double* __restrict__ dtd, const double* __restrict__ hd,
                    double* __restrict__ hdo, const double* __restrict__ lmd, const double* __restrict__ rcd,
                    double* __restrict__ rsc) {
  const int nfl = count[COUNT_RECOMPUTED];
  if ((int)blockIdx.x >= nfl) return;
  extern __shared__ double lds[];
  const int lane = (int)(threadIdx.x & (WAVE - 1)), wv = (int)(threadIdx.x / WAVE);
  double* ws = lds;
  double* ts = ws + L_SIZE;
#if RATES_GLOBAL
  double* rk = rsc + (size_t)blockIdx.x * (4 * NR);  // this workgroup's slot: the rates, then their tangents
  double* rkd = rk + 2 * NR;
#else
  double* rk = ws + L_KF;
  double* rkd = ts + J_KD;
#endif
  double* W = ts + J_SIZE;
  double* dinv = W + NS * NS;
  double* yv = dinv + NS;
  int* pr = (int*)(yv + NS);
  volatile int* cmd = pr + NS;
  if (wv != 0) {
    for (;;) {
      __syncthreads();
      if (*cmd == 0) return;
      dense_lu(W, pr, dinv, lane, wv);
    }
  }
  double c[NT];
  LeadRow a21;
  const bool ad = adapt != 0;
  int q = (int)blockIdx.x;
  while (q < nfl) {
    const int cell = list[q];
    const double Tc = Ta[cell], dt = dta[cell], nd = Pa[cell] / (KBOLTZ * Tc);
    const double* xr = x0 + (size_t)cell * NS;
    const double* xdr = x0d + (size_t)cell * NS;
    const double* Jc = NJ ? Jin + (size_t)cell * NJ : nullptr;
    const double* lmc = lm ? lm + (size_t)cell * NR : nullptr;
    cell_rates(ws, rk, Tc, nd, net, Jc, lane, lmc);
    const double Tdc = Td ? Td[cell] : 0.0, Pdc = Pd ? Pd[cell] : 0.0;
    const double ndd = nd * (Pdc / Pa[cell] - Tdc / Tc);
    rates_tangent(ts, rkd, Tc, Tdc, nd, ndd, net, Jc, (NJ && Jd) ? Jd + (size_t)cell * NJ : nullptr, lmc,
                  lmd ? lmd + (size_t)cell * NR : nullptr, rcd, lane);
    const double h0 = isfinite(hin[cell]) ? hin[cell] : dt;
    const double m = fmin(h0, dt);
    const double dtdc = (ad && dtd) ? dtd[cell] : 0.0;
    const double h0d = !ad ? 0.0 : (isfinite(hin[cell]) ? (hd ? hd[cell] : 0.0) : dtdc);
    const double md = h0 < dt ? h0d : (h0 > dt ? dtdc : 0.5 * (h0d + dtdc));  // m's tangent
    int ntry = 0;
    for (int pass = 0;; pass++) {
      const gptr<int> perm = opaque(net)->perm;
      double Nl = lane < K0 ? xr[perm[lane]] : 0.0, Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
      double Ndl = lane < K0 ? xdr[perm[lane]] : 0.0, Ndt = lane < NT ? xdr[perm[K0 + lane]] : 0.0;
      const double split = pow(REC_RETRY_SPLIT, (double)pass), cap = dt * split;
      const double start = pass == 0 ? m : fmin(m, cap);
      const double sd = pass == 0 || cap > m ? md : (cap < m ? dtdc * split : 0.5 * (md + dtdc * split));
      double hdot = !ad ? 0.0 : (start > H_MIN ? sd : (start < H_MIN ? 0.0 : 0.5 * sd));
      double tdot = 0.0, pendot = 0.0, phadot = 0.0;
      Step s = {fmax(start, H_MIN), 0.0, 0.0, 0.0, 0.0, false};
      int it = dt > 0.0 ? 0 : max_steps;
      while (it < max_steps && s.t < dt * CALL_END) {
        const double ha = fmin(s.h, dt - s.t), igh = 1.0 / (GAMMA * ha), ih = 1.0 / ha;
        const double rem = dt - s.t;
        const double had = ad ? (s.h < rem ? hdot : (s.h > rem ? dtdc - tdot : 0.5 * (hdot + dtdc - tdot))) : 0.0;
        double Fl = 0.0, Ft = 0.0, ul = 0.0, ut = 0.0, u0l = 0.0, u0t = 0.0;
        const double ntot = rhs<true>(Nl, Nt, Fl, Ft, u0l, u0t, net, nd, ws, rk, lane);
        jacobian(c, a21, Nl, Nt, u0l, u0t, ntot, ul, ut, net, ws, lane);
        // W = I / (gamma h) - Jn - w 1^T, dense (recover.cuh:124-130)
        to_lds(yv, ul, ut, lane);
        for (int e = lane; e < K0 * NS; e += WAVE) {
          const int i = e / NS;
          W[e] = (e == i * (NS + 1) ? ws[e] + igh : ws[e]) - yv[i];
        }
        if (lane < NT) {
          double* row = W + (K0 + lane) * NS;
#pragma unroll
          for (int k = 0; k < K0; k++) row[k] = a21[k] - ut;
#pragma unroll
          for (int j = 0; j < NT; j++) row[K0 + j] = (j == lane ? c[j] + igh : c[j]) - ut;
        }
        if (lane == 0) *cmd = 1;
        __syncthreads();
        dense_lu(W, pr, dinv, lane, 0);
        double K1l, K1t, K2l, K2t, K3l, K3t, K4l, K4t;
        to_lds(yv, Fl, Ft, lane);
        dense_solve(W, pr, dinv, yv, lane);
        from_lds(yv, K1l, K1t, lane);
        to_lds(yv, fma(RODAS_C21 * ih, K1l, Fl), fma(RODAS_C21 * ih, K1t, Ft), lane);
        dense_solve(W, pr, dinv, yv, lane);
        from_lds(yv, K2l, K2t, lane);
        double Yl = Nl + 2.0 * K1l, Yt = Nt + 2.0 * K1t;
        rhs<false>(Yl, Yt, Fl, Ft, u0l, u0t, net, nd, ws, rk, lane);
        to_lds(yv, fma(ih, K1l - K2l, Fl), fma(ih, K1t - K2t, Ft), lane);
        dense_solve(W, pr, dinv, yv, lane);
        from_lds(yv, K3l, K3t, lane);
        Yl = Nl + 2.0 * K1l + K3l;
        Yt = Nt + 2.0 * K1t + K3t;
        rhs<false>(Yl, Yt, Fl, Ft, u0l, u0t, net, nd, ws, rk, lane);
        to_lds(yv, fma(ih, fma(-RODAS_C43, K3l, K1l - K2l), Fl), fma(ih, fma(-RODAS_C43, K3t, K1t - K2t), Ft), lane);
        dense_solve(W, pr, dinv, yv, lane);
        from_lds(yv, K4l, K4t, lane);
        bool ok = false;
        const bool first = s.t == 0.0;
        const double en = try_error(Nl, Nt, K1l, K1t, K3l, K3t, K4l, K4t, Yl, Yt, first, rtol, atol, lane, ok);
        double end = 0.0;  // the error's tangent (adaptive contract)
        if (ok || ad) {
          // the try's tangent on its dense factors, as the tangent kernel's on the sparse ones
          double Fel, Fet, Gl, Gt, Ql, Qt, dl, dtt;
          double K1dl, K1dt, K2dl, K2dt, K3dl, K3dt, K4dl, K4dt;
          const double sg = had * ih * ih / GAMMA, hq = had * ih * ih;
          rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K1l, K1t, Fel, Fet, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
          K1dl = Fel + Gl;
          K1dt = Fet + Gt;
          if (ad) {
            K1dl = fma(sg, K1l, K1dl);
            K1dt = fma(sg, K1t, K1dt);
          }
          dense_solve_pair(W, pr, dinv, yv, K1dl, K1dt, lane);
          rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K2l, K2t, Ql, Qt, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
          K2dl = fma(RODAS_C21 * ih, K1dl, Fel) + Gl;
          K2dt = fma(RODAS_C21 * ih, K1dt, Fet) + Gt;
          if (ad) {
            K2dl += sg * K2l - hq * RODAS_C21 * K1l;
            K2dt += sg * K2t - hq * RODAS_C21 * K1t;
          }
          dense_solve_pair(W, pr, dinv, yv, K2dl, K2dt, lane);
          const double Y3el = Ndl + 2.0 * K1dl, Y3et = Ndt + 2.0 * K1dt;
          rhs_tangent<false>(Nl + 2.0 * K1l, Nt + 2.0 * K1t, Y3el, Y3et, 0.0, 0.0, Ql, Qt, dl, dtt, net, nd, ndd, ts,
                             rk, rkd, lane);
          rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K3l, K3t, dl, dtt, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
          K3dl = fma(ih, K1dl - K2dl, Ql) + Gl;
          K3dt = fma(ih, K1dt - K2dt, Qt) + Gt;
          if (ad) {
            K3dl += sg * K3l - hq * (K1l - K2l);
            K3dt += sg * K3t - hq * (K1t - K2t);
          }
          dense_solve_pair(W, pr, dinv, yv, K3dl, K3dt, lane);
          rhs_tangent<false>(Nl + 2.0 * K1l + K3l, Nt + 2.0 * K1t + K3t, Y3el + K3dl, Y3et + K3dt, 0.0, 0.0, Ql, Qt,
                             dl, dtt, net, nd, ndd, ts, rk, rkd, lane);
          rhs_tangent<true>(Nl, Nt, Ndl, Ndt, K4l, K4t, dl, dtt, Gl, Gt, net, nd, ndd, ts, rk, rkd, lane);
          K4dl = fma(ih, fma(-RODAS_C43, K3dl, K1dl - K2dl), Ql) + Gl;
          K4dt = fma(ih, fma(-RODAS_C43, K3dt, K1dt - K2dt), Qt) + Gt;
          if (ad) {
            K4dl += sg * K4l - hq * fma(-RODAS_C43, K3l, K1l - K2l);
            K4dt += sg * K4t - hq * fma(-RODAS_C43, K3t, K1t - K2t);
          }
          dense_solve_pair(W, pr, dinv, yv, K4dl, K4dt, lane);
          const double ydl = Ndl + 2.0 * K1dl + K3dl + K4dl, ydt = Ndt + 2.0 * K1dt + K3dt + K4dt;
          if (ad)
            end = error_tangent(Nl, Nt, Yl, Yt, K4l, K4t, Ndl, Ndt, ydl, ydt, K4dl, K4dt, first, en, rtol, atol, lane);
          if (ok) {
            Ndl = lane < K0 && Yl >= 0.0 ? ydl : 0.0;
            Ndt = lane < NT && Yt >= 0.0 ? ydt : 0.0;
          }
        }
        if (ok) {
          Nl = fmax(Yl, 0.0);
          Nt = fmax(Yt, 0.0);
          s.t += ha;
          s.nok += 1.0;
          tdot += had;
        }
        const double fac = step_factor(en, ok, s.prej, s.pen, s.pha, ha);
        if (ad) {
          hdot = fma(had, fac, ha * step_factor_tangent(en, end, ok, s.prej, s.pen, pendot, s.pha, phadot, ha, had));
          pendot = end;
          phadot = had;
        }
        s.h = ha * fac;
        s.prej = !ok;
        s.pen = en;
        s.pha = ha;
        it++;
        ntry++;
      }
      // the guard's test of the result, which goes to the cell's output row (a later pass overwrites it)
      const int dn = s.t >= dt * CALL_END ? 1 : 0;
      const int code = __builtin_amdgcn_readfirstlane(
          guard_wave(Nl, Nt, xr, xout + (size_t)cell * NS, dn != 0, s.nok, atol, net, lane));
      if (code == GUARD_DRIFT && pass < REC_RETRIES) continue;  // only drifted: again, with a smaller first try
      double* xdo = xd + (size_t)cell * NS;
      if (code == 0) {
        guard_tangent(Nl, Nt, Ndl, Ndt, xr, xdr, xdo, net, lane);
      } else {
        for (int i = lane; i < NS; i += WAVE) xdo[i] = DNAN;
      }
      int nx = 0;
      if (lane == 0) {
        hio[cell] = s.h;
        if (hdo) hdo[cell] = !ad ? 0.0 : (code == 0 ? hdot : DNAN);
        doneo[cell] = code ? 0 : dn;  // done: finished and passing the guard
        noko[cell] = s.nok;
        redo[cell] = code ? 2 : 1;
        if (tries) tries[cell] += (double)ntry;
        atomicAdd(count + (code ? COUNT_STILL_FAILED : COUNT_RECOVERED), 1);
        nx = (int)gridDim.x + atomicAdd(count + COUNT_RECOVERY_QUEUE, 1);
      }
      q = __builtin_amdgcn_readfirstlane(nx);
      break;
    }
  }
  if (lane == 0) *cmd = 0;
  __syncthreads();  // the other wavefronts leave
}

// The rate table's cotangent grc [NR][9] (its nine Arrhenius columns) from the cells' cotangents of their log rate
// multipliers, glm [B][NR]. A multiplier and the law enter a reaction's two coefficients alike: d ln kf = d ln kr =
// (1 - wA) d ln k + d lm, with k the law before the collision limit and wA the limit's share (1 above the limit, 1/2
// at it, 0 below: the limit's factor KCOLL / q moves with k). So grc[r][j] = sum over the cells of
// (1 - wA) (d ln k / d rc[r][j]) glm[cell][r], each lane's cells in index order and the lanes' shares by a fixed-order
// wave sum, the same bits on every run. Cells with skip[cell] != 0 (the guard's still-failed ones, whose cotangents
// are NaN) are left out. One wavefront per reaction; a photolysis branch has no law (0).
extern "C" __global__ void __launch_bounds__(WAVE)
    mxc_rc_cotangent(NetP __restrict__ net, const double* __restrict__ Ta, const double* __restrict__ Pa,
                     const double* __restrict__ glm, const int* __restrict__ skip, int B, double* __restrict__ grc) {
  const int r = (int)blockIdx.x, lane = (int)threadIdx.x;
  if (r >= NR) return;
  const NetP nt = opaque(net);
  const int fl = nt->rf[r];
  double acc[9];
#pragma unroll
  for (int j = 0; j < 9; j++) acc[j] = 0.0;
  if (!(fl & FL_PHOTO)) {
    const gptr<double> q = nt->rc + RCW * r;
    const u64 w = nt->rxe[r];
    const int mf = (fl >> FL_NF_SHIFT) & 3, mr = (fl >> FL_NR_SHIFT) & 3;
    const bool tb = (fl & FL_THREE_BODY) != 0;
    for (int cell = lane; cell < B; cell += WAVE) {
      const double gl = glm[(size_t)cell * NR + r];
      if ((skip && skip[cell] != 0) || gl == 0.0) continue;
      const double T = Ta[cell], nd = Pa[cell] / (KBOLTZ * T);
      const double lT = log(T), iT = 1.0 / T, lKT = log(KBOLTZ * T / P0BAR);
      const gptr<double> cf = nt->nasa + (T < NASA_SPLIT_T ? 0 : NS * 9);
      double dg = 0.0;
#pragma unroll
      for (int m = 0; m < 6; m++) {
        const int sp = species_of(w, m);
        const double gs = sp < NS ? gibbs_rt(cf + 9 * sp, T, lT, iT) : 0.0;
        dg += m < 3 ? -gs : gs;
      }
      const double k = law(q, fl, T, lT, iT, nd), c3 = tb ? nd : 1.0;
      const double f = k * c3, b = (k * exp(dg + (double)(mr - mf) * lKT)) * c3;
      const double qb = fmax(bimolecular(mf, tb, f, nd), bimolecular(mr, tb, b, nd));
      const double share = qb > KCOLL ? 0.0 : (qb == KCOLL ? 0.5 : 1.0);  // 1 - wA
      if (share == 0.0 || !(k > 0.0)) continue;  // no law derivative left (k = 0: no rate to scale)
      const Dual Tq{T, 0.0}, lTq{lT, 0.0}, iTq{iT, 0.0}, ndq{nd, 0.0};
#pragma unroll
      for (int j = 0; j < 9; j++) {
        double unit[9];
#pragma unroll
        for (int i = 0; i < 9; i++) unit[i] = i == j ? 1.0 : 0.0;
        const Dual kd = law_dual(q, unit, fl, Tq, lTq, iTq, ndq);
        acc[j] += share * (kd.d / kd.v) * gl;
      }
    }
  }
  wave_sums(acc);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 9; j++) grc[9 * r + j] = acc[j];
  }
}
