// The call kernel: a persistent grid of wavefronts, each stepping one cell at a time from an atomic queue through
// its rates, Rodas3 tries, step control and the guard of its result (the reference's call kernel,
// maxichem/kernels/call.cuh:5-178, and its guard kernel, guard.cuh:114-140).

// One try's four stages from the factors of M0 (call.cuh:119-158): K1 = W^-1 F, K2 = W^-1 (F + 4 K1 / h),
// K3 = W^-1 (F(N + 2 K1) + (K1 - K2) / h), K4 = W^-1 (F(N + 2 K1 + K3) + (K1 - K2 - 8/3 K3) / h), with
// W^-1 y = M0^-1 y + c0 (1^T M0^-1 y) M0^-1 w by Sherman-Morrison.

// M0's factorization and its solves, by the trailing block's type: FP64; or, in the mixed form, FP32 with each solve
// refined in FP64 (false when the refinement did not converge).
__device__ __forceinline__ void factor_m0(double (&c)[NT], LeadRow& a21, double& dl, double& dtr, double* ws,
                                          int lane) {
  factor(c, a21, dl, dtr, ws, lane);
}

template <int NV>
__device__ __forceinline__ bool solve_m0(double (&yl)[NV], double (&yt)[NV], const double (&c)[NT],
                                         const LeadRow& a21, double dl, double dtr, double, double, double,
                                         NetP, double* ws, int lane) {
  solve<NV>(yl, yt, c, a21, dl, dtr, ws, lane);
  return true;
}

#if MIXED
__device__ __forceinline__ void factor_m0(float (&c)[NT], LeadRow& a21, double& dl, float& dtr, double* ws, int lane) {
  factor_lead(c, a21, dl, ws, lane);
  factor_trailing32(c, dtr, ws, lane);
}

template <int NV>
__device__ __forceinline__ bool solve_m0(double (&yl)[NV], double (&yt)[NV], const float (&c)[NT], const LeadRow& a21,
                                         double dl, float dtr, double igh, double wl, double wt, NetP net,
                                         double* ws, int lane) {
  return msolve<NV>(yl, yt, c, a21, dl, dtr, igh, wl, wt, net, ws, lane);
}
#endif

// What a try's stages leave besides K1, K3 and K4, for a tangent that reuses its factors: K2, M0^-1 w (lead / trailing
// parts), the Sherman-Morrison factor c0 and the factors' 1 / U_ii of this lane's lead row.
struct StageKeep {
  double K2l, K2t, Ul, Ut, c0, dl;
};

// One try's stages from the state N at the step ha: F, w and -Jn at N, M0 = I / (gamma ha) - Jn factored, then K1 .. K4
// (the lead / trailing parts). Returns false when a solve's FP32 refinement did not converge (the mixed form only).
// keep, when not null, receives StageKeep's values (the arithmetic is the same either way).
template <typename T, typename DT = T>
__device__ __forceinline__ bool stages(T (&c)[NT], LeadRow& a21, double Nl, double Nt, double ha, double rtol,
                                       double atol, NetP net, double nd, double* ws, const double* rk,
                                       int lane, double& K1l, double& K1t, double& K3l, double& K3t, double& K4l,
                                       double& K4t, StageKeep* keep = nullptr, DT* dtr_keep = nullptr) {
  const double igh = 1.0 / (GAMMA * ha), ih = 1.0 / ha;
  double Fl = 0.0, Ft = 0.0, ul = 0.0, ut = 0.0, u0l = 0.0, u0t = 0.0;
  const double ntot = rhs<true, WORDS_LDS != 0>(Nl, Nt, Fl, Ft, u0l, u0t, net, nd, ws, rk, lane);
  jacobian<WORDS_LDS != 0>(c, a21, Nl, Nt, u0l, u0t, ntot, ul, ut, net, ws, lane);
  add_diagonal(c, igh, ws, lane);
  double dl;
  T dtr;
  factor_m0(c, a21, dl, dtr, ws, lane);
  const double wl = atol + rtol * fabs(Nl), wt = atol + rtol * fabs(Nt);  // the refinement's error weights
#define MXC_SOLVE(YL, YT) \
  if (!solve_m0(YL, YT, c, a21, dl, dtr, igh, wl, wt, net, ws, lane)) return false
  double y2l[2] = {Fl, ul}, y2t[2] = {Ft, ut};
  MXC_SOLVE(y2l, y2t);
  const double Ul = y2l[1], Ut = y2t[1];  // M0^-1 w
  const double c0 = 1.0 / (1.0 - vsum(Ul, Ut, lane));
  K1l = y2l[0];
  K1t = y2t[0];
  double sk = vsum(K1l, K1t, lane) * c0;
  K1l += sk * Ul;
  K1t += sk * Ut;
  double y1l[1] = {fma(RODAS_C21 * ih, K1l, Fl)}, y1t[1] = {fma(RODAS_C21 * ih, K1t, Ft)};
  MXC_SOLVE(y1l, y1t);
  sk = vsum(y1l[0], y1t[0], lane) * c0;
  const double K2l = y1l[0] + sk * Ul, K2t = y1t[0] + sk * Ut;
  double Yl = Nl + 2.0 * K1l, Yt = Nt + 2.0 * K1t;
  rhs<false, WORDS_LDS != 0>(Yl, Yt, Fl, Ft, u0l, u0t, net, nd, ws, rk, lane);
  y1l[0] = fma(ih, K1l - K2l, Fl);
  y1t[0] = fma(ih, K1t - K2t, Ft);
  MXC_SOLVE(y1l, y1t);
  sk = vsum(y1l[0], y1t[0], lane) * c0;
  K3l = y1l[0] + sk * Ul;
  K3t = y1t[0] + sk * Ut;
  Yl = Nl + 2.0 * K1l + K3l;
  Yt = Nt + 2.0 * K1t + K3t;
  rhs<false, WORDS_LDS != 0>(Yl, Yt, Fl, Ft, u0l, u0t, net, nd, ws, rk, lane);
  y1l[0] = fma(ih, fma(-RODAS_C43, K3l, K1l - K2l), Fl);
  y1t[0] = fma(ih, fma(-RODAS_C43, K3t, K1t - K2t), Ft);
  MXC_SOLVE(y1l, y1t);
#undef MXC_SOLVE
  sk = vsum(y1l[0], y1t[0], lane) * c0;
  K4l = y1l[0] + sk * Ul;
  K4t = y1t[0] + sk * Ut;
  if (keep) *keep = {K2l, K2t, Ul, Ut, c0, dl};
  if (dtr_keep) *dtr_keep = dtr;
  return true;
}

extern "C" __global__ void __launch_bounds__(WAVE* CPW) __attribute__((amdgpu_waves_per_eu(CALL_WAVES_PER_SIMD)))
    mxc_call(NetP __restrict__ net, const double* __restrict__ x0, const double* __restrict__ Ta,
             const double* __restrict__ Pa, const double* __restrict__ dta, double* __restrict__ hio,
             double* __restrict__ xout, int* __restrict__ doneo, double* __restrict__ noko,
             int* __restrict__ status, int* __restrict__ redo, int* __restrict__ list, int* __restrict__ count,
             double* __restrict__ tries, const int* __restrict__ order, const double* __restrict__ Jin, int B,
             double rtol, double atol, int max_steps, const double* __restrict__ xprop,
             double* __restrict__ rscratch, int* __restrict__ fp64) {
  extern __shared__ double lds[];
  const int lane = (int)(threadIdx.x & (WAVE - 1)), wv = (int)(threadIdx.x / WAVE);
  double* ws = lds + wv * L_SIZE;
#if RATES_GLOBAL
  double* rk = rscratch + (size_t)((int)blockIdx.x * CPW + wv) * (2 * NR);  // this wave's slot
#else
  double* rk = ws + L_KF;
#endif
  const int waves = (int)gridDim.x * CPW;
  int q = (int)blockIdx.x * CPW + wv;
#if WORDS_LDS
  {  // the stream words into the workgroup's LDS, after its cells' regions
    const NetP nt = opaque(net);
    u32* lw = (u32*)(lds + CPW * L_SIZE);
    for (int i = (int)threadIdx.x; i < SV_ROUNDS * WAVE; i += WAVE * CPW) lw[W_SV + i] = nt->svw[i];
    for (int i = (int)threadIdx.x; i < DIAG_ROUNDS * WAVE; i += WAVE * CPW) lw[W_DIAG + i] = nt->diagw[i];
    for (int i = (int)threadIdx.x; i < LEAD_ROUNDS * WAVE; i += WAVE * CPW) lw[W_LEAD + i] = nt->leadw[i];
    for (int i = (int)threadIdx.x; i < WORDS_TOTAL - W_CH; i += WAVE * CPW) lw[W_CH + i] = nt->chw[i];
    __syncthreads();
  }
#endif
  double c[NT];  // the trailing block of M0 and its factors (in the mixed form, once a cell's FP32 refinement failed)
  LeadRow a21;  // the trailing row's lead columns of M0, then of L
  while (q < B) {
    const int cell = order ? order[q] : q;
    const double Tc = Ta[cell], dt = dta[cell], nd = Pa[cell] / (KBOLTZ * Tc);
    const double* xr = x0 + (size_t)cell * NS;
    const gptr<int> perm = opaque(net)->perm;
    double Nl = lane < K0 ? xr[perm[lane]] : 0.0, Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
    // the screen (common.cuh:244-279): a NaN enters the sum for a negative or non-finite x or J
    double p = (lane < K0 ? (Nl >= 0.0 ? Nl : DNAN) : 0.0) + (lane < NT ? (Nt >= 0.0 ? Nt : DNAN) : 0.0);
    double pm = fmax(lane < K0 ? Nl : 0.0, lane < NT ? Nt : 0.0);
    const double* Jc = NJ ? Jin + (size_t)cell * NJ : nullptr;
    for (int b = lane; b < NJ; b += WAVE) p += Jc[b] >= 0.0 ? 0.0 * Jc[b] : DNAN;
    const double ps = wave_sum(p), pmax = wave_max(pm);
    const bool fine = fabs(ps - 1.0) <= XSUMTOL && pmax > 0.0 && isfinite(Tc) && isfinite(nd) && nd > 0.0 &&
                      isfinite(dt) && dt >= 0.0;
    const int st = __builtin_amdgcn_readfirstlane(
        Tc < TMIN ? REFUSED_COLD : (Tc > TMAX ? REFUSED_HOT : (fine ? 0 : REFUSED_BAD_INPUT)));
    if (lane == 0) status[cell] = st;
    if (st == 0) cell_rates(ws, rk, Tc, nd, net, Jc, lane);
    const double hin = hio[cell];
    const double h0 = isfinite(hin) ? hin : dt;  // a non-finite h acts as dt, and is counted
    if (lane == 0 && !isfinite(hin)) atomicAdd(count + COUNT_NONFINITE_H, 1);
    Step s = {fmax(fmin(h0, dt), H_MIN), 0.0, 0.0, 0.0, 0.0, false};
    int it = (dt > 0.0 && st == 0) ? 0 : max_steps;  // a refused cell takes no try
    int ntry = 0;
    // a proposed end state (a row whose first entry is not NaN) is tried first: one try of the whole dt from it
    bool verifying = false;
    if (xprop != nullptr && it == 0) {
      const double* xq = xprop + (size_t)cell * NS;
      if (!isnan(xq[0])) {
        verifying = true;
        Nl = lane < K0 ? xq[perm[lane]] : 0.0;
        Nt = lane < NT ? xq[perm[K0 + lane]] : 0.0;
      }
    }
    // the mixed form starts a cell's call in FP32, unless fp64 (per cell; may be null) says that an earlier call fell
    // back for it; a refinement that does not converge puts the cell on FP64, and marks it there
    [[maybe_unused]] bool fp32 = !(fp64 != nullptr && fp64[cell] != 0);
    while (it < max_steps && s.t < dt * CALL_END) {
      const double ha = verifying ? dt : fmin(s.h, dt - s.t);
      double K1l, K1t, K3l, K3t, K4l, K4t;
#if MIXED
      if (fp32) {
        float cf[NT];  // the trailing block in FP32
        if (!stages(cf, a21, Nl, Nt, ha, rtol, atol, net, nd, ws, rk, lane, K1l, K1t, K3l, K3t, K4l, K4t)) {
          fp32 = false;  // this try again in FP64, and the cell's later tries of the call
          if (lane == 0 && fp64 != nullptr) fp64[cell] = 1;
          continue;
        }
      } else {
        stages(c, a21, Nl, Nt, ha, rtol, atol, net, nd, ws, rk, lane, K1l, K1t, K3l, K3t, K4l, K4t);
      }
#else
      stages(c, a21, Nl, Nt, ha, rtol, atol, net, nd, ws, rk, lane, K1l, K1t, K3l, K3t, K4l, K4t);
#endif
      bool ok = false;
      double Yl, Yt;
      const double en = try_error(Nl, Nt, K1l, K1t, K3l, K3t, K4l, K4t, Yl, Yt, s.t == 0.0, rtol, atol, lane, ok);
      if (verifying) {
        // the proposal is kept when the try passes and moves it by no more than the tolerance: the end state is
        // then stationary over dt; otherwise the cell starts again from its input
        const double ml = lane < K0 ? fabs(Yl - Nl) / (atol + rtol * fmax(fabs(Nl), fabs(Yl))) : 0.0;
        const double mt = lane < NT ? fabs(Yt - Nt) / (atol + rtol * fmax(fabs(Nt), fabs(Yt))) : 0.0;
        const bool kept = ok && wave_max(fmax(ml, mt)) <= 1.0;
        verifying = false;
        it++;
        ntry++;
        if (kept) {
          Nl = fmax(Yl, 0.0);
          Nt = fmax(Yt, 0.0);
          s.t = dt;
          s.nok = 1.0;
          s.h = ha * step_factor(en, true, false, 0.0, 0.0, ha);
        } else {
          Nl = lane < K0 ? xr[perm[lane]] : 0.0;
          Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
        }
        continue;
      }
      if (ok) {  // accepted: the new state clamped at 0 (common.cuh:111-118)
        Nl = fmax(Yl, 0.0);
        Nt = fmax(Yt, 0.0);
        s.t += ha;
        s.nok += 1.0;
      }
      const double fac = step_factor(en, ok, s.prej, s.pen, s.pha, ha);
      s.h = ha * fac;
      s.prej = !ok;
      s.pen = en;
      s.pha = ha;
      it++;
      ntry++;
    }
    // the result, and the guard while the state is in the lanes: a refused cell is counted by its status (and is
    // not done, also at dt = 0); a stepped one whose result fails the guard is counted by what failed and listed
    // for the recovery kernel
    const bool fin = st == 0 && s.t >= dt * CALL_END;
    double* xo = xout + (size_t)cell * NS;
    int code = 0;
    if (st == 0)
      code = guard_wave(Nl, Nt, xr, xo, fin, s.nok, atol, net, lane);
    else
      write_state(Nl, Nt, xo, net, lane);
    if (lane == 0) {
      hio[cell] = s.h;
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
