// The mixed-precision form: the trailing block of M0 = I / (gamma h) - Jn is factored and its triangular solves run in
// FP32; each solve of M0 y = b is then refined with residuals r = b - M0 y taken in FP64 without the matrix, from the
// dv terms of the try (M0 y = y / (gamma h) - S (dv/dn y)), at most REFINE_SWEEPS times. A refinement that does not
// converge sends the try, and the cell's later tries of the call, to the FP64 factorization (the call kernel's
// fallback, as LAPACK's dsgesv). The rates, the right-hand sides, the lead rows of the factorization (LDS), the error
// estimate and the guard stay FP64.

constexpr int REFINE_SWEEPS = 3;
constexpr double REFINE_TOL = 1e-4;  // a sweep's correction below this many error weights ends the refinement

__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ void pin_f(float& a, float& b) { asm volatile("" : "+v"(a), "+v"(b)); }

// The trailing block's LU in registers, FP32 (as factor_trailing). dt: 1 / U_ii of trailing row lane.
__device__ __forceinline__ void factor_trailing32(float (&c)[NT], float& dt, double* ws, int lane) {
  float* dinv = (float*)(ws + L_UNION);
#pragma unroll
  for (int k = 0; k < NT; k++) {
    const float dk = 1.0f / readlane_f(c[k], k);
    if (lane == 0) dinv[k] = dk;
    if (lane != k) c[k] *= dk;
    const float l = lane > k ? c[k] : 0.0f;
#pragma unroll
    for (int j = k + 1; j < NT; j++)
      if ((TRAIL_FILL[k] >> j) & 1ull) c[j] = fmaf(-l, readlane_f(c[j], k), c[j]);  // structural zeros left out
  }
  wave_sync();
  dt = lane < NT ? dinv[lane] : 1.0f;
  wave_sync();
}

// The trailing block's solves in FP32 (as solve_trailing), from and to FP64 vectors.
template <int NV>
__device__ __forceinline__ void solve_trailing32(double (&ytd)[NV], const float (&c)[NT], float dt, int lane) {
  float yt[NV], rec[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) yt[v] = (float)ytd[v];
#pragma unroll
  for (int k = 0; k < NT; k++) {
    const bool own = lane_after(lane, (double)yt[0]) == k;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const float yk = readlane_f(yt[v], k);
      rec[v] = own ? yt[v] : (k == 0 ? 0.0f : rec[v]);
      if (k + 1 < NT) yt[v] = fmaf(-c[k], yk, yt[v]);
      pin_f(rec[v], yt[v]);
    }
  }
#pragma unroll
  for (int k = NT - 1; k >= 0; k--) {
    const bool own = lane_after(lane, (double)rec[0]) == k;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const float zk = readlane_f(rec[v], k);
      yt[v] = own ? rec[v] : (k == NT - 1 ? 0.0f : yt[v]);
      if (k > 0) rec[v] = fmaf(-c[k], zk, rec[v]);
      pin_f(rec[v], yt[v]);
    }
  }
#pragma unroll
  for (int v = 0; v < NV; v++) ytd[v] = (double)(yt[v] * dt);
}

// y = M0 x for NV vectors (lead / trailing parts, elimination order), FP64: x / (gamma h) - S (dv/dn x), with the dv
// terms of the try's start (DV) and the S v stream; the exchange vectors XV hold x by species.
template <int NV>
__device__ __forceinline__ void m0_times(const double (&xl)[NV], const double (&xt)[NV], double (&yl)[NV],
                                         double (&yt)[NV], double igh, NetP net, double* ws, int lane) {
  double* XV = ws + L_XV;
  wave_sync();
#pragma unroll
  for (int v = 0; v < NV; v++) {
    if (lane < K0) XV[v * (NS + 1) + lane] = xl[v];
    if (lane < NT) XV[v * (NS + 1) + K0 + lane] = xt[v];
    if (lane == 0) XV[v * (NS + 1) + NS] = 0.0;
  }
  net = opaque(net);
  const gptr<u64> rxe = net->rxe;
  const gptr<int> rf = net->rf;
  const gptr<u32> svw = net->svw;
  const gptr<u32> rowp = net->rowp;
  const double* DV = ws + L_DV;
  double* V = ws + L_V;
  double* part = ws + L_PART;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double* x = XV + v * (NS + 1);
    wave_sync();
#pragma unroll
    for (int q = 0; q < RROUNDS; q++) {  // dv_r/dn . x per reaction
      const int r = q * WAVE + lane;
      if (r >= NR) break;
      const u64 w = rxe[r];
      const int fl = rf[r];
      double d = 0.0;
      if (!(fl & FL_INERT)) {
        int o = (int)(w >> 48);
        const int mf = (fl >> FL_NF_SHIFT) & 3, mr = (fl >> FL_NR_SHIFT) & 3;
        if (mf >= 1) d = fma(DV[o], x[species_of(w, 0)], d);
        if (mf >= 2) d = fma(DV[o + 1], x[species_of(w, 1)], d);
        if (mf >= 3) d = fma(DV[o + 2], x[species_of(w, 2)], d);
        if (!(fl & FL_PHOTO)) {
          o += mf;
          if (mr >= 1) d = fma(-DV[o], x[species_of(w, 3)], d);
          if (mr >= 2) d = fma(-DV[o + 1], x[species_of(w, 4)], d);
          if (mr >= 3) d = fma(-DV[o + 2], x[species_of(w, 5)], d);
        }
      }
      V[r] = d;
    }
    if (lane == 0) V[NR] = 0.0;
    wave_sync();
    double acc = 0.0;
#pragma unroll 1
    for (int q0 = 0; q0 < SV_ROUNDS; q0 += STREAM_BATCH) {  // S (dv/dn x), as the right-hand side's S v
      u32 wd[STREAM_BATCH];
      double dd[STREAM_BATCH];
#pragma unroll
      for (int i = 0; i < STREAM_BATCH; i++) wd[i] = q0 + i < SV_ROUNDS ? svw[(q0 + i) * WAVE + lane] : (u32)NR;
#pragma unroll
      for (int i = 0; i < STREAM_BATCH; i++) dd[i] = V[wd[i] & SV_R_MASK];
#pragma unroll
      for (int i = 0; i < STREAM_BATCH; i++) {
        acc += term<SV_DBL>(wd[i], dd[i]);
        if ((wd[i] >> SV_END) & 1u) {
          part[(wd[i] >> SV_PIECE) & SV_PIECE_MASK] = acc;
          acc = 0.0;
        }
      }
    }
    wave_sync();
    double sl = 0.0, st = 0.0;
    if (lane < K0) {
      const u32 rp = rowp[lane];
      for (int e = 0; e < (int)(rp >> 16); e++) sl += part[(rp & 0xffffu) + e];
    }
    if (lane < NT) {
      const u32 rp = rowp[K0 + lane];
      for (int e = 0; e < (int)(rp >> 16); e++) st += part[(rp & 0xffffu) + e];
    }
    yl[v] = lane < K0 ? fma(igh, xl[v], -sl) : 0.0;
    yt[v] = lane < NT ? fma(igh, xt[v], -st) : 0.0;
  }
  wave_sync();
}

// M0^-1 b for NV vectors with the FP32 trailing factors and FP64 refinement; wl, wt: the error weights
// atol + rtol |N| of the try's state. b comes in yl, yt and the solution goes out there. Returns whether the last
// sweep's correction fell below REFINE_TOL weights: where M0 is too ill-conditioned for FP32 (large steps of stiff
// chemistry) the refinement stalls or diverges, and the try then goes to FP64.
template <int NV>
__device__ __forceinline__ bool msolve(double (&yl)[NV], double (&yt)[NV], const float (&c)[NT], const LeadRow& a21,
                                       double dl, float dt, double igh, double wl, double wt, NetP net,
                                       double* ws, int lane) {
  double bl[NV], bt[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) bl[v] = yl[v], bt[v] = yt[v];
  solve_lead_forward(yl, yt, a21, ws, lane);
  solve_trailing32(yt, c, dt, lane);
  solve_lead_backward(yl, yt, dl, ws, lane);
#pragma unroll 1
  for (int sweep = 0; sweep < REFINE_SWEEPS; sweep++) {
    double rl[NV], rt[NV];
    m0_times(yl, yt, rl, rt, igh, net, ws, lane);
#pragma unroll
    for (int v = 0; v < NV; v++) rl[v] = bl[v] - rl[v], rt[v] = bt[v] - rt[v];
    solve_lead_forward(rl, rt, a21, ws, lane);
    solve_trailing32(rt, c, dt, lane);
    solve_lead_backward(rl, rt, dl, ws, lane);
    // converged when every lane's correction is within REFINE_TOL weights; a NaN or infinite correction (a broken
    // FP32 factor) fails the comparison, so the try goes to FP64
    bool fine = true;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      yl[v] += rl[v];
      yt[v] += rt[v];
      fine = fine && (lane >= K0 || fabs(rl[v]) / wl <= REFINE_TOL) && (lane >= NT || fabs(rt[v]) / wt <= REFINE_TOL);
    }
    if (wave_all(fine)) return true;
  }
  return false;
}
