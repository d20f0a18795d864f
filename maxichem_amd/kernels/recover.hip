// The recovery kernel: recomputes the cells the guard listed, one wavefront per cell, from their inputs, with the
// call kernel's rates, rhs, Jacobian, stages and step control, but the dense W = I / (gamma h) - Jn - w 1^T factored
// by LU with threshold partial pivoting (the reference's recover kernel, maxichem/kernels/recover.cuh:3-310).

// Dense LU of W (NS x NS, row-major, LDS) in place, shared by the workgroup's REC_WAVES wavefronts (wv: this one's
// index): row k is exchanged with the row of the largest |entry| of column k (ties to the lower row; a NaN is never
// chosen) only when |W_kk| < PIVTAU times it or is not finite. pr[i]: the original row now in row i; dinv[i] =
// 1 / U_ii. Every wavefront reads column k once (rows k + 1 + l (+ WAVE) in lane l's registers, W_kk in all lanes)
// and finds the pivot itself from it, so all take the same branches and meet at the same barriers: the largest
// entry from a wave reduction of the lanes' own largest, then its lowest row from a second one. The pivot and the
// multipliers come from those registers. Wavefront wv updates the blocks of REC_BLOCK columns b = wv, wv +
// REC_WAVES, ... from column k + 1, so every entry takes the same operations in the same order as with one
// wavefront. Wavefront 0 writes column k's multipliers at the start of column k + 1: no other wavefront reads
// column k again, and the next exchange, which moves them, comes after a barrier.
constexpr int REC_BLOCK = 8;
constexpr int REC_ROWS = (NS + WAVE - 2) / WAVE;  // rows below a pivot per lane

__device__ void dense_lu(double* W, int* pr, double* dinv, int lane, int wv) {
  if (wv == 0)
    for (int i = lane; i < NS; i += WAVE) pr[i] = i;
  __syncthreads();
  double l[REC_ROWS];  // column k's multipliers, rows k + 1 + lane + WAVE p
  for (int k = 0; k < NS; k++) {
    if (wv == 0 && k > 0) {
#pragma unroll
      for (int p = 0; p < REC_ROWS; p++) {
        const int i = k + lane + WAVE * p;  // rows below pivot k - 1
        if (i < NS) W[i * NS + k - 1] = l[p];
      }
    }
    double v[REC_ROWS];
#pragma unroll
    for (int p = 0; p < REC_ROWS; p++) {
      const int i = k + 1 + lane + WAVE * p;
      v[p] = i < NS ? W[i * NS + k] : 0.0;
    }
    const double akk = W[k * NS + k];
    double best = -1.0;
    int bi = NS;
    if (lane == 0 && fabs(akk) > best) {  // row k first: each lane sees its rows in increasing order
      best = fabs(akk);
      bi = k;
    }
#pragma unroll
    for (int p = 0; p < REC_ROWS; p++) {
      const int i = k + 1 + lane + WAVE * p;
      if (i < NS && fabs(v[p]) > best) {
        best = fabs(v[p]);
        bi = i;
      }
    }
    const double top = wave_max(best);
    bi = wave_min(best == top ? bi : NS);
    const bool exchange = bi != NS && bi != k && !(fabs(akk) >= PIVTAU * top);
    double piv = akk;
    if (exchange) {
      __syncthreads();  // every wavefront has read column k
      for (int j = lane + WAVE * wv; j < NS; j += WAVE * REC_WAVES) {
        const double a = W[k * NS + j];
        W[k * NS + j] = W[bi * NS + j];
        W[bi * NS + j] = a;
      }
      if (wv == 0 && lane == 0) {
        const int t = pr[k];
        pr[k] = pr[bi];
        pr[bi] = t;
      }
      __syncthreads();
      // row bi's entry is the pivot now, and row bi holds row k's
      const int ob = bi - k - 1, lb = ob % WAVE, pb = ob / WAVE;
      double vb = v[0];
#pragma unroll
      for (int p = 1; p < REC_ROWS; p++) vb = pb == p ? v[p] : vb;
      piv = readlane(vb, lb);
#pragma unroll
      for (int p = 0; p < REC_ROWS; p++) v[p] = (p == pb && lane == lb) ? akk : v[p];
    }
    const double dk = 1.0 / piv;
    if (wv == 0 && lane == 0) dinv[k] = dk;
#pragma unroll
    for (int p = 0; p < REC_ROWS; p++) {
      const int i = k + 1 + lane + WAVE * p;
      l[p] = i < NS ? v[p] * dk : 0.0;
      if (i < NS) {
        for (int j0 = k + 1 + REC_BLOCK * wv; j0 < NS; j0 += REC_BLOCK * REC_WAVES) {
          if (j0 + REC_BLOCK <= NS) {  // a whole block: its loads go out together
            double u[REC_BLOCK], w[REC_BLOCK];
#pragma unroll
            for (int q = 0; q < REC_BLOCK; q++) {
              u[q] = W[k * NS + j0 + q];
              w[q] = W[i * NS + j0 + q];
            }
#pragma unroll
            for (int q = 0; q < REC_BLOCK; q++) W[i * NS + j0 + q] = fma(-l[p], u[q], w[q]);
          } else {
            for (int j = j0; j < NS; j++) W[i * NS + j] = fma(-l[p], W[k * NS + j], W[i * NS + j]);
          }
        }
      }
    }
    __syncthreads();  // the updates are in before the next column is read
  }
}

// y <- W^-1 y with dense_lu's factors; y is the vector in LDS (elimination order). Up to ROW_REGISTERS species, this
// lane's row of the factors (row lane, clamped in bounds) is loaded into registers before the substitutions, so the
// chain from step to step is a readlane and an fma with no LDS wait (beyond that the row would not fit the
// registers, and its entries are read in the steps). Rows WAVE and beyond (a second row per lane) read their entries
// in the steps. Every update is kept by a select, as the substitutions' order fixes the bits.
constexpr int ROW_REGISTERS = 80;
constexpr bool ROW_IN_REGISTERS = NS <= ROW_REGISTERS;

__device__ void dense_solve(const double* W, const int* pr, const double* dinv, double* y, int lane) {
  double a[2];
#pragma unroll
  for (int p = 0; p < 2; p++) a[p] = lane + WAVE * p < NS ? y[pr[lane + WAVE * p]] : 0.0;
  const int r0 = lane < NS ? lane : NS - 1, r1 = lane + WAVE < NS ? lane + WAVE : NS - 1;  // rows in bounds
  double w0[NS];
  if constexpr (ROW_IN_REGISTERS) {
#pragma unroll
    for (int k = 0; k < NS; k++) w0[k] = W[r0 * NS + k];
  }
  const double d0 = dinv[r0], d1 = dinv[r1];
#pragma unroll
  for (int k = 0; k < NS - 1; k++) {
    const double yk = readlane(k < WAVE ? a[0] : a[1], k & (WAVE - 1));
    a[0] = lane > k ? fma(-(ROW_IN_REGISTERS ? w0[k] : W[r0 * NS + k]), yk, a[0]) : a[0];
    if (NS > WAVE) a[1] = lane + WAVE > k && lane + WAVE < NS ? fma(-W[r1 * NS + k], yk, a[1]) : a[1];
  }
#pragma unroll
  for (int k = NS - 1; k > 0; k--) {
    const double zk = readlane(k < WAVE ? a[0] : a[1], k & (WAVE - 1)) * readlane(k < WAVE ? d0 : d1, k & (WAVE - 1));
    a[0] = lane < k ? fma(-(ROW_IN_REGISTERS ? w0[k] : W[r0 * NS + k]), zk, a[0]) : a[0];
    if (NS > WAVE) a[1] = lane + WAVE < k ? fma(-W[r1 * NS + k], zk, a[1]) : a[1];
  }
  wave_sync();
  if (lane < NS) y[lane] = a[0] * d0;
  if (lane + WAVE < NS) y[lane + WAVE] = a[1] * d1;
  wave_sync();
}

// A lead / trailing vector pair into the LDS vector y (elimination order) and back.
__device__ __forceinline__ void to_lds(double* y, double vl, double vt, int lane) {
  if (lane < K0) y[lane] = vl;
  if (lane < NT) y[K0 + lane] = vt;
  wave_sync();
}

__device__ __forceinline__ void from_lds(const double* y, double& vl, double& vt, int lane) {
  vl = lane < K0 ? y[lane] : 0.0;
  vt = lane < NT ? y[K0 + lane] : 0.0;
  wave_sync();
}

// A workgroup of REC_WAVES wavefronts per listed cell: wavefront 0 steps the cell; the others wait at a workgroup
// barrier for each of its dense factorizations (cmd 1: factor W, 0: the workgroup's cells are done), so every
// wavefront passes the same barriers.
extern "C" __global__ void __launch_bounds__(WAVE* REC_WAVES)
    mxc_recover(NetP __restrict__ net, const double* __restrict__ x0, const double* __restrict__ Ta,
                const double* __restrict__ Pa, const double* __restrict__ dta, const double* __restrict__ hin,
                double* __restrict__ xout, double* __restrict__ hio, int* __restrict__ doneo,
                double* __restrict__ noko, int* __restrict__ redo, const int* __restrict__ list,
                int* __restrict__ count, double* __restrict__ tries, const double* __restrict__ Jin, double rtol,
                double atol, int max_steps, double* __restrict__ rscratch) {
  const int nfl = count[COUNT_RECOMPUTED];
  if ((int)blockIdx.x >= nfl) return;
  extern __shared__ double lds[];
  const int lane = (int)(threadIdx.x & (WAVE - 1)), wv = (int)(threadIdx.x / WAVE);
  double* ws = lds;
#if RATES_GLOBAL
  double* rk = rscratch + (size_t)blockIdx.x * (2 * NR);  // this workgroup's slot
#else
  double* rk = ws + L_KF;
#endif
  double* W = lds + L_SIZE;
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
  int q = (int)blockIdx.x;
  while (q < nfl) {
    const int cell = list[q];
    const double Tc = Ta[cell], dt = dta[cell], nd = Pa[cell] / (KBOLTZ * Tc);
    const double* xr = x0 + (size_t)cell * NS;
    const double* Jc = NJ ? Jin + (size_t)cell * NJ : nullptr;
    // a listed cell passed the call kernel's screen, so it is not screened again
    cell_rates(ws, rk, Tc, nd, net, Jc, lane);
    const double h0 = isfinite(hin[cell]) ? hin[cell] : dt;
    const double m = fmin(h0, dt);
    int ntry = 0;
    for (int pass = 0;; pass++) {
      const gptr<int> perm = opaque(net)->perm;
      double Nl = lane < K0 ? xr[perm[lane]] : 0.0, Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
      const double start = pass == 0 ? m : fmin(m, dt * pow(REC_RETRY_SPLIT, (double)pass));
      Step s = {fmax(start, H_MIN), 0.0, 0.0, 0.0, 0.0, false};
      int it = dt > 0.0 ? 0 : max_steps;
      while (it < max_steps && s.t < dt * CALL_END) {
        const double ha = fmin(s.h, dt - s.t), igh = 1.0 / (GAMMA * ha), ih = 1.0 / ha;
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
        const double en = try_error(Nl, Nt, K1l, K1t, K3l, K3t, K4l, K4t, Yl, Yt, s.t == 0.0, rtol, atol, lane, ok);
        if (ok) {
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
      // the guard's test of the result, which goes to the cell's output row (a later pass overwrites it)
      const int dn = s.t >= dt * CALL_END ? 1 : 0;
      const int code = __builtin_amdgcn_readfirstlane(
          guard_wave(Nl, Nt, xr, xout + (size_t)cell * NS, dn != 0, s.nok, atol, net, lane));
      if (code == GUARD_DRIFT && pass < REC_RETRIES) continue;  // only drifted: again, with a smaller first try
      int nx = 0;
      if (lane == 0) {
        hio[cell] = s.h;
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
