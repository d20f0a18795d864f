// The guard: one thread per cell checks the call's result and puts its element shares back (the reference's guard
// kernel, maxichem/kernels/guard.cuh:8-140).

constexpr int GUARD_NONFINITE = 4;
constexpr int GUARD_UNFINISHED = 8;
constexpr int GUARD_DRIFT = 16;

// The guard's test of a stepped cell's result (network order) and, when it passes, the share correction
// x' = x (1 + sum_e lambda_e a_ie), clamped to 1 +- GMAX and renormalised (guard.cuh:24-108). Returns the GUARD_*
// bits, 0 when it passed.
__device__ int check_cell(const double* __restrict__ x0, double* __restrict__ xo, int done, double nok,
                          const double* __restrict__ atoms, double atol) {
  double e0[NEL], e1[NEL];
#pragma unroll
  for (int e = 0; e < NEL; e++) e0[e] = e1[e] = 0.0;
  bool xfin = true;
  for (int i = 0; i < NS; i++) {
    const double a = x0[i], b = xo[i];
    xfin = xfin && isfinite(b);
#pragma unroll
    for (int e = 0; e < NEL; e++) {
      e0[e] = fma(atoms[i * NEL + e], a, e0[e]);
      e1[e] = fma(atoms[i * NEL + e], b, e1[e]);
    }
  }
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    s0 += e0[e];
    s1 += e1[e];
  }
  bool drift = false;
#pragma unroll
  for (int e = 0; e < NEL; e++)
    drift = drift || fabs(e1[e] / s1 - e0[e] / s0) > DRIFT * e0[e] / s0 + 2.0 * atol * nok * ETOT / s0;
  const int code = (xfin ? 0 : GUARD_NONFINITE) | (done ? 0 : GUARD_UNFINISHED) | (drift ? GUARD_DRIFT : 0);
  if (code) return code;
  // (A^T X A) lambda = e0 s1 / s0 - e1 by elimination; an element whose pivot vanishes is left out
  double m[NEL * NEL], m0[NEL], y[NEL], lam[NEL];
  const double r = s1 / s0;
#pragma unroll
  for (int e = 0; e < NEL * NEL; e++) m[e] = 0.0;
  for (int i = 0; i < NS; i++) {
    const double* ai = atoms + i * NEL;
#pragma unroll
    for (int e = 0; e < NEL; e++) {
      if (ai[e] == 0.0) continue;
#pragma unroll
      for (int f = e; f < NEL; f++)
        if (ai[f] != 0.0) m[e * NEL + f] = fma(ai[e] * ai[f], xo[i], m[e * NEL + f]);
    }
  }
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    m0[e] = m[e * NEL + e];
    y[e] = fma(e0[e], r, -e1[e]);
  }
#pragma unroll
  for (int k = 0; k < NEL; k++) {
    if (!(m[k * NEL + k] > ELEMENT_PIVOT_MIN * m0[k])) continue;
#pragma unroll
    for (int j = k + 1; j < NEL; j++) {
      const double l = m[k * NEL + j] / m[k * NEL + k];
#pragma unroll
      for (int f = j; f < NEL; f++) m[j * NEL + f] = fma(-l, m[k * NEL + f], m[j * NEL + f]);
      y[j] = fma(-l, y[k], y[j]);
    }
  }
#pragma unroll
  for (int k = NEL - 1; k >= 0; k--) {
    lam[k] = 0.0;
    if (!(m[k * NEL + k] > ELEMENT_PIVOT_MIN * m0[k])) continue;
    double b = y[k];
#pragma unroll
    for (int f = k + 1; f < NEL; f++) b = fma(-m[k * NEL + f], lam[f], b);
    lam[k] = b / m[k * NEL + k];
  }
  double sum = 0.0;
  for (int i = 0; i < NS; i++) {
    const double* ai = atoms + i * NEL;
    double g = 1.0;
#pragma unroll
    for (int e = 0; e < NEL; e++)
      if (ai[e] != 0.0) g = fma(lam[e], ai[e], g);
    g = fmin(fmax(g, 1.0 - GMAX), 1.0 + GMAX);
    sum = fma(xo[i], g, sum);
    xo[i] *= g;
  }
  if (!(isfinite(sum) && sum > 0.0)) return GUARD_NONFINITE;
  const double is = 1.0 / sum;
  for (int i = 0; i < NS; i++) xo[i] *= is;
  return 0;
}

// Counts a refused cell by its status; tests a stepped one, and lists it for the recovery when it fails.
extern "C" __global__ void __launch_bounds__(GUARD_THREADS)
    mxc_guard(const Net* __restrict__ net, const double* __restrict__ x0, double* __restrict__ xout,
              const int* __restrict__ doneo, const double* __restrict__ noko, const int* __restrict__ status,
              int* __restrict__ redo, int* __restrict__ list, int* __restrict__ count, int B, double atol) {
  const int cell = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (cell >= B) return;
  const int s = status[cell];
  int r = 0;
  if (s == REFUSED_COLD)
    atomicAdd(count + COUNT_COLD, 1);
  else if (s == REFUSED_HOT)
    atomicAdd(count + COUNT_HOT, 1);
  else if (s == REFUSED_BAD_INPUT)
    atomicAdd(count + COUNT_BAD_INPUT, 1);
  else {
    const int code = check_cell(x0 + (size_t)cell * NS, xout + (size_t)cell * NS, doneo[cell], noko[cell],
                                net->atoms, atol);
    if (code & GUARD_NONFINITE) atomicAdd(count + COUNT_NONFINITE, 1);
    if (code & GUARD_UNFINISHED) atomicAdd(count + COUNT_UNFINISHED, 1);
    if (code & GUARD_DRIFT) atomicAdd(count + COUNT_DRIFT, 1);
    if (code) {
      r = 1;
      list[atomicAdd(count + COUNT_RECOMPUTED, 1)] = cell;
    }
  }
  redo[cell] = r;
}
