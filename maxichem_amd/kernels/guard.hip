// The guard of a stepped cell, run by the wavefront that stepped it while the cell's state is still in its lanes:
// the test of the result and, when it passes, the element-share correction (the reference's guard kernel,
// maxichem/kernels/guard.cuh:8-140, whose thread per cell re-reads the input and the result from HBM).
//
// Each lane holds two species of the cell (the lead row `lane` and the trailing row K0 + lane of the elimination
// order), so every sum over species is a lane product summed by DPP (wave_sums); the element-by-element system is
// NEL x NEL and is solved the same way on every lane.

constexpr int GUARD_NONFINITE = 4;
constexpr int GUARD_UNFINISHED = 8;
constexpr int GUARD_DRIFT = 16;

// Packed index of the entry (e, f), e <= f, of a symmetric NEL x NEL matrix stored by rows of its upper triangle.
__host__ __device__ constexpr int sym_index(int e, int f) { return e * NEL - e * (e - 1) / 2 + (f - e); }
constexpr int NSYM = NEL * (NEL + 1) / 2;

// Checks the cell's result x = N / sum N (N: this lane's lead and trailing entries) against its input row x0
// (network order) and writes the result, corrected when it passes, to out (network order). Returns the GUARD_*
// bits, 0 when it passed:
// - non-finite: an entry of x is not finite;
// - unfinished: the call did not reach dt (done false);
// - drift: an element's share moved by more than DRIFT relative plus what clamping nok accepted steps at -atol can
//   add, |e1/s1 - e0/s0| > DRIFT e0/s0 + 2 atol nok ETOT / s0 (e: the element's atoms, s: all atoms).
// A passing result is moved onto the input's element shares by x_i <- x_i g_i, g_i = 1 + sum_e lambda_e a_ie
// clamped to 1 +- GMAX, with (A^T X A) lambda = e0 s1 / s0 - e1 solved by LDL^T (an element whose pivot falls to
// ELEMENT_PIVOT_MIN of its diagonal is left out), then renormalised; a correction whose sum is not finite and
// positive makes the cell non-finite (x g is written as it is).
__device__ __forceinline__ int guard_wave(double Nl, double Nt, const double* __restrict__ x0,
                                          double* __restrict__ out, bool done, double nok, double atol,
                                          NetP net, int lane) {
  net = opaque(net);
  const gptr<int> perm = net->perm;
  const gptr<double> atoms = net->atoms;
  const bool hl = lane < K0, ht = lane < NT;
  const int il = hl ? perm[lane] : 0, it = ht ? perm[K0 + lane] : 0;
  // the input's entries and each species' atoms, loaded before the state's sum is waited for
  const double al = hl ? x0[il] : 0.0, at = ht ? x0[it] : 0.0;
  double Al[NEL], At[NEL];
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    Al[e] = hl ? atoms[il * NEL + e] : 0.0;
    At[e] = ht ? atoms[it * NEL + e] : 0.0;
  }
  const double s = vsum(Nl, Nt, lane);
  double xl = hl ? Nl / s : 0.0, xt = ht ? Nt / s : 0.0;
  // element totals of the input (e0, first NEL) and of the result (e1)
  double e01[2 * NEL];
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    e01[e] = fma(Al[e], al, At[e] * at);
    e01[NEL + e] = fma(Al[e], xl, At[e] * xt);
  }
  wave_sums(e01);
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    s0 += e01[e];
    s1 += e01[NEL + e];
  }
  bool drift = false;
#pragma unroll
  for (int e = 0; e < NEL; e++) {
    const double r0 = e01[e] / s0;
    drift = drift || fabs(e01[NEL + e] / s1 - r0) > DRIFT * r0 + 2.0 * atol * nok * ETOT / s0;
  }
  const bool xfin = wave_all(isfinite(xl) && isfinite(xt));
  const int code = (xfin ? 0 : GUARD_NONFINITE) | (done ? 0 : GUARD_UNFINISHED) | (drift ? GUARD_DRIFT : 0);
  if (code == 0) {
    // A^T X A, the upper triangle: each lane's two species' a_ie a_if x_i
    double m[NSYM];
#pragma unroll
    for (int e = 0; e < NEL; e++) {
#pragma unroll
      for (int f = e; f < NEL; f++) m[sym_index(e, f)] = fma(Al[e] * Al[f], xl, (At[e] * At[f]) * xt);
    }
    wave_sums(m);
    // LDL^T of the symmetric system on every lane alike (wave-uniform values)
    const double r = s1 / s0;
    double y[NEL], d0[NEL], lam[NEL];
    bool live[NEL];
#pragma unroll
    for (int e = 0; e < NEL; e++) {
      y[e] = fma(e01[e], r, -e01[NEL + e]);
      d0[e] = m[sym_index(e, e)];
    }
#pragma unroll
    for (int k = 0; k < NEL; k++) {
      live[k] = m[sym_index(k, k)] > ELEMENT_PIVOT_MIN * d0[k];
      if (!live[k]) continue;
      const double ik = 1.0 / m[sym_index(k, k)];
#pragma unroll
      for (int j = k + 1; j < NEL; j++) {
        const double l = m[sym_index(k, j)] * ik;
#pragma unroll
        for (int f = j; f < NEL; f++) m[sym_index(j, f)] = fma(-l, m[sym_index(k, f)], m[sym_index(j, f)]);
        y[j] = fma(-l, y[k], y[j]);
      }
    }
#pragma unroll
    for (int k = NEL - 1; k >= 0; k--) {
      double b = y[k];
#pragma unroll
      for (int f = k + 1; f < NEL; f++) b = fma(-m[sym_index(k, f)], lam[f], b);
      lam[k] = live[k] ? b / m[sym_index(k, k)] : 0.0;
    }
    double gl = 1.0, gt = 1.0;
#pragma unroll
    for (int e = 0; e < NEL; e++) {
      if (Al[e] != 0.0) gl = fma(lam[e], Al[e], gl);
      if (At[e] != 0.0) gt = fma(lam[e], At[e], gt);
    }
    xl *= fmin(fmax(gl, 1.0 - GMAX), 1.0 + GMAX);
    xt *= fmin(fmax(gt, 1.0 - GMAX), 1.0 + GMAX);
    const double sum = vsum(xl, xt, lane);
    if (isfinite(sum) && sum > 0.0) {
      const double is = 1.0 / sum;
      xl *= is;
      xt *= is;
    } else {
      if (hl) out[il] = xl;
      if (ht) out[it] = xt;
      return GUARD_NONFINITE;
    }
  }
  if (hl) out[il] = xl;
  if (ht) out[it] = xt;
  return code;
}
