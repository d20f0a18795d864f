// Test harness kernels, compiled after the solver's kernel files (tests/probes.py): the guard, the recovery's dense LU
// and the step-size law, each alone on given inputs.

// The guard alone: cell b's result state N = xres[b] (network order; the wave's lanes take it as the call kernel
// holds it) checked against its input x0[b]; the written result goes to xout[b] and the guard's bits to code[b].
// One wavefront per cell.
extern "C" __global__ void __launch_bounds__(WAVE)
    mxc_guard_probe(NetP __restrict__ net, const double* __restrict__ x0, const double* __restrict__ xres,
                    const int* __restrict__ done, const double* __restrict__ nok, double* __restrict__ xout,
                    int* __restrict__ code, double atol) {
  const int lane = (int)threadIdx.x, cell = (int)blockIdx.x;
  const gptr<int> perm = net->perm;
  const double* xr = xres + (size_t)cell * NS;
  const double Nl = lane < K0 ? xr[perm[lane]] : 0.0, Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
  const int c = guard_wave(Nl, Nt, x0 + (size_t)cell * NS, xout + (size_t)cell * NS, done[cell] != 0, nok[cell],
                           atol, net, lane);
  if (lane == 0) code[cell] = c;
}

// The recovery's dense LU with threshold pivoting and its solve, on matrix b of W [count][NS][NS] (row-major, index
// order): the REC_WAVES wavefronts of the workgroup share the factorization, the first one solves with y[b]. W[b]
// is overwritten with the factors (unit L below the diagonal, U on and above), rows[b] with the original row of
// each factor row, dinv[b] with 1 / U_ii, y[b] with the solution.
extern "C" __global__ void __launch_bounds__(WAVE* REC_WAVES)
    mxc_dense_probe(double* __restrict__ Wg, double* __restrict__ dinvg, int* __restrict__ rows,
                    double* __restrict__ yg) {
  extern __shared__ double lds[];
  const int lane = (int)(threadIdx.x & (WAVE - 1)), wv = (int)(threadIdx.x / WAVE), b = (int)blockIdx.x;
  double* W = lds;
  double* dinv = W + NS * NS;
  double* y = dinv + NS;
  int* pr = (int*)(y + NS);
  for (int e = (int)threadIdx.x; e < NS * NS; e += WAVE * REC_WAVES) W[e] = Wg[(size_t)b * NS * NS + e];
  for (int i = (int)threadIdx.x; i < NS; i += WAVE * REC_WAVES) y[i] = yg[(size_t)b * NS + i];
  __syncthreads();
  dense_lu(W, pr, dinv, lane, wv);
  if (wv == 0) dense_solve(W, pr, dinv, y, lane);
  __syncthreads();
  for (int e = (int)threadIdx.x; e < NS * NS; e += WAVE * REC_WAVES) Wg[(size_t)b * NS * NS + e] = W[e];
  for (int i = (int)threadIdx.x; i < NS; i += WAVE * REC_WAVES) {
    dinvg[(size_t)b * NS + i] = dinv[i];
    rows[(size_t)b * NS + i] = pr[i];
    yg[(size_t)b * NS + i] = y[i];
  }
}

// The step-size law alone: row i of in is (en, ok, prej, pen, pha, ha) (flags as 0 / 1); out[i] = step_factor's
// factor on ha for the next try. One thread per row.
extern "C" __global__ void mxc_step_probe(const double* __restrict__ in, double* __restrict__ out, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const double* r = in + 6 * i;
  out[i] = step_factor(r[0], r[1] != 0.0, r[2] != 0.0, r[3], r[4], r[5]);
}
