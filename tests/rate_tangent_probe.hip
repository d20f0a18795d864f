// Test harness, appended to the tangent kernel's translation unit (the solver's kernel files and jvp.hip): each
// cell's rate coefficients (cell_rates) and their tangents (rates_tangent) for given tangents of T and of the number
// density, one wavefront per cell. Outputs kf, kr [cells][NR] and their tangents dkf, dkr [cells][NR].
extern "C" __global__ void __launch_bounds__(WAVE)
    mxc_rate_tangent_probe(NetP __restrict__ net, const double* __restrict__ Ta, const double* __restrict__ Pa,
                           const double* __restrict__ Td, const double* __restrict__ ndd,
                           const double* __restrict__ Jin, double* __restrict__ kf, double* __restrict__ kr,
                           double* __restrict__ dkf, double* __restrict__ dkr) {
  extern __shared__ double lds[];
  const int lane = (int)threadIdx.x, cell = (int)blockIdx.x;
  double* ws = lds;
  double* ts = ws + L_SIZE;
  double* rk = ws + L_KF;
  double* rkd = ts + J_KD;
  const double T = Ta[cell], nd = Pa[cell] / (KBOLTZ * T);
  const double* Jc = NJ ? Jin + (size_t)cell * NJ : nullptr;
  cell_rates(ws, rk, T, nd, net, Jc, lane);
  rates_tangent(ts, rkd, T, Td[cell], nd, ndd[cell], net, Jc, nullptr, nullptr, nullptr, nullptr, lane);
  wave_sync();
  for (int r = lane; r < NR; r += WAVE) {
    kf[(size_t)cell * NR + r] = rk[r];
    kr[(size_t)cell * NR + r] = rk[NR + r];
    dkf[(size_t)cell * NR + r] = rkd[r];
    dkr[(size_t)cell * NR + r] = rkd[NR + r];
  }
}
