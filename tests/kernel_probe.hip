// Test harness: one Rodas3 try's pieces for one cell, compiled after the solver's kernel files (tests/test_kernel_pieces.py).
// out (doubles, elimination order): [0] wave sums and maxima of known lane values, then F, w, -Jn (dense NS x NS),
// M0^-1 F, K1, K2, K3, K4. One workgroup of one wavefront.

extern "C" __global__ void __launch_bounds__(WAVE)
    mxc_probe(NetP __restrict__ net, const double* __restrict__ xr, double Tc, double Pc, double ha,
              const double* __restrict__ Jc, double* __restrict__ out) {
  extern __shared__ double lds[];
  double* ws = lds;
  const int lane = (int)threadIdx.x;
  // reductions of known values: sum of lane + 0.5 and of 2^-lane, maximum of (lane * 37) % 64
  out[0] = wave_sum((double)lane + 0.5);
  out[1] = wave_sum(ldexp(1.0, -lane));
  out[2] = wave_max((double)((lane * 37) % 64));
  out[3] = readlane((double)lane * 3.25, 17);
  double* o = out + 8;
  const double nd = Pc / (KBOLTZ * Tc);
  double c[NT];
  LeadRow a21;
  cell_rates(ws, ws + L_KF, Tc, nd, net, Jc, lane);
  const gptr<int> perm = net->perm;
  double Nl = lane < K0 ? xr[perm[lane]] : 0.0, Nt = lane < NT ? xr[perm[K0 + lane]] : 0.0;
  double Fl = 0.0, Ft = 0.0, ul = 0.0, ut = 0.0, u0l = 0.0, u0t = 0.0;
  const double ntot = rhs<true>(Nl, Nt, Fl, Ft, u0l, u0t, net, nd, ws, ws + L_KF, lane);
  jacobian(c, a21, Nl, Nt, u0l, u0t, ntot, ul, ut, net, ws, lane);
  auto put = [&](double* dst, double vl, double vt) {
    if (lane < K0) dst[lane] = vl;
    if (lane < NT) dst[K0 + lane] = vt;
  };
  put(o, Fl, Ft);
  put(o + NS, ul, ut);
  double* jn = o + 2 * NS;  // -Jn, dense
  for (int e = lane; e < K0 * NS; e += WAVE) jn[e] = ws[e];
  if (lane < NT) {
#pragma unroll
    for (int k = 0; k < K0; k++) jn[(K0 + lane) * NS + k] = a21[k];
    for (int j = 0; j < NT; j++) jn[(K0 + lane) * NS + K0 + j] = c[j];
  }
  double* lu = jn + NS * NS;  // the factors of M0 (unit L below the diagonal, U on and above), then 1 / U_ii
  double* v = lu + NS * NS + NS;
  const double igh = 1.0 / (GAMMA * ha), ih = 1.0 / ha;
  add_diagonal(c, igh, ws, lane);
  double dl, dtr;
  factor(c, a21, dl, dtr, ws, lane);
  wave_sync();
  for (int e = lane; e < K0 * NS; e += WAVE) lu[e] = ws[e];
  if (lane < NT) {
#pragma unroll
    for (int k = 0; k < K0; k++) lu[(K0 + lane) * NS + k] = a21[k];
    for (int j = 0; j < NT; j++) lu[(K0 + lane) * NS + K0 + j] = c[j];
  }
  put(lu + NS * NS, dl, dtr);
  double y2l[2] = {Fl, ul}, y2t[2] = {Ft, ut};
  solve<2>(y2l, y2t, c, a21, dl, dtr, ws, lane);
  put(v, y2l[0], y2t[0]);
  const double Ul = y2l[1], Ut = y2t[1];
  const double c0 = 1.0 / (1.0 - vsum(Ul, Ut, lane));
  double sk = vsum(y2l[0], y2t[0], lane) * c0;
  const double K1l = y2l[0] + sk * Ul, K1t = y2t[0] + sk * Ut;
  put(v + NS, K1l, K1t);
  double y1l[1] = {fma(RODAS_C21 * ih, K1l, Fl)}, y1t[1] = {fma(RODAS_C21 * ih, K1t, Ft)};
  solve<1>(y1l, y1t, c, a21, dl, dtr, ws, lane);
  sk = vsum(y1l[0], y1t[0], lane) * c0;
  const double K2l = y1l[0] + sk * Ul, K2t = y1t[0] + sk * Ut;
  put(v + 2 * NS, K2l, K2t);
  double Yl = Nl + 2.0 * K1l, Yt = Nt + 2.0 * K1t;
  rhs<false>(Yl, Yt, Fl, Ft, u0l, u0t, net, nd, ws, ws + L_KF, lane);
  y1l[0] = fma(ih, K1l - K2l, Fl);
  y1t[0] = fma(ih, K1t - K2t, Ft);
  solve<1>(y1l, y1t, c, a21, dl, dtr, ws, lane);
  sk = vsum(y1l[0], y1t[0], lane) * c0;
  const double K3l = y1l[0] + sk * Ul, K3t = y1t[0] + sk * Ut;
  put(v + 3 * NS, K3l, K3t);
  Yl = Nl + 2.0 * K1l + K3l;
  Yt = Nt + 2.0 * K1t + K3t;
  rhs<false>(Yl, Yt, Fl, Ft, u0l, u0t, net, nd, ws, ws + L_KF, lane);
  y1l[0] = fma(ih, fma(-RODAS_C43, K3l, K1l - K2l), Fl);
  y1t[0] = fma(ih, fma(-RODAS_C43, K3t, K1t - K2t), Ft);
  solve<1>(y1l, y1t, c, a21, dl, dtr, ws, lane);
  sk = vsum(y1l[0], y1t[0], lane) * c0;
  put(v + 4 * NS, y1l[0] + sk * Ul, y1t[0] + sk * Ut);
}
