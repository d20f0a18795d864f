// The learned warm start's network: BF16 matrix products on the MFMA units with FP32 accumulation, the GELU
// activation, the loss gradient, column sums and Adam. Rows are samples (cells); every matrix is row-major, and the
// sizes are padded by the host to multiples of the tiles (64 rows, 64 columns, 32 inner).
//
// A workgroup of four wavefronts computes one 64 x 64 tile of C = A B. Each step stages a 64 x 32 slice of A and a
// 32 x 64 slice of B in LDS (both with the inner index contiguous), and each wavefront accumulates its 32 x 32
// quarter with gfx950's v_mfma_f32_16x16x32_bf16, one product per 16 x 16 block and slice: lane l holds A row l % 16
// and B column l % 16 at inner indices 8 (l / 16) .. 8 (l / 16) + 7 (16 bytes, one LDS read each), and receives C
// rows 4 (l / 16) .. + 3 of column l % 16.

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int TILE_M = 64, TILE_N = 64, TILE_K = 32, LDS_K = TILE_K + 8;  // 8: keeps rows off one bank

__device__ __forceinline__ float bf2f(bf16_t v) { return __uint_as_float((unsigned)v << 16); }

// Round to nearest even; NaN stays NaN (a quiet bit is kept), infinities stay.
__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

// GELU, the erf form, and its derivative.
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2 };

// C [M, N] = A B (+ bias, GELU); A (m, k) at A[m sam + k sak], B (k, n) at B[k sbk + n sbn]. EPI_F32 writes
// alpha C to Cf; EPI_BIAS writes C + bias to Cf; EPI_BIAS_GELU writes z = C + bias to Cf and GELU(z) in BF16 to Cb.
// Cf and Cb rows are ldc apart.
template <int EPI>
__device__ __forceinline__ void gemm(const bf16_t* __restrict__ A, long sam, long sak, const bf16_t* __restrict__ B,
                                     long sbk, long sbn, int K, int ntn, const float* __restrict__ bias,
                                     float* __restrict__ Cf, bf16_t* __restrict__ Cb, long ldc, float alpha) {
  __shared__ __attribute__((aligned(16))) bf16_t As[TILE_M * LDS_K];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[TILE_N * LDS_K];  // B's slice transposed: column n's inner
                                                                        // indices contiguous
  const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
  // a one-dimensional grid of ntn tiles per row of tiles
  const long m0 = (long)(blockIdx.x / ntn) * TILE_M, n0 = (long)(blockIdx.x % ntn) * TILE_N;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  float4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += TILE_K) {
    // 2048 elements of each slice over 256 threads: element e = t + 256 r, row e / 32, inner e % 32
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int e = t + 256 * r, row = e >> 5, kk = e & 31;
      As[row * LDS_K + kk] = A[(m0 + row) * sam + (long)(k0 + kk) * sak];
      Bs[row * LDS_K + kk] = B[(long)(k0 + kk) * sbk + (n0 + row) * sbn];
    }
    __syncthreads();
    {
      const int kq = 8 * (lane >> 4);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        a[i] = *(const bf16x8*)(As + (wm + 16 * i + (lane & 15)) * LDS_K + kq);
        b[i] = *(const bf16x8*)(Bs + (wn + 16 * i + (lane & 15)) * LDS_K + kq);
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long row = m0 + wm + 16 * i + 4 * (lane >> 4) + q, col = n0 + wn + 16 * j + (lane & 15);
        const float v = acc[i][j][q];
        if (EPI == EPI_F32) {
          Cf[row * ldc + col] = alpha * v;
        } else {
          const float z = v + bias[col];
          Cf[row * ldc + col] = z;
          if (EPI == EPI_BIAS_GELU) Cb[row * ldc + col] = f2bf(gelu(z));
        }
      }
}

extern "C" __global__ void __launch_bounds__(256)
    mlp_gemm_f32(const bf16_t* A, long sam, long sak, const bf16_t* B, long sbk, long sbn, int K, int ntn, float* Cf,
                 long ldc, float alpha) {
  gemm<EPI_F32>(A, sam, sak, B, sbk, sbn, K, ntn, nullptr, Cf, nullptr, ldc, alpha);
}

extern "C" __global__ void __launch_bounds__(256)
    mlp_gemm_bias_gelu(const bf16_t* A, long sam, long sak, const bf16_t* B, long sbk, long sbn, int K, int ntn,
                       const float* bias, float* Cf, bf16_t* Cb, long ldc) {
  gemm<EPI_BIAS_GELU>(A, sam, sak, B, sbk, sbn, K, ntn, bias, Cf, Cb, ldc, 1.0f);
}

extern "C" __global__ void __launch_bounds__(256)
    mlp_gemm_bias(const bf16_t* A, long sam, long sak, const bf16_t* B, long sbk, long sbn, int K, int ntn,
                  const float* bias, float* Cf, long ldc) {
  gemm<EPI_BIAS>(A, sam, sak, B, sbk, sbn, K, ntn, bias, Cf, nullptr, ldc, 1.0f);
}

// The loss and its gradient on the output Y [M, ldy] against targets T: over rows < rows and columns < cols, the
// mean of w_c (Y - T)^2; dY = 2 w_c (Y - T) / (rows cols) in BF16 (zero in the padding), the sum of squares added to
// loss[0].
extern "C" __global__ void __launch_bounds__(256)
    mlp_loss(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ w, int rows, int cols,
             long ldy, long total, bf16_t* __restrict__ dY, float* __restrict__ loss) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (e < total) {
    const long r = e / ldy;
    const int c = (int)(e - r * ldy);
    float g = 0.0f;
    if (r < rows && c < cols) {
      const float d = Y[e] - T[r * cols + c];
      s = w[c] * d * d;
      g = 2.0f * w[c] * d / ((float)rows * (float)cols);
    }
    dY[e] = f2bf(g);
  }
  // one atomic per wavefront
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0 && s != 0.0f) atomicAdd(loss, s);
}

// dZ = dA GELU'(Z) in BF16, elementwise over n values.
extern "C" __global__ void __launch_bounds__(256)
    mlp_gelu_backward(const float* __restrict__ dA, const float* __restrict__ Z, bf16_t* __restrict__ dZ, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dZ[e] = f2bf(dA[e] * gelu_grad(Z[e]));
}

// Column sums of D [rows, cols] (BF16), in two fixed-order passes: mlp_column_sum adds the rows of one block of
// rows_per_block rows for 64 columns (4 row phases, then the phases) into partial [row blocks, cols]; mlp_column_total
// adds the row blocks. Block b of the first pass takes columns 64 (b % col_blocks) on, rows rows_per_block
// (b / col_blocks) on.
extern "C" __global__ void __launch_bounds__(256)
    mlp_column_sum(const bf16_t* __restrict__ D, int rows, int cols, int rows_per_block, int col_blocks,
                   float* __restrict__ partial) {
  __shared__ float part[4][64];
  const int cb = (int)blockIdx.x % col_blocks, rb = (int)blockIdx.x / col_blocks;
  const int c = cb * 64 + (int)(threadIdx.x & 63), ph = (int)(threadIdx.x >> 6);
  const int r0 = rb * rows_per_block, r1 = min(rows, r0 + rows_per_block);
  float s = 0.0f;
  if (c < cols)
    for (int r = r0 + ph; r < r1; r += 4) s += bf2f(D[(long)r * cols + c]);
  part[ph][threadIdx.x & 63] = s;
  __syncthreads();
  if (ph == 0 && c < cols)
    partial[(long)rb * cols + c] = (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

extern "C" __global__ void __launch_bounds__(256)
    mlp_column_total(const float* __restrict__ partial, int blocks, int cols, float* __restrict__ g) {
  const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= cols) return;
  float s = 0.0f;
  for (int b = 0; b < blocks; b++) s += partial[(long)b * cols + c];
  g[c] = s;
}

// One Adam step on n parameters (FP32 master copy p, moments m and v, gradient g), with the BF16 copy pb for the
// products; bc1 = 1 - beta1^t and bc2 = 1 - beta2^t.
extern "C" __global__ void __launch_bounds__(256)
    mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
             bf16_t* __restrict__ pb, long n, float lr, float beta1, float beta2, float eps, float bc1, float bc2,
             float decay) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = beta1 * m[e] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[e] + (1.0f - beta2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  const float pi = p[e] - lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + decay * p[e]);
  p[e] = pi;
  if (pb) pb[e] = f2bf(pi);
}

// FP32 to BF16, elementwise.
extern "C" __global__ void __launch_bounds__(256)
    mlp_to_bf16(const float* __restrict__ x, bf16_t* __restrict__ y, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e]);
}

// ---- the solver's side: features, targets and decoded predictions ----
// A cell's features (row of F [rows, ldf], BF16): (log10 max(x_i, 1e-30) for the ns species, ln T, ln P, ln dt,
// log10 J_b for the nj branches with J = 0 as -20), standardised by mu and sd; rows past the cells and columns past
// the features are 0.
extern "C" __global__ void __launch_bounds__(256)
    mlp_features(const double* __restrict__ x, const double* __restrict__ T, const double* __restrict__ P,
                 const double* __restrict__ dt, const double* __restrict__ J, int cells, int ns, int nj,
                 const float* __restrict__ mu, const float* __restrict__ sd, bf16_t* __restrict__ F, int rows,
                 int ldf) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)rows * ldf) return;
  const int r = (int)(e / ldf), c = (int)(e - (long)r * ldf), nf = ns + 3 + nj;
  float v = 0.0f;
  if (r < cells && c < nf) {
    double f;
    if (c < ns) f = log10(fmax(x[(long)r * ns + c], 1e-30));
    else if (c == ns) f = log(T[r]);
    else if (c == ns + 1) f = log(P[r]);
    else if (c == ns + 2) f = log(fmax(dt[r], 1e-30));
    else {
      const double j = J ? J[(long)r * nj + (c - ns - 3)] : 0.0;
      f = j > 0.0 ? log10(j) : -20.0;
    }
    v = (float)((f - (double)mu[c]) / (double)sd[c]);
  }
  F[e] = f2bf(v);
}

// A call's targets (row of Tg [cells, ns + 2], FP32, standardised): ln h (the step the call ends with), ln(1 +
// accepted steps) and log10 max(x_out, 1e-30) - log10 max(x_in, 1e-30) per species.
extern "C" __global__ void __launch_bounds__(256)
    mlp_targets(const double* __restrict__ x0, const double* __restrict__ x1, const double* __restrict__ h,
                const double* __restrict__ steps, int cells, int ns, const float* __restrict__ mu,
                const float* __restrict__ sd, float* __restrict__ Tg) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int no = ns + 2;
  if (e >= (long)cells * no) return;
  const int r = (int)(e / no), c = (int)(e - (long)r * no);
  double f;
  if (c == 0) f = log(fmax(h[r], 1e-30));
  else if (c == 1) f = log1p(fmax(steps[r], 0.0));
  else f = log10(fmax(x1[(long)r * ns + c - 2], 1e-30)) - log10(fmax(x0[(long)r * ns + c - 2], 1e-30));
  Tg[e] = (float)((f - (double)mu[c]) / (double)sd[c]);
}

// The predictions of one cell from the network's output row Y [rows, ldy] (standardised by mu and sd):
// h0 = exp(ln h) within [hmin, dt]; cost = exp(ln(1 + steps)) - 1; and the proposal xp = x 10^dlog10x, scaled to sum 1
// and then moved onto the cell's element shares: x_i exp(sum_e lambda_e a_ie) with lambda from Newton steps on
// the shares (atoms [ns, nel]); a proposal that does not settle within the steps gets xp[0] = NaN (no proposal). With
// xp null only h0 and cost are made.
extern "C" __global__ void __launch_bounds__(256)
    mlp_decode(const float* __restrict__ Y, int ldy, const float* __restrict__ mu, const float* __restrict__ sd,
               const double* __restrict__ x, const double* __restrict__ dt, const double* __restrict__ atoms, int cells,
               int ns, int nel, double hmin, double* __restrict__ h0, double* __restrict__ cost,
               double* __restrict__ xp) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= cells) return;
  const float* y = Y + (long)r * ldy;
  const double lh = (double)y[0] * sd[0] + mu[0];
  h0[r] = fmin(fmax(exp(lh), hmin), fmax(dt[r], hmin));
  cost[r] = expm1(fmax((double)y[1] * sd[1] + mu[1], 0.0));
  if (xp == nullptr) return;
  const double* xr = x + (long)r * ns;
  double* o = xp + (long)r * ns;
  constexpr int MAXEL = 8;
  double e0[MAXEL], lam[MAXEL];
  double s0 = 0.0, sum = 0.0;
  for (int k = 0; k < nel; k++) e0[k] = 0.0, lam[k] = 0.0;
  for (int i = 0; i < ns; i++) {
    const double d = (double)y[2 + i] * sd[2 + i] + mu[2 + i];
    const double v = fmax(xr[i], 1e-30) * exp10(fmin(d, 30.0));
    o[i] = v;
    sum += v;
    for (int k = 0; k < nel; k++) e0[k] += atoms[i * nel + k] * xr[i];
  }
  for (int k = 0; k < nel; k++) s0 += e0[k];
  for (int i = 0; i < ns; i++) o[i] /= sum;
  // Newton on the shares: f_k(lambda) = e_k(x') / sum_e e(x') - e0_k / s0 with x'_i = o_i exp(lambda . a_i)
  bool settled = false;
  for (int it = 0; it < 8 && !settled; it++) {
    double e1[MAXEL], m[MAXEL * MAXEL], s1 = 0.0;
    for (int k = 0; k < nel; k++) e1[k] = 0.0;
    for (int k = 0; k < nel * nel; k++) m[k] = 0.0;
    for (int i = 0; i < ns; i++) {
      double g = 0.0;
      for (int k = 0; k < nel; k++) g += lam[k] * atoms[i * nel + k];
      const double v = o[i] * exp(g);
      for (int k = 0; k < nel; k++) {
        const double ak = atoms[i * nel + k];
        if (ak == 0.0) continue;
        e1[k] += ak * v;
        for (int l = 0; l < nel; l++) m[k * nel + l] += ak * atoms[i * nel + l] * v;
      }
    }
    for (int k = 0; k < nel; k++) s1 += e1[k];
    double res[MAXEL], worst = 0.0;
    for (int k = 0; k < nel; k++) {
      res[k] = e0[k] / s0 * s1 - e1[k];  // the element totals the shares ask for, minus the present ones
      worst = fmax(worst, fabs(res[k]) / fmax(e0[k] / s0 * s1, 1e-300));
    }
    settled = worst < 1e-12;
    if (settled) break;
    // m lambda_step = res (m = A^T diag(x') A, the totals' derivative in lambda; the sum's change is second order
    // near the answer), by elimination without exchanges (m is symmetric positive semidefinite); a vanishing pivot
    // leaves its element out
    double m0[MAXEL];
    for (int k = 0; k < nel; k++) m0[k] = m[k * nel + k];
    for (int k = 0; k < nel; k++) {
      if (!(m[k * nel + k] > 1e-14 * m0[k])) continue;
      for (int j = k + 1; j < nel; j++) {
        const double l = m[j * nel + k] / m[k * nel + k];
        for (int c = k; c < nel; c++) m[j * nel + c] -= l * m[k * nel + c];
        res[j] -= l * res[k];
      }
    }
    for (int k = nel - 1; k >= 0; k--) {
      if (!(m[k * nel + k] > 1e-14 * m0[k])) continue;
      double b = res[k];
      for (int c = k + 1; c < nel; c++) b -= m[k * nel + c] * (m[c * nel + c] > 1e-14 * m0[c] ? res[c] : 0.0);
      res[k] = b / m[k * nel + k];
    }
    for (int k = 0; k < nel; k++) lam[k] += (m[k * nel + k] > 1e-14 * m0[k]) ? fmin(fmax(res[k], -5.0), 5.0) : 0.0;
  }
  if (!settled) {
    o[0] = __longlong_as_double(0x7ff8000000000000ll);
    return;
  }
  double s = 0.0;
  for (int i = 0; i < ns; i++) {
    double g = 0.0;
    for (int k = 0; k < nel; k++) g += lam[k] * atoms[i * nel + k];
    o[i] *= exp(g);
    s += o[i];
  }
  for (int i = 0; i < ns; i++) o[i] /= s;
}
