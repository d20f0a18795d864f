// The GCM side of a chemistry call, fused: mass mixing ratios to mole fractions and back, element moles and the
// tracers' totals, each one pass over the [cells, S] array with its input checks counted in the same pass (the
// reference's maxichem/coupling/composition.py:102-233 runs each as several PyTorch kernels, one per operation and
// check). HBM-bound. A wavefront takes CELLS_PER_WAVE cells at a time: lane l holds species l, l + 64, ... of each
// (S <= 64 * MAX_ROUNDS), every row's loads are issued before any of its sums, and each row is read once and written
// once. Row sums reduce within rows of 16 lanes by DPP, then over the four rows by v_readlane, in a fixed order that
// leaves the same wave-uniform bits for every lane.
//
// flags (int32, set by the host to {0, 0, INT_MAX, 0, 0, 0, INT_MAX, 0} before a launch):
//   [0] non-finite entries of the [cells, S] input   [1] negative entries of it
//   [2] first cell whose row sum (n or x * mw) is not above 0   [3] molar masses not finite and positive
//   [4] per-cell masses not finite or negative   [5] non-finite atoms   [6] first cell of [4]

constexpr int WAVE = 64;
constexpr int WAVES_PER_BLOCK = 4;  // blocks of 256 threads
constexpr int CELLS_PER_WAVE = 4;
constexpr int CELLS_PER_BLOCK = WAVES_PER_BLOCK * CELLS_PER_WAVE;
constexpr int MAX_ROUNDS = 4;  // S <= 256: R = ceil(S / 64) rounds of lanes, a template parameter
constexpr int MAX_ELEMENTS = 8;
constexpr int DPP_XOR1 = 0xb1, DPP_XOR2 = 0x4e, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp32<CTRL>((int)(b & 0xffffffffll)), hi = dpp32<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(dpp32<CTRL>(__float_as_int(v)));
}
__device__ __forceinline__ double readlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// the sum over the 64 lanes, the same bits in every lane (all lanes active)
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp<DPP_XOR1>(v);
  v += dpp<DPP_XOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  v += dpp<DPP_MIRROR>(v);
  return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

__device__ __forceinline__ int wave_sum(int v) {
  v += dpp32<DPP_XOR1>(v);
  v += dpp32<DPP_XOR2>(v);
  v += dpp32<DPP_HALF_MIRROR>(v);
  v += dpp32<DPP_MIRROR>(v);
  return (__builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16)) +
         (__builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48));
}

__device__ __forceinline__ int wave_count(bool b) { return __popcll(__ballot(b)); }

template <typename T>
__device__ __forceinline__ bool finite(T v) {
  return isfinite(v);
}

// block 0's first wavefront counts the molar masses that are not finite and positive
template <typename T>
__device__ __forceinline__ void check_mw(const T* __restrict__ mw, int S, int* __restrict__ flags) {
  if (blockIdx.x != 0 || threadIdx.x >= WAVE) return;
  int bad = 0;
  for (int j = (int)threadIdx.x; j < S; j += WAVE) bad += wave_count(!(finite(mw[j]) && mw[j] > T(0)));
  if (threadIdx.x == 0 && bad) atomicAdd(flags + 3, bad);
}

// this wave's first cell in a launch of ceil(B / CELLS_PER_BLOCK) blocks
__device__ __forceinline__ int first_cell() {
  return ((int)blockIdx.x * WAVES_PER_BLOCK + (int)(threadIdx.x / WAVE)) * CELLS_PER_WAVE;
}

// x = n / sum(n), n = q / mw, per cell (composition.py:102-159); negatives counted unless signed_ (the signed form
// for output keeps them)
template <typename T, int R>
__device__ void to_mole_r(const T* __restrict__ q, const T* __restrict__ mw, T* __restrict__ x, int B, int S,
                        int signed_, int* __restrict__ flags) {
  check_mw(mw, S, flags);
  const int lane = (int)(threadIdx.x & (WAVE - 1)), c0 = first_cell();
  if (c0 >= B) return;
  T w[R], v[CELLS_PER_WAVE][R];
#pragma unroll
  for (int r = 0; r < R; r++) w[r] = r * WAVE + lane < S ? mw[r * WAVE + lane] : T(1);
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++)
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = r * WAVE + lane;
      v[c][r] = c0 + c < B && j < S ? q[(size_t)(c0 + c) * S + j] : T(0);
    }
  int nonfinite = 0, negative = 0;
  T total[CELLS_PER_WAVE];
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++) {
    T t = T(0);
#pragma unroll
    for (int r = 0; r < R; r++) {
      nonfinite += !finite(v[c][r]);
      negative += v[c][r] < T(0);
      v[c][r] = v[c][r] / w[r];
      t += v[c][r];
    }
    total[c] = t;
  }
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++) total[c] = wave_sum(total[c]);
  nonfinite = wave_sum(nonfinite);
  negative = signed_ ? 0 : wave_sum(negative);
  if (lane == 0) {
    if (nonfinite) atomicAdd(flags, nonfinite);
    if (negative) atomicAdd(flags + 1, negative);
    if (!signed_)
      for (int c = 0; c < CELLS_PER_WAVE; c++)
        if (c0 + c < B && !(total[c] > T(0))) {
          atomicMin(flags + 2, c0 + c);
          break;
        }
  }
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++)
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = r * WAVE + lane;
      if (c0 + c < B && j < S) x[(size_t)(c0 + c) * S + j] = v[c][r] / total[c];
    }
}

// q = x mw (mass / sum(x mw)), per cell (composition.py:162-192)
template <typename T, int R>
__device__ void to_mass_r(const T* __restrict__ x, const T* __restrict__ mw, const T* __restrict__ mass,
                        T* __restrict__ q, int B, int S, int* __restrict__ flags) {
  check_mw(mw, S, flags);
  const int lane = (int)(threadIdx.x & (WAVE - 1)), c0 = first_cell();
  if (c0 >= B) return;
  T w[R], v[CELLS_PER_WAVE][R], m[CELLS_PER_WAVE];
#pragma unroll
  for (int r = 0; r < R; r++) w[r] = r * WAVE + lane < S ? mw[r * WAVE + lane] : T(0);
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++) {
    m[c] = c0 + c < B ? mass[c0 + c] : T(0);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = r * WAVE + lane;
      v[c][r] = c0 + c < B && j < S ? x[(size_t)(c0 + c) * S + j] : T(0);
    }
  }
  int nonfinite = 0, negative = 0;
  T weight[CELLS_PER_WAVE];
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++) {
    T t = T(0);
#pragma unroll
    for (int r = 0; r < R; r++) {
      nonfinite += !finite(v[c][r]);
      negative += v[c][r] < T(0);
      v[c][r] = v[c][r] * w[r];
      t += v[c][r];
    }
    weight[c] = t;
  }
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++) weight[c] = wave_sum(weight[c]);
  nonfinite = wave_sum(nonfinite);
  negative = wave_sum(negative);
  if (lane == 0) {
    if (nonfinite) atomicAdd(flags, nonfinite);
    if (negative) atomicAdd(flags + 1, negative);
    for (int c = 0; c < CELLS_PER_WAVE; c++)
      if (c0 + c < B && !(weight[c] > T(0))) {
        atomicMin(flags + 2, c0 + c);
        break;
      }
    for (int c = 0; c < CELLS_PER_WAVE; c++)
      if (c0 + c < B && !(finite(m[c]) && m[c] >= T(0))) {
        atomicAdd(flags + 4, 1);
        atomicMin(flags + 6, c0 + c);
      }
  }
#pragma unroll
  for (int c = 0; c < CELLS_PER_WAVE; c++) {
    const T ratio = m[c] / weight[c];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = r * WAVE + lane;
      if (c0 + c < B && j < S) q[(size_t)(c0 + c) * S + j] = v[c][r] * ratio;
    }
  }
}

// the sum over each group of 8 lanes (pairs, quads, half-rows by DPP), the same bits in the group's lanes
template <typename T>
__device__ __forceinline__ T group8_sum(T v) {
  v += dpp<DPP_XOR1>(v);
  v += dpp<DPP_XOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  return v;
}

constexpr int GROUP = 8;           // element_moles: lanes per cell
constexpr int GROUP_UNROLL = 8;    // species per lane whose loads are issued together
constexpr int GROUP_CELLS = WAVE * WAVES_PER_BLOCK / GROUP;  // cells per block

// e_k = sum_j (q_j / mw_j) atoms[j][k] per cell (composition.py:195-217), written to out [B][E] when out is given:
// 8 lanes per cell, lane s taking species s, s + 8, ... (8 cells' rows per load, 64 bytes of each), summed over the
// group by DPP. With cell_mass (the tracers' totals, composition.py:213-233), the block's share of
// sum_c sum_j q_cj cell_mass_c and sum_c e_ck cell_mass_c goes to partial[block][1 + E]: the groups of a wave, then
// the block's four waves, added in a fixed order.
template <typename T>
__device__ void element_moles(const T* __restrict__ q, const T* __restrict__ mw, const T* __restrict__ atoms,
                              const T* __restrict__ cell_mass, T* __restrict__ out, T* __restrict__ partial, int B,
                              int S, int E, int* __restrict__ flags) {
  __shared__ T part[WAVES_PER_BLOCK][1 + MAX_ELEMENTS];
  check_mw(mw, S, flags);
  if (blockIdx.x == 0 && threadIdx.x < WAVE) {
    int bad = 0;
    for (int i = (int)threadIdx.x; i < S * E; i += WAVE) bad += wave_count(!finite(atoms[i]));
    if (threadIdx.x == 0 && bad) atomicAdd(flags + 5, bad);
  }
  const int lane = (int)(threadIdx.x & (WAVE - 1)), wv = (int)(threadIdx.x / WAVE), sub = lane & (GROUP - 1);
  const int cell = (int)blockIdx.x * GROUP_CELLS + (int)(threadIdx.x / GROUP);
  const bool live = cell < B;
  T e[MAX_ELEMENTS], qs = T(0);
#pragma unroll
  for (int k = 0; k < MAX_ELEMENTS; k++) e[k] = T(0);
  int nonfinite = 0;
  const T cm = cell_mass == nullptr ? T(1) : live ? cell_mass[cell] : T(0);
  if (live) {
    const T* row = q + (size_t)cell * S;
    for (int j0 = 0; j0 < S; j0 += GROUP * GROUP_UNROLL) {
      T v[GROUP_UNROLL];
#pragma unroll
      for (int u = 0; u < GROUP_UNROLL; u++) {
        const int j = j0 + u * GROUP + sub;
        v[u] = j < S ? row[j] : T(0);
      }
#pragma unroll
      for (int u = 0; u < GROUP_UNROLL; u++) {
        const int j = j0 + u * GROUP + sub;
        if (j < S) {
          nonfinite += !finite(v[u]);
          qs += v[u] * cm;
          const T n = v[u] / mw[j];
          const T* a = atoms + (size_t)j * E;
#pragma unroll
          for (int k = 0; k < MAX_ELEMENTS; k++)
            if (k < E) e[k] += n * a[k];
        }
      }
    }
    if (cell_mass != nullptr && sub == 0 && !(finite(cm) && cm >= T(0))) {
      atomicAdd(flags + 4, 1);
      atomicMin(flags + 6, cell);
    }
  }
  // every lane from here: the DPP sums take the whole wave
#pragma unroll
  for (int k = 0; k < MAX_ELEMENTS; k++)
    if (k < E) e[k] = group8_sum(e[k]);
  if (out != nullptr && live && sub < E) {
    T mine = e[0];
#pragma unroll
    for (int k = 1; k < MAX_ELEMENTS; k++)
      if (sub == k) mine = e[k];
    out[(size_t)cell * E + sub] = mine;
  }
  nonfinite = wave_sum(nonfinite);
  if (lane == 0 && nonfinite) atomicAdd(flags, nonfinite);
  if (partial == nullptr) return;
  T acc[1 + MAX_ELEMENTS];
  acc[0] = wave_sum(qs);
#pragma unroll
  for (int k = 0; k < MAX_ELEMENTS; k++) acc[1 + k] = k < E ? wave_sum(sub == 0 ? e[k] * cm : T(0)) : T(0);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k <= MAX_ELEMENTS; k++) part[wv][k] = acc[k];
  __syncthreads();
  if (threadIdx.x <= (unsigned)E) {
    T sum = T(0);
    for (int w = 0; w < WAVES_PER_BLOCK; w++) sum += part[w][threadIdx.x];
    partial[(size_t)blockIdx.x * (1 + E) + threadIdx.x] = sum;
  }
}

// the body for R = ceil(S / 64) rounds (a wave-uniform branch)
#define MXC_BY_ROUNDS(FN, ...)                 \
  switch ((S + WAVE - 1) / WAVE) {              \
    case 0:                                     \
    case 1: FN<T, 1>(__VA_ARGS__); break;       \
    case 2: FN<T, 2>(__VA_ARGS__); break;       \
    case 3: FN<T, 3>(__VA_ARGS__); break;       \
    default: FN<T, 4>(__VA_ARGS__); break;      \
  }

template <typename T>
__device__ void to_mole(const T* q, const T* mw, T* x, int B, int S, int signed_, int* flags) {
  MXC_BY_ROUNDS(to_mole_r, q, mw, x, B, S, signed_, flags)
}
template <typename T>
__device__ void to_mass(const T* x, const T* mw, const T* mass, T* q, int B, int S, int* flags) {
  MXC_BY_ROUNDS(to_mass_r, x, mw, mass, q, B, S, flags)
}

// out[k] = sum over rows of partial[row][k]: one block of 256 threads per column, each thread's rows in order, then
// the threads in a fixed tree
template <typename T>
__device__ void column_sums(const T* __restrict__ partial, int rows, int width, T* __restrict__ out) {
  __shared__ T s[WAVE * WAVES_PER_BLOCK];
  const int k = (int)blockIdx.x, t = (int)threadIdx.x;
  T v = T(0);
  for (int r = t; r < rows; r += WAVE * WAVES_PER_BLOCK) v += partial[(size_t)r * width + k];
  s[t] = v;
  __syncthreads();
  for (int h = WAVE * WAVES_PER_BLOCK / 2; h > 0; h >>= 1) {
    if (t < h) s[t] += s[t + h];
    __syncthreads();
  }
  if (t == 0) out[k] = s[0];
}

#define MXC_COUPLING(T, SUFFIX)                                                                                     \
  extern "C" __global__ void __launch_bounds__(WAVE* WAVES_PER_BLOCK)                                              \
      mxc_to_mole_##SUFFIX(const T* q, const T* mw, T* x, int B, int S, int signed_, int* flags) {                  \
    to_mole<T>(q, mw, x, B, S, signed_, flags);                                                                      \
  }                                                                                                                  \
  extern "C" __global__ void __launch_bounds__(WAVE* WAVES_PER_BLOCK)                                              \
      mxc_to_mass_##SUFFIX(const T* x, const T* mw, const T* mass, T* q, int B, int S, int* flags) {                \
    to_mass<T>(x, mw, mass, q, B, S, flags);                                                                         \
  }                                                                                                                  \
  extern "C" __global__ void __launch_bounds__(WAVE* WAVES_PER_BLOCK)                                              \
      mxc_element_moles_##SUFFIX(const T* q, const T* mw, const T* atoms, const T* cell_mass, T* out, T* partial,   \
                                 int B, int S, int E, int* flags) {                                                  \
    element_moles<T>(q, mw, atoms, cell_mass, out, partial, B, S, E, flags);                                         \
  }                                                                                                                  \
  extern "C" __global__ void __launch_bounds__(WAVE* WAVES_PER_BLOCK)                                              \
      mxc_column_sums_##SUFFIX(const T* partial, int rows, int width, T* out) {                                      \
    column_sums<T>(partial, rows, width, out);                                                                       \
  }

MXC_COUPLING(double, f64)
MXC_COUPLING(float, f32)
