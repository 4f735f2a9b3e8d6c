// gp_dct.hip — Poisson solve of the electrostatic density as dense DCT
// contractions (Poisson2d::solve, electrostatic_placement.rs:725-776).
//
// phi = C_H^T [ (C_H rho C_W^T) .* invLambda ] C_W, written as four
// "rows x basis" products; every field of every task is a block of rows and
// the basis is a static K x N operand:
//   1. Rt_f  = (rho_f C_W^T)^T            rows F*H, K = N = W, transposed store
//   2. St'_f = (Rt_f C_H^T) .* invLambda^T rows F*W, K = N = H, spectral epilogue
//   3. T_f   = (St'_f C_H)^T               rows F*W, K = N = H, transposed store
//   4. phi_f = T_f C_W                     rows F*H, K = N = W
// A launch covers the tiles of every task of one group (a block plan); a tile never straddles two
// tasks.  The electrostatic energy 1/2 <rho_bar, phi> equals
// 1/2 sum invLambda S^2 by Parseval (orthonormal basis); it is accumulated in
// FP64 in the epilogue of step 2.  The DC coefficient has invLambda = 0, so
// removing the mean of rho (electrostatic_placement.rs:739-744) changes
// neither phi nor the energy.

#include "gp_common.h"

namespace texo {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short ushort8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ double wave_sum(double v) { return lane_reduce<64>(v, OpAdd{}); }

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf16_to_float(unsigned short h) {
  return __uint_as_float(static_cast<unsigned int>(h) << 16);
}

// Power of two that brings max |a| into [2^14, 2^15), the top of FP16's
// range, so that the hi/lo split keeps 22 significant bits.
__device__ __forceinline__ float pow2_scale(float amax) {
  if (!(amax > 0.f)) return 1.f;
  int e = 0;
  (void)frexpf(amax, &e);
  e = e < -100 ? -100 : e;
  return ldexpf(1.f, 15 - e);
}

__device__ __forceinline__ void split_f16(float x, _Float16 &hi, _Float16 &lo) {
  hi = static_cast<_Float16>(x);
  lo = static_cast<_Float16>(x - static_cast<float>(hi));  // x - hi is exact in FP32
}

__device__ __forceinline__ half4_t half4_of(uint2 v) {
  half4_t h;
  __builtin_memcpy(&h, &v, 8);
  return h;
}

__device__ __forceinline__ float wave_max_f(float v) { return lane_reduce<64>(v, OpMaxF{}); }
// |x| as bits: ordered like the magnitudes, a NaN above infinity
__device__ __forceinline__ unsigned int abs_bits(float x) { return __float_as_uint(x) & 0x7fffffffu; }

// Stage geometry resolved per block from the task.
struct Stage {
  const float *A;
  float *out;
  unsigned int *amax_out;  // max |out| per output tile, for the next stage (fast path), or null
  int M, K, N, R;  // rows of this task, inner, columns, rows per field
  bool transposed;
  const float *inv_lambda;
  int W;
  double *energy;  // per-task partials [row blocks][col blocks][MAXF] or null
};

template <int S>
__device__ __forceinline__ Stage stage_of(const Task &t) {
  const int W = t.j.W, H = t.j.H, F = t.j.F;
  Stage s;
  s.W = W;
  s.inv_lambda = nullptr;
  s.energy = nullptr;
  // stages 0-2: the next stage's input range; stage 3: |phi| for the bound
  // on a trial's gradient
  s.amax_out = t.amax_part + (size_t)t.nb_conv * MAXF + (size_t)S * t.nb_tile;
  if constexpr (S == 0) {
    s.A = t.rho;
    s.out = t.t0;
    s.M = F * H;
    s.K = W;
    s.N = W;
    s.R = H;
    s.transposed = true;
  } else if constexpr (S == 1) {
    s.A = t.t0;
    s.out = t.t1;
    s.M = F * W;
    s.K = H;
    s.N = H;
    s.R = W;
    s.transposed = false;
    s.inv_lambda = t.j.inv_lambda;
    s.energy = t.part_energy;
  } else if constexpr (S == 2) {
    s.A = t.t1;
    s.out = t.t0;
    s.M = F * W;
    s.K = H;
    s.N = H;
    s.R = W;
    s.transposed = true;
  } else {
    s.A = t.t0;
    s.out = t.phi;
    s.M = F * H;
    s.K = W;
    s.N = W;
    s.R = H;
    s.transposed = false;
  }
  return s;
}

__device__ __forceinline__ void store_element(const Stage &e, int row, int col, float v, double (&energy)[MAXF]) {
  const int f = row / e.R, r = row - f * e.R;
  if (e.inv_lambda) {
    const float lam = e.inv_lambda[(size_t)col * e.W + r];
    const double dv = static_cast<double>(v);
    const double contribution = 0.5 * dv * dv * static_cast<double>(lam);
#pragma unroll
    for (int k = 0; k < MAXF; ++k)
      if (k == f) energy[k] += contribution;
    v = v * lam;
  }
  if (e.transposed)
    e.out[(size_t)f * e.R * e.N + (size_t)col * e.R + r] = v;
  else
    e.out[(size_t)row * e.N + col] = v;
}

// rows row0..row0+3 of column col; row0 % 4 == 0 and R % 4 == 0 keep the
// four rows inside one field
// Stores rows row0..row0+3 of column col; `v` leaves holding the stored values.
__device__ __forceinline__ void store4(const Stage &e, int row0, int col, float4_t &v, double (&energy)[MAXF]) {
  const int f = row0 / e.R, r0 = row0 - f * e.R;
  if (e.inv_lambda) {
    const float4 lam = *reinterpret_cast<const float4 *>(e.inv_lambda + (size_t)col * e.W + r0);
    const float l[4] = {lam.x, lam.y, lam.z, lam.w};
    double c = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double dv = static_cast<double>(v[q]);
      c += 0.5 * dv * dv * static_cast<double>(l[q]);
      v[q] = v[q] * l[q];
    }
#pragma unroll
    for (int k = 0; k < MAXF; ++k)
      if (k == f) energy[k] += c;
  }
  if (e.transposed) {
    *reinterpret_cast<float4 *>(e.out + (size_t)f * e.R * e.N + (size_t)col * e.R + r0) =
        make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) e.out[(size_t)(row0 + q) * e.N + col] = v[q];
  }
}

__device__ __forceinline__ void write_energy(const Stage &e, int local_rb, int bx, int gx, double (&energy)[MAXF]) {
  if (!e.energy) return;
#pragma unroll
  for (int k = 0; k < MAXF; ++k) energy[k] = wave_sum(energy[k]);
  if ((threadIdx.x & 63) == 0) {
    double *p = e.energy + ((size_t)local_rb * gx + bx) * MAXF;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) p[k] = energy[k];
  }
}

// ---------------------------------------------------------------------------
// Fast shape (W, H multiples of 32): one 256-thread block per 32x32 output
// tile, K split over its four wavefronts (wave w takes the 32-wide k-steps
// w, w+4, ...).  The four partial tiles meet in LDS and are summed in wave
// order, so results do not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int SPLIT = 4;
constexpr int FAST_THREADS = 64 * SPLIT;

// Epilogue of one 32x32 output tile over the whole block: thread t owns
// column t % 32 and rows 4 (t / 32) .. 4 (t / 32) + 3, whose summed values
// it holds in `v`.  Every summation order (energy, range) is fixed by the
// tile alone, so any kernel that hands it the same `v` writes the same bits.
// Block-wide, with a barrier before its partials are read.
__device__ __forceinline__ void tile_epilogue(const Stage &e, float4_t v, int m0, int n0, int local_rb, int bx,
                                              int gx) {
  const int col = threadIdx.x & 31, rg = threadIdx.x >> 5;
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
  store4(e, m0 + 4 * rg, n0 + col, v, energy);
  if (e.amax_out) {
    // the 32 rows of a fast-shape tile lie in one field; v now holds the
    // stored values (after the spectral scaling in stage 1)
    __shared__ unsigned int sm[SPLIT];
    unsigned int mx = max(max(abs_bits(v[0]), abs_bits(v[1])), max(abs_bits(v[2]), abs_bits(v[3])));
    mx = lane_reduce<64>(mx, OpMaxI{});
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) e.amax_out[(size_t)local_rb * gx + bx] = max(max(sm[0], sm[1]), max(sm[2], sm[3]));
  }
  if (!e.energy) return;
  __shared__ double se[SPLIT][MAXF];
#pragma unroll
  for (int k = 0; k < MAXF; ++k) energy[k] = wave_sum(energy[k]);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < MAXF; ++k) se[wave][k] = energy[k];
  }
  __syncthreads();
  if (threadIdx.x < MAXF) {
    const int k = threadIdx.x;
    e.energy[((size_t)local_rb * gx + bx) * MAXF + k] = ((se[0][k] + se[1][k]) + se[2][k]) + se[3][k];
  }
}

// Sums the per-wave partial tiles red[w][32][33] in wave order and runs the
// tile epilogue.
__device__ __forceinline__ void fast_epilogue(const Stage &e, float (*red)[32][33], int m0, int n0, int local_rb,
                                              int bx, int gx) {
  const int col = threadIdx.x & 31, rg = threadIdx.x >> 5;
  float4_t v;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = 4 * rg + q;
    v[q] = ((red[0][row][col] + red[1][row][col]) + red[2][row][col]) + red[3][row][col];
  }
  tile_epilogue(e, v, m0, n0, local_rb, bx, gx);
}

// max |input| (float bits) of field f of stage S's input, from the partials
// its producer wrote: the k_convert blocks covering field f for stage 0,
// else the 32x32 output tiles of field f of stage S-1.  Block-wide.
template <int S>
__device__ unsigned int input_amax(const Task &t, int f) {
  const int W = t.j.W, H = t.j.H;
  const unsigned int *p;
  int n, stride;
  if constexpr (S == 0) {
    const size_t G = (size_t)W * H;
    const int lo = (int)(f * G / BLOCK), hi = (int)(((f + 1) * G + BLOCK - 1) / BLOCK);
    p = t.amax_part + (size_t)lo * MAXF + f;
    n = hi - lo;
    stride = MAXF;
  } else {
    // producer stage S-1: rows per field R, columns N (both multiples of 32)
    const int R = (S - 1) == 0 ? H : W, N = (S - 1) == 0 ? W : H;
    n = (R / 32) * (N / 32);
    p = t.amax_part + (size_t)t.nb_conv * MAXF + (size_t)(S - 1) * t.nb_tile + (size_t)f * n;
    stride = 1;
  }
  unsigned int m = 0u;
  for (int i = threadIdx.x; i < n; i += FAST_THREADS) m = max(m, p[(size_t)i * stride]);
  m = lane_reduce<64>(m, OpMaxI{});
  __shared__ unsigned int sm[FAST_THREADS / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  return max(max(sm[0], sm[1]), max(sm[2], sm[3]));
}

// gfx950's double-rate v_mfma_f32_16x16x32_{bf16,f16}: lane l holds
// A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16], i = 0..7, eight
// halves (16 bytes) each; the accumulator layout is that of 16x16x16.
template <int P>
__device__ __forceinline__ float4_t mfma32(uint4 a, uint4 b, float4_t c) {
  if constexpr (P == 1)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, a), __builtin_bit_cast(half8_t, b), c,
                                                  0, 0, 0);
}

template <int P>
__device__ __forceinline__ void split_pair(float x, unsigned short &hi, unsigned short &lo) {
  if constexpr (P == 1) {
    hi = bf16_rne(x);
    lo = bf16_rne(x - bf16_to_float(hi));
  } else {
    _Float16 h, l;
    split_f16(x, h, l);
    hi = __builtin_bit_cast(unsigned short, h);
    lo = __builtin_bit_cast(unsigned short, l);
  }
}

// Split MFMA of one 32x32 output tile's k-steps w, w + 4, ... (32 wide),
// P = 1: BF16x3, P = 2: FP16-scaled (A multiplied by `sa` before the split).
// Each step issues hi*hi + hi*lo + lo*hi with FP32 accumulation.
template <int P>
__device__ __forceinline__ void split_steps(const float *a0, const float *a1, const uint4 *bh0, const uint4 *bh1,
                                            const uint4 *bl0, const uint4 *bl1, int K, int wave, float sa,
                                            float4_t (&acc)[2][2]) {
#pragma unroll 2
  for (int k0 = 32 * wave; k0 < K; k0 += 32 * SPLIT) {
    const float *ap[2] = {a0 + k0, a1 + k0};
    uint4 ahi[2], alo[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float4 v0 = *reinterpret_cast<const float4 *>(ap[i]), v1 = *reinterpret_cast<const float4 *>(ap[i] + 4);
      const float f8[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
      ushort8_t h, l;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        unsigned short hs, ls;
        split_pair<P>(P == 2 ? f8[s] * sa : f8[s], hs, ls);
        h[s] = hs;
        l[s] = ls;
      }
      ahi[i] = __builtin_bit_cast(uint4, h);
      alo[i] = __builtin_bit_cast(uint4, l);
    }
    const uint4 bhi[2] = {bh0[k0 / 8], bh1[k0 / 8]};
    const uint4 blo[2] = {bl0[k0 / 8], bl1[k0 / 8]};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = mfma32<P>(ahi[i], bhi[j], acc[i][j]);
        acc[i][j] = mfma32<P>(ahi[i], blo[j], acc[i][j]);
        acc[i][j] = mfma32<P>(alo[i], bhi[j], acc[i][j]);
      }
  }
}

// BF16x3 split-precision MFMA.  A (FP32 activations) is split on the fly
// into hi + lo BF16, B (static basis) is pre-split and pre-transposed so each
// fragment is one 16-byte load.
template <int S>
__global__ void __launch_bounds__(FAST_THREADS) k_gemm_bf16_fast(const Task *__restrict__ tasks,
                                                                 BlockPlan p) {
  __shared__ float red[SPLIT][32][33];
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's tiles: 32-row blocks x 32-column tiles, row-major
  const int gx = (e.N + 31) / 32, rb = bm.y / gx, bx = bm.y - rb * gx;
  const uint4 *Bth = reinterpret_cast<const uint4 *>(t.j.bthi[S]);
  const uint4 *Btl = reinterpret_cast<const uint4 *>(t.j.btlo[S]);
  const int K = e.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = rb * 32, n0 = bx * 32;
  const int q = lane >> 4, r = lane & 15;
  const float *a0 = e.A + (size_t)(m0 + r) * K + 8 * q;
  const float *a1 = e.A + (size_t)(m0 + 16 + r) * K + 8 * q;
  const size_t k8 = (size_t)K / 8;
  const uint4 *bh0 = Bth + (size_t)(n0 + r) * k8 + q, *bh1 = Bth + (size_t)(n0 + 16 + r) * k8 + q;
  const uint4 *bl0 = Btl + (size_t)(n0 + r) * k8 + q, *bl1 = Btl + (size_t)(n0 + 16 + r) * k8 + q;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  split_steps<1>(a0, a1, bh0, bh1, bl0, bl1, K, wave, 1.f, acc);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int s = 0; s < 4; ++s) red[wave][16 * i + 4 * q + s][16 * j + r] = acc[i][j][s];
  __syncthreads();
  fast_epilogue(e, red, m0, n0, rb, bx, gx);
}

// FP16-scaled split precision on MFMA.  The producer of A (k_convert, or the
// previous stage's epilogue) recorded max |A| per field; A is scaled by a
// power of two into the top of FP16's range and split into hi + lo FP16 (22
// significant bits); the basis is pre-scaled, pre-split and pre-transposed.  hi*hi + hi*lo + lo*hi run on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation; the product of the two
// power-of-two scales is divided out exactly before the epilogue.
template <int S>
__global__ void __launch_bounds__(FAST_THREADS) k_gemm_f16s_fast(const Task *__restrict__ tasks,
                                                                 BlockPlan p) {
  __shared__ float red[SPLIT][32][33];
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's tiles: 32-row blocks x 32-column tiles, row-major
  const int gx = (e.N + 31) / 32, rb = bm.y / gx, bx = bm.y - rb * gx;
  const uint4 *Bth = reinterpret_cast<const uint4 *>(t.j.fthi[S]);
  const uint4 *Btl = reinterpret_cast<const uint4 *>(t.j.ftlo[S]);
  const int K = e.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = rb * 32, n0 = bx * 32;
  const int q = lane >> 4, r = lane & 15;
  const float *a0 = e.A + (size_t)(m0 + r) * K + 8 * q;
  const float *a1 = e.A + (size_t)(m0 + 16 + r) * K + 8 * q;
  // max |A| over the 32 rows' field, from the partials of the producer of A
  const float sa = pow2_scale(__uint_as_float(input_amax<S>(t, m0 / e.R)));
  const float unscale = 1.f / (sa * t.j.f16_scale[S]);  // exact: a power of two
  const size_t k8 = (size_t)K / 8;
  const uint4 *bh0 = Bth + (size_t)(n0 + r) * k8 + q, *bh1 = Bth + (size_t)(n0 + 16 + r) * k8 + q;
  const uint4 *bl0 = Btl + (size_t)(n0 + r) * k8 + q, *bl1 = Btl + (size_t)(n0 + 16 + r) * k8 + q;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  split_steps<2>(a0, a1, bh0, bh1, bl0, bl1, K, wave, sa, acc);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int s = 0; s < 4; ++s) red[wave][16 * i + 4 * q + s][16 * j + r] = acc[i][j][s] * unscale;
  __syncthreads();
  fast_epilogue(e, red, m0, n0, rb, bx, gx);
}

// Split MFMA (BF16x3, P = 1, or FP16-scaled, P = 2) over 32 rows x 128
// columns per block, for K <= KMAX: the block splits its 32 rows of A into
// hi + lo once, into LDS, and each wave owns one 32x32 column tile over the
// whole K, its B fragments coming from the pre-split basis in L2.  The wave
// accumulates each class of 32-wide k-steps modulo 4 (the k-steps that wave
// c of the 32x32 kernel takes) on its own, in the same order, and sums the
// scaled class results in class order; its epilogue then replays the 32x32
// kernel's, the four 64-lane quarters of that kernel's block one after the
// other: the stored values, energy and range partials are bit for bit those
// of the 32x32 kernel, so a task's results do not depend on which of the two
// kernels its launch picked.
// Below this many wide tiles (one per CU) a launch leaves CUs idle and the
// 32x32 split-K kernel is faster; from one per CU the wide kernel wins, also
// beside the other job group's concurrent launches.
constexpr int WIDE_MIN_TILES = 256;
constexpr int WIDE_KMAX = 512;

template <int KMAX>
union WideLds {
  struct {
    unsigned short hi[32][KMAX + 8], lo[32][KMAX + 8];  // row stride KMAX + 8 halves
  } a;
  float red[SPLIT][32][33];
};


template <int S, int P, int KMAX>
__global__ void __launch_bounds__(FAST_THREADS) k_gemm_wide(const Task *__restrict__ tasks, BlockPlan p) {
  __shared__ WideLds<KMAX> lds;
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's wide tiles: 32-row blocks x 128-column tiles, row-major
  const int gw = e.N / (32 * SPLIT), rb = bm.y / gw, bw = bm.y - rb * gw;
  const int K = e.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = rb * 32, n0 = (bw * SPLIT + wave) * 32;
  const int q = lane >> 4, r = lane & 15;
  float sa = 1.f, unscale = 1.f;
  if constexpr (P == 2) {
    // max |A| over the 32 rows' field, from the partials of the producer of A
    sa = pow2_scale(__uint_as_float(input_amax<S>(t, m0 / e.R)));
    unscale = 1.f / (sa * t.j.f16_scale[S]);  // exact: a power of two
  }
  // stage A: 32 rows x K, coalesced float4 loads, every load of the thread
  // in flight before the first split (a loop of load, split, store waits
  // out one memory latency per iteration)
  const int k4 = K / 4;
  constexpr int AL = 32 * (KMAX / 4) / FAST_THREADS;  // loads per thread at K = KMAX
  float4 va[AL];
#pragma unroll
  for (int u = 0; u < AL; ++u) {
    const int idx = threadIdx.x + u * FAST_THREADS;
    if (idx < 32 * k4) {
      const int row = idx / k4, c = idx - row * k4;
      va[u] = *reinterpret_cast<const float4 *>(e.A + (size_t)(m0 + row) * K + 4 * c);
    }
  }
#pragma unroll
  for (int u = 0; u < AL; ++u) {
    const int idx = threadIdx.x + u * FAST_THREADS;
    if (idx < 32 * k4) {
      const int row = idx / k4, c = idx - row * k4;
      float4 v = va[u];
      if constexpr (P == 2) v = make_float4(v.x * sa, v.y * sa, v.z * sa, v.w * sa);
      ushort4 hi, lo;
      split_pair<P>(v.x, hi.x, lo.x);
      split_pair<P>(v.y, hi.y, lo.y);
      split_pair<P>(v.z, hi.z, lo.z);
      split_pair<P>(v.w, hi.w, lo.w);
      *reinterpret_cast<ushort4 *>(&lds.a.hi[row][4 * c]) = hi;
      *reinterpret_cast<ushort4 *>(&lds.a.lo[row][4 * c]) = lo;
    }
  }
  __syncthreads();
  const uint4 *Bth = reinterpret_cast<const uint4 *>(P == 1 ? t.j.bthi[S] : t.j.fthi[S]);
  const uint4 *Btl = reinterpret_cast<const uint4 *>(P == 1 ? t.j.btlo[S] : t.j.ftlo[S]);
  const size_t k8 = (size_t)K / 8;
  const uint4 *bh0 = Bth + (size_t)(n0 + r) * k8 + q, *bh1 = Bth + (size_t)(n0 + 16 + r) * k8 + q;
  const uint4 *bl0 = Btl + (size_t)(n0 + r) * k8 + q, *bl1 = Btl + (size_t)(n0 + 16 + r) * k8 + q;
  float4_t v[2][2];
#pragma unroll
  for (int c = 0; c < SPLIT; ++c) {
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k = 32 * c; k < K; k += 32 * SPLIT) {
      const uint4 bhi[2] = {bh0[k / 8], bh1[k / 8]};
      const uint4 blo[2] = {bl0[k / 8], bl1[k / 8]};
      uint4 ahi[2], alo[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ahi[i] = *reinterpret_cast<const uint4 *>(&lds.a.hi[16 * i + r][k + 8 * q]);
        alo[i] = *reinterpret_cast<const uint4 *>(&lds.a.lo[16 * i + r][k + 8 * q]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = mfma32<P>(ahi[i], bhi[j], acc[i][j]);
          acc[i][j] = mfma32<P>(ahi[i], blo[j], acc[i][j]);
          acc[i][j] = mfma32<P>(alo[i], bhi[j], acc[i][j]);
        }
    }
    // the 32x32 kernel sums its waves' scaled partials in wave order
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) v[i][j] = c == 0 ? acc[i][j] * unscale : v[i][j] + acc[i][j] * unscale;
  }
  __syncthreads();  // every wave is done with A in LDS
  // accumulator lane l holds D[4 (l / 16) + u][l % 16] of each 16x16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int u = 0; u < 4; ++u) lds.red[wave][16 * i + 4 * q + u][16 * j + r] = v[i][j][u];
  __syncthreads();
  // quarter h of the 32x32 kernel's block: its thread 64 h + lane owns
  // column lane % 32 and rows 4 (2 h + lane / 32) .. + 3
  const int gx = e.N / 32, bx = n0 / 32, col = lane & 31;
  double se[SPLIT][MAXF];
  unsigned int mx = 0u;
#pragma unroll
  for (int h = 0; h < SPLIT; ++h) {
    const int row0 = 4 * (2 * h + (lane >> 5));
    float4_t w;
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = lds.red[wave][row0 + u][col];
    double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
    store4(e, m0 + row0, n0 + col, w, energy);
    mx = max(mx, max(max(abs_bits(w[0]), abs_bits(w[1])), max(abs_bits(w[2]), abs_bits(w[3]))));
#pragma unroll
    for (int k = 0; k < MAXF; ++k) se[h][k] = e.energy ? wave_sum(energy[k]) : 0.0;
  }
  if (e.amax_out) {
    mx = lane_reduce<64>(mx, OpMaxI{});
    if (lane == 0) e.amax_out[(size_t)rb * gx + bx] = mx;
  }
  if (e.energy && lane < MAXF) {
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < MAXF; ++k)
      if (k == lane) tot = ((se[0][k] + se[1][k]) + se[2][k]) + se[3][k];
    e.energy[((size_t)rb * gx + bx) * MAXF + lane] = tot;
  }
}

// FP32 vector-ALU: each wave owns the 32-wide k-chunks w, w+4, ..., stages
// them through its own LDS tiles and accumulates 4x4 outputs per lane.
template <int S>
__global__ void __launch_bounds__(FAST_THREADS) k_gemm_f32_fast(const Task *__restrict__ tasks,
                                                                BlockPlan p) {
  __shared__ float As[SPLIT][32][33];
  __shared__ float Bs[SPLIT][32][36];
  __shared__ float red[SPLIT][32][33];
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's tiles: 32-row blocks x 32-column tiles, row-major
  const int gx = (e.N + 31) / 32, rb = bm.y / gx, bx = bm.y - rb * gx;
  const float *B = t.j.b32[S];
  const int K = e.K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tr = lane >> 3, tc = lane & 7;
  const int m0 = rb * 32, n0 = bx * 32;
  float acc[4][4] = {};
  // every wave runs the same number of rounds so the block barriers match
  const int chunks = K / 32, rounds = (chunks + SPLIT - 1) / SPLIT;
  for (int c = 0; c < rounds; ++c) {
    const int chunk = wave + SPLIT * c;
    const bool active = chunk < chunks;
    const int k0 = 32 * chunk;
    if (active) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int idx = lane + 64 * s;
        const int rr = idx >> 3, cc = (idx & 7) * 4;
        const float4 a = *reinterpret_cast<const float4 *>(e.A + (size_t)(m0 + rr) * K + k0 + cc);
        As[wave][rr][cc] = a.x;
        As[wave][rr][cc + 1] = a.y;
        As[wave][rr][cc + 2] = a.z;
        As[wave][rr][cc + 3] = a.w;
        *reinterpret_cast<float4 *>(&Bs[wave][rr][cc]) =
            *reinterpret_cast<const float4 *>(B + (size_t)(k0 + rr) * e.N + n0 + cc);
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll 8
      for (int kk = 0; kk < 32; ++kk) {
        const float4 b = *reinterpret_cast<const float4 *>(&Bs[wave][kk][4 * tc]);
        const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float a = As[wave][4 * tr + i][kk];
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a, bv[j], acc[i][j]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) red[wave][4 * tr + i][4 * tc + j] = acc[i][j];
  __syncthreads();
  fast_epilogue(e, red, m0, n0, rb, bx, gx);
}

// ---------------------------------------------------------------------------
// Any grid size: 32x32 tiles with bounds checks (prime widths, 7x5 tests).
// ---------------------------------------------------------------------------
template <int S>
__global__ void __launch_bounds__(64) k_gemm_bf16_any(const Task *__restrict__ tasks, BlockPlan p) {
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's tiles: 32-row blocks x 32-column tiles, row-major
  const int gx = (e.N + 31) / 32, rb = bm.y / gx, bx = bm.y - rb * gx;
  // the transposed split basis: element (k, n) at [n * K + k]
  const unsigned short *Bth = t.j.bthi[S], *Btl = t.j.btlo[S];
  const int K = e.K;
  const int lane = threadIdx.x;
  const int m0 = rb * 32, n0 = bx * 32;
  const int q = lane >> 4, r = lane & 15;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    short4_t ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = m0 + 16 * i + r, col = n0 + 16 * i + r;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = k0 + 4 * q + s;
        const float v = (row < e.M && k < K) ? e.A[(size_t)row * K + k] : 0.f;
        const unsigned short h = bf16_rne(v);
        ahi[i][s] = static_cast<short>(h);
        alo[i][s] = static_cast<short>(bf16_rne(v - bf16_to_float(h)));
        const bool in = col < e.N && k < K;
        bhi[i][s] = in ? static_cast<short>(Bth[(size_t)col * K + k]) : 0;
        blo[i][s] = in ? static_cast<short>(Btl[(size_t)col * K + k]) : 0;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[i], blo[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo[i], bhi[j], acc[i][j], 0, 0, 0);
      }
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int row = m0 + 16 * i + 4 * q + s, col = n0 + 16 * j + r;
        if (row < e.M && col < e.N) store_element(e, row, col, acc[i][j][s], energy);
      }
  write_energy(e, rb, bx, gx, energy);
}

// FP16-scaled split precision, any grid size (bounds-checked 32x32 tiles).
template <int S>
__global__ void __launch_bounds__(64) k_gemm_f16s_any(const Task *__restrict__ tasks, BlockPlan p) {
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's tiles: 32-row blocks x 32-column tiles, row-major
  const int gx = (e.N + 31) / 32, rb = bm.y / gx, bx = bm.y - rb * gx;
  const unsigned short *Bth = t.j.fthi[S], *Btl = t.j.ftlo[S];
  const int K = e.K;
  const int lane = threadIdx.x;
  const int m0 = rb * 32, n0 = bx * 32;
  const int q = lane >> 4, r = lane & 15;
  float amax = 0.f;
  for (int idx = lane; idx < 32 * K; idx += 64) {
    const int row = m0 + idx / K;
    if (row < e.M) amax = fmaxf(amax, fabsf(e.A[(size_t)row * K + idx % K]));
  }
  const float sa = pow2_scale(wave_max_f(amax));
  const float unscale = 1.f / (sa * t.j.f16_scale[S]);
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    half4_t ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = m0 + 16 * i + r, col = n0 + 16 * i + r;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = k0 + 4 * q + s;
        const float v = (row < e.M && k < K) ? e.A[(size_t)row * K + k] * sa : 0.f;
        _Float16 h, l;
        split_f16(v, h, l);
        ahi[i][s] = h;
        alo[i][s] = l;
        const bool in = col < e.N && k < K;
        unsigned short bh = in ? Bth[(size_t)col * K + k] : 0, bl = in ? Btl[(size_t)col * K + k] : 0;
        _Float16 fh, fl;
        __builtin_memcpy(&fh, &bh, 2);
        __builtin_memcpy(&fl, &bl, 2);
        bhi[i][s] = fh;
        blo[i][s] = fl;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi[i], blo[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16f16(alo[i], bhi[j], acc[i][j], 0, 0, 0);
      }
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int row = m0 + 16 * i + 4 * q + s, col = n0 + 16 * j + r;
        if (row < e.M && col < e.N) store_element(e, row, col, acc[i][j][s] * unscale, energy);
      }
  write_energy(e, rb, bx, gx, energy);
}

template <int S>
__global__ void __launch_bounds__(64) k_gemm_f32_any(const Task *__restrict__ tasks, BlockPlan p) {
  __shared__ float As[32][33];
  __shared__ float Bs[32][33];
  int2 bm;
  if (!plan_block(p, bm)) return;
  const Task &t = tasks[bm.x];
  const Stage e = stage_of<S>(t);
  // the task's tiles: 32-row blocks x 32-column tiles, row-major
  const int gx = (e.N + 31) / 32, rb = bm.y / gx, bx = bm.y - rb * gx;
  const float *B = t.j.b32[S];
  const int K = e.K;
  const int lane = threadIdx.x;
  const int tr = lane >> 3, tc = lane & 7;
  const int m0 = rb * 32, n0 = bx * 32;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int idx = lane; idx < 32 * 32; idx += 64) {
      const int rr = idx >> 5, cc = idx & 31;
      As[rr][cc] = (m0 + rr < e.M && k0 + cc < K) ? e.A[(size_t)(m0 + rr) * K + k0 + cc] : 0.f;
      Bs[rr][cc] = (k0 + rr < K && n0 + cc < e.N) ? B[(size_t)(k0 + rr) * e.N + n0 + cc] : 0.f;
    }
    __syncthreads();
    for (int kk = 0; kk < 32; ++kk) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float a = As[4 * tr + i][kk];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a, Bs[kk][4 * tc + j], acc[i][j]);
      }
    }
    __syncthreads();
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = m0 + 4 * tr + i, col = n0 + 4 * tc + j;
      if (row < e.M && col < e.N) store_element(e, row, col, acc[i][j], energy);
    }
  write_energy(e, rb, bx, gx, energy);
}

}  // namespace

bool poisson_fast_shape(int W, int H) { return W % 32 == 0 && H % 32 == 0; }
// split-MFMA precision with both extents multiples of the wide kernel's 128
// columns and within its LDS rows
bool poisson_wide_ok(int precision, int W, int H) {
  return (precision == 1 || precision == 2) && W % (32 * SPLIT) == 0 && H % (32 * SPLIT) == 0 &&
         W <= WIDE_KMAX && H <= WIDE_KMAX;
}

int poisson_energy_blocks(int W, int H, int F) {
  // stage 2: rows F*W in 32-row blocks, columns H in 32-column blocks
  return ((F * W + 31) / 32) * ((H + 31) / 32);
}

template <int S, int P>
void launch_wide(const Task *tasks, const BlockPlan &p, int kmax, hipStream_t s) {
  const dim3 grid(plan_grid(p));
  if (kmax <= 256)
    hipLaunchKernelGGL((k_gemm_wide<S, P, 256>), grid, dim3(FAST_THREADS), 0, s, tasks, p);
  else
    hipLaunchKernelGGL((k_gemm_wide<S, P, WIDE_KMAX>), grid, dim3(FAST_THREADS), 0, s, tasks, p);
}

template <int S>
void launch_stage(const Task *tasks, const DctLaunch &d, bool wide, hipStream_t s) {
  constexpr bool h_rows = S == 0 || S == 3;  // stages with F*H rows
  const BlockPlan &p = wide ? (h_rows ? d.h_wide : d.w_wide) : (h_rows ? d.h : d.w);
  const dim3 grid(plan_grid(p));
  if (wide) {
    if (d.precision == 1)
      launch_wide<S, 1>(tasks, p, d.kmax, s);
    else
      launch_wide<S, 2>(tasks, p, d.kmax, s);
  } else if (d.precision == 1) {
    if (d.fast)
      hipLaunchKernelGGL(k_gemm_bf16_fast<S>, grid, dim3(FAST_THREADS), 0, s, tasks, p);
    else
      hipLaunchKernelGGL(k_gemm_bf16_any<S>, grid, dim3(64), 0, s, tasks, p);
  } else if (d.precision == 2) {
    if (d.fast)
      hipLaunchKernelGGL(k_gemm_f16s_fast<S>, grid, dim3(FAST_THREADS), 0, s, tasks, p);
    else
      hipLaunchKernelGGL(k_gemm_f16s_any<S>, grid, dim3(64), 0, s, tasks, p);
  } else {
    if (d.fast)
      hipLaunchKernelGGL(k_gemm_f32_fast<S>, grid, dim3(FAST_THREADS), 0, s, tasks, p);
    else
      hipLaunchKernelGGL(k_gemm_f32_any<S>, grid, dim3(64), 0, s, tasks, p);
  }
}

void launch_poisson(const Task *tasks, const DctLaunch &d, hipStream_t s) {
  if (d.h.total <= 0) return;
  timer_begin(g_timer, "dct", s);
  const bool wide = d.wide_ok && d.h_wide.total >= WIDE_MIN_TILES;
  launch_stage<0>(tasks, d, wide, s);
  launch_stage<1>(tasks, d, wide, s);
  launch_stage<2>(tasks, d, wide, s);
  launch_stage<3>(tasks, d, wide, s);
  timer_end(g_timer, "dct", s, d.n_tasks);
}

}  // namespace texo
