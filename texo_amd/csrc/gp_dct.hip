// gp_dct.hip — Poisson solve of the electrostatic density as dense DCT
// contractions (Poisson2d::solve, electrostatic_placement.rs:725-776).
//
// phi = C_H^T [ (C_H rho C_W^T) .* invLambda ] C_W, written as four
// "stacked rows x basis" products so that every field (and every batched job)
// stacks along M and the basis stays a static K x N operand:
//   1. Rt_f  = (rho_f C_W^T)^T            M = F*H, K = N = W, transposed store
//   2. St'_f = (Rt_f C_H^T) .* invLambda^T M = F*W, K = N = H, spectral epilogue
//   3. T_f   = (St'_f C_H)^T               M = F*W, K = N = H, transposed store
//   4. phi_f = T_f C_W                     M = F*H, K = N = W
// The electrostatic energy 1/2 <rho_bar, phi> equals 1/2 sum invLambda S^2 by
// Parseval (orthonormal basis); it is accumulated in FP64 in the epilogue of
// step 2.  The DC coefficient has invLambda = 0, so removing the mean of rho
// (electrostatic_placement.rs:739-744) changes neither phi nor the energy.

#include "gp_common.h"

namespace texo {

namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ void block_sum_fields(double (&v)[MAXF], double *out) {
  __shared__ double s[MAXF][BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < MAXF; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < MAXF; ++k) s[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / 64;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
      double acc = 0.0;
      for (int w = 0; w < nw; ++w) acc += s[k][w];
      out[k] = acc;
    }
  }
}

struct Epilogue {
  int M, N, R;             // rows, cols, rows per field
  bool transposed;         // store Out_f[col][r]
  const float *inv_lambda; // spectral epilogue when non-null
  int spectral_width;      // W (layout ky*W + kx)
  double *energy;          // [blocks][MAXF]
};

__device__ __forceinline__ void store_element(const Epilogue &e, float *__restrict__ out, int row, int col, float v,
                                              double (&energy)[MAXF]) {
  const int f = row / e.R, r = row - f * e.R;
  if (e.inv_lambda) {
    const float lam = e.inv_lambda[(size_t)col * e.spectral_width + r];
    const double dv = static_cast<double>(v);
    const double contribution = 0.5 * dv * dv * static_cast<double>(lam);
#pragma unroll
    for (int k = 0; k < MAXF; ++k)
      if (k == f) energy[k] += contribution;
    v = v * lam;
  }
  if (e.transposed)
    out[(size_t)f * e.R * e.N + (size_t)col * e.R + r] = v;
  else
    out[(size_t)row * e.N + col] = v;
}

// ---------------------------------------------------------------------------
// FP32 vector-ALU contraction: 32x32 tile, 256 threads, 2x2 outputs each.
// ---------------------------------------------------------------------------
constexpr int T32 = 32;

__global__ void __launch_bounds__(BLOCK) k_gemm_f32(const float *__restrict__ A, const float *__restrict__ B,
                                                    float *__restrict__ out, int K, Epilogue e) {
  __shared__ float As[T32][T32 + 1];
  __shared__ float Bs[T32][T32 + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * T32, n0 = blockIdx.x * T32;
  float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
  for (int k0 = 0; k0 < K; k0 += T32) {
    for (int idx = threadIdx.x; idx < T32 * T32; idx += BLOCK) {
      const int r = idx / T32, c = idx % T32;
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < e.M && gk < K) ? A[(size_t)gm * K + gk] : 0.f;
      const int bk = k0 + r, bn = n0 + c;
      Bs[r][c] = (bk < K && bn < e.N) ? B[(size_t)bk * e.N + bn] : 0.f;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < T32; ++kk) {
      const float a0 = As[ty][kk], a1 = As[ty + 16][kk];
      const float b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
      acc[0][0] = fmaf(a0, b0, acc[0][0]);
      acc[0][1] = fmaf(a0, b1, acc[0][1]);
      acc[1][0] = fmaf(a1, b0, acc[1][0]);
      acc[1][1] = fmaf(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int row = m0 + ty + 16 * i, col = n0 + tx + 16 * j;
      if (row < e.M && col < e.N) store_element(e, out, row, col, acc[i][j], energy);
    }
  if (e.energy) block_sum_fields(energy, e.energy + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * MAXF);
}

// ---------------------------------------------------------------------------
// BF16x3 split-precision MFMA contraction.  A (dynamic, FP32) is split on the
// fly into hi + lo BF16; B (static basis) is pre-split.  Each 16x16x16 MFMA
// step issues hi*hi + hi*lo + lo*hi with FP32 accumulation.
// Block tile 64x64, four waves of 32x32 (2x2 MFMA tiles), K chunk 32.
//
// v_mfma_f32_16x16x16_bf16 operand layout (64 lanes): A lane l holds
// A[l%16][4*(l/16) + i], B lane l holds B[4*(l/16) + i][l%16], and the
// accumulator lane l holds D[4*(l/16) + i][l%16], i = 0..3.
// ---------------------------------------------------------------------------
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf16_to_float(unsigned short h) {
  return __uint_as_float(static_cast<unsigned int>(h) << 16);
}

constexpr int TM = 64, TN = 64, TK = 32;

template <bool SPLIT>
__global__ void __launch_bounds__(BLOCK) k_gemm_bf16(const float *__restrict__ A,
                                                     const unsigned short *__restrict__ Bhi,
                                                     const unsigned short *__restrict__ Blo,
                                                     float *__restrict__ out, int K, Epilogue e) {
  __shared__ float As[TM][TK + 1];
  __shared__ unsigned short Bh[TK][TN + 2];
  __shared__ unsigned short Bl[TK][TN + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int idx = threadIdx.x; idx < TM * TK; idx += BLOCK) {
      const int r = idx / TK, c = idx % TK;
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < e.M && gk < K) ? A[(size_t)gm * K + gk] : 0.f;
    }
    for (int idx = threadIdx.x; idx < TK * TN; idx += BLOCK) {
      const int r = idx / TN, c = idx % TN;
      const int bk = k0 + r, bn = n0 + c;
      const bool in = bk < K && bn < e.N;
      Bh[r][c] = in ? Bhi[(size_t)bk * e.N + bn] : 0;
      if (SPLIT) Bl[r][c] = in ? Blo[(size_t)bk * e.N + bn] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TK; ks += 16) {
      const int kq = ks + 4 * (lane >> 4);
      short4_t ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm + 16 * i + (lane & 15);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float v = As[r][kq + t];
          const unsigned short h = bf16_rne(v);
          ahi[i][t] = static_cast<short>(h);
          alo[i][t] = static_cast<short>(bf16_rne(v - bf16_to_float(h)));
        }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = wn + 16 * j + (lane & 15);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          bhi[j][t] = static_cast<short>(Bh[kq + t][c]);
          blo[j][t] = SPLIT ? static_cast<short>(Bl[kq + t][c]) : 0;
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
          if (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[i], blo[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo[i], bhi[j], acc[i][j], 0, 0, 0);
          }
        }
    }
    __syncthreads();
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + t;
        const int col = n0 + wn + 16 * j + (lane & 15);
        if (row < e.M && col < e.N) store_element(e, out, row, col, acc[i][j][t], energy);
      }
  if (e.energy) block_sum_fields(energy, e.energy + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * MAXF);
}

// ---------------------------------------------------------------------------
// Fast paths for M % 32 == 0, N % 32 == 0, K % 16 == 0 (every power-of-two
// grid from 32 up): one wavefront computes a 32x32 output tile, operand
// fragments are loaded straight from global memory (L2-resident: the basis
// is static and the activations were just written) into registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store4(const Epilogue &e, float *__restrict__ out, int row0, int col, float4_t v,
                                       double (&energy)[MAXF]) {
  // rows row0..row0+3 of column col; row0 % 4 == 0 and R % 4 == 0 keep the
  // four rows inside one field
  const int f = row0 / e.R, r0 = row0 - f * e.R;
  if (e.inv_lambda) {
    const float4 lam = *reinterpret_cast<const float4 *>(e.inv_lambda + (size_t)col * e.spectral_width + r0);
    const float l[4] = {lam.x, lam.y, lam.z, lam.w};
    double c = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double dv = static_cast<double>(v[t]);
      c += 0.5 * dv * dv * static_cast<double>(l[t]);
      v[t] = v[t] * l[t];
    }
#pragma unroll
    for (int k = 0; k < MAXF; ++k)
      if (k == f) energy[k] += c;
  }
  if (e.transposed) {
    *reinterpret_cast<float4 *>(out + (size_t)f * e.R * e.N + (size_t)col * e.R + r0) =
        make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int t = 0; t < 4; ++t) out[(size_t)(row0 + t) * e.N + col] = v[t];
  }
}

template <bool SPLIT>
__global__ void __launch_bounds__(64) k_gemm_bf16_fast(const float *__restrict__ A, const uint2 *__restrict__ Bth,
                                                       const uint2 *__restrict__ Btl, float *__restrict__ out, int K,
                                                       Epilogue e) {
  const int lane = threadIdx.x;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int q = lane >> 4, r = lane & 15;
  const float *a0 = A + (size_t)(m0 + r) * K + 4 * q;
  const float *a1 = A + (size_t)(m0 + 16 + r) * K + 4 * q;
  const size_t k4 = (size_t)K / 4;
  const uint2 *bh0 = Bth + (size_t)(n0 + r) * k4 + q, *bh1 = Bth + (size_t)(n0 + 16 + r) * k4 + q;
  const uint2 *bl0 = Btl + (size_t)(n0 + r) * k4 + q, *bl1 = Btl + (size_t)(n0 + 16 + r) * k4 + q;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int k0 = 0; k0 < K; k0 += 16) {
    const float4 av[2] = {*reinterpret_cast<const float4 *>(a0 + k0), *reinterpret_cast<const float4 *>(a1 + k0)};
    const uint2 bhv[2] = {bh0[k0 / 4], bh1[k0 / 4]};
    uint2 blv[2] = {make_uint2(0, 0), make_uint2(0, 0)};
    if (SPLIT) {
      blv[0] = bl0[k0 / 4];
      blv[1] = bl1[k0 / 4];
    }
    short4_t ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float f4[4] = {av[i].x, av[i].y, av[i].z, av[i].w};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const unsigned short h = bf16_rne(f4[t]);
        ahi[i][t] = static_cast<short>(h);
        alo[i][t] = static_cast<short>(bf16_rne(f4[t] - bf16_to_float(h)));
      }
      bhi[i] = short4_t{static_cast<short>(bhv[i].x & 0xffff), static_cast<short>(bhv[i].x >> 16),
                        static_cast<short>(bhv[i].y & 0xffff), static_cast<short>(bhv[i].y >> 16)};
      blo[i] = short4_t{static_cast<short>(blv[i].x & 0xffff), static_cast<short>(blv[i].x >> 16),
                        static_cast<short>(blv[i].y & 0xffff), static_cast<short>(blv[i].y >> 16)};
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[i], bhi[j], acc[i][j], 0, 0, 0);
        if (SPLIT) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[i], blo[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo[i], bhi[j], acc[i][j], 0, 0, 0);
        }
      }
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) store4(e, out, m0 + 16 * i + 4 * q, n0 + 16 * j + r, acc[i][j], energy);
  if (e.energy) {
#pragma unroll
    for (int k = 0; k < MAXF; ++k) energy[k] = wave_sum(energy[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < MAXF; ++k) e.energy[(size_t)(blockIdx.y * gridDim.x + blockIdx.x) * MAXF + k] = energy[k];
    }
  }
}

// FP32 vector-ALU fast path: one wavefront per 32x32 tile, 4x4 outputs per
// lane, K staged through LDS in chunks of 32.
__global__ void __launch_bounds__(64) k_gemm_f32_fast(const float *__restrict__ A, const float *__restrict__ B,
                                                      float *__restrict__ out, int K, Epilogue e) {
  __shared__ float As[32][33];
  __shared__ float Bs[32][36];
  const int lane = threadIdx.x;
  const int tr = lane >> 3, tc = lane & 7;  // rows 4*tr.., cols 4*tc..
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int idx = lane + 64 * t;  // 256 float4 per 32x32 tile
      const int rr = idx >> 3, cc = (idx & 7) * 4;
      const float4 a = *reinterpret_cast<const float4 *>(A + (size_t)(m0 + rr) * K + k0 + cc);
      As[rr][cc] = a.x;
      As[rr][cc + 1] = a.y;
      As[rr][cc + 2] = a.z;
      As[rr][cc + 3] = a.w;
      const float4 b = *reinterpret_cast<const float4 *>(B + (size_t)(k0 + rr) * e.N + n0 + cc);
      *reinterpret_cast<float4 *>(&Bs[rr][cc]) = b;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; ++kk) {
      const float4 b = *reinterpret_cast<const float4 *>(&Bs[kk][4 * tc]);
      const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float a = As[4 * tr + i][kk];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a, bv[j], acc[i][j]);
      }
    }
    __syncthreads();
  }
  double energy[MAXF] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float4_t v = float4_t{acc[0][j], acc[1][j], acc[2][j], acc[3][j]};
    store4(e, out, m0 + 4 * tr, n0 + 4 * tc + j, v, energy);
  }
  if (e.energy) {
#pragma unroll
    for (int k = 0; k < MAXF; ++k) energy[k] = wave_sum(energy[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < MAXF; ++k) e.energy[(size_t)(blockIdx.y * gridDim.x + blockIdx.x) * MAXF + k] = energy[k];
    }
  }
}

bool fast_shape(int M, int K, int N, int R) { return M % 32 == 0 && N % 32 == 0 && K % 32 == 0 && R % 4 == 0; }

void stage(const DctPlan &p, int which, const float *A, float *out, int M, int K, int N, int R, bool transposed,
           const float *inv_lambda, double *energy, int *nb, hipStream_t s) {
  Epilogue e;
  e.M = M;
  e.N = N;
  e.R = R;
  e.transposed = transposed;
  e.inv_lambda = inv_lambda;
  e.spectral_width = p.W;
  e.energy = energy;
  if (fast_shape(M, K, N, R)) {
    dim3 grid(N / 32, M / 32);
    if (nb) *nb = grid.x * grid.y;
    if (p.precision == 1)
      hipLaunchKernelGGL(k_gemm_bf16_fast<true>, grid, dim3(64), 0, s, A,
                         reinterpret_cast<const uint2 *>(p.bthi[which]), reinterpret_cast<const uint2 *>(p.btlo[which]),
                         out, K, e);
    else
      hipLaunchKernelGGL(k_gemm_f32_fast, grid, dim3(64), 0, s, A, p.b32[which], out, K, e);
    return;
  }
  if (p.precision == 1) {  // BF16x3
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    if (nb) *nb = grid.x * grid.y;
    hipLaunchKernelGGL(k_gemm_bf16<true>, grid, dim3(BLOCK), 0, s, A, p.bhi[which], p.blo[which], out, K, e);
  } else {
    dim3 grid((N + T32 - 1) / T32, (M + T32 - 1) / T32);
    if (nb) *nb = grid.x * grid.y;
    hipLaunchKernelGGL(k_gemm_f32, grid, dim3(BLOCK), 0, s, A, p.b32[which], out, K, e);
  }
}

}  // namespace

int poisson_energy_blocks(const DctPlan &p) {
  const int M = p.F * p.W, N = p.H, K = p.H, R = p.W;
  if (fast_shape(M, K, N, R)) return (N / 32) * (M / 32);
  if (p.precision == 1) return ((N + TN - 1) / TN) * ((M + TM - 1) / TM);
  return ((N + T32 - 1) / T32) * ((M + T32 - 1) / T32);
}

void launch_poisson(const DctPlan &p, const float *rho, float *phi, double *energy_partials, int *nb_energy,
                    hipStream_t s) {
  const int W = p.W, H = p.H, F = p.F;
  timer_begin(g_timer, "dct", s);
  stage(p, 0, rho, p.t0, F * H, W, W, H, true, nullptr, nullptr, nullptr, s);
  stage(p, 1, p.t0, p.t1, F * W, H, H, W, false, p.inv_lambda, energy_partials, nb_energy, s);
  stage(p, 2, p.t1, p.t0, F * W, H, H, W, true, nullptr, nullptr, nullptr, s);
  stage(p, 3, p.t0, phi, F * H, W, W, H, false, nullptr, nullptr, nullptr, s);
  timer_end(g_timer, "dct", s);
}

}  // namespace texo
