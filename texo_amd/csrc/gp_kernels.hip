// gp_kernels.hip — scatter/gather and elementwise kernels of the electrostatic
// global-placement loop for gfx950 (64-lane wavefronts).
//
// Every kernel here is HBM/latency bound; reductions are deterministic:
// integer sums use int64 atomics (exact, order-free), maxima use atomicMax on
// the bit patterns of non-negative doubles (order-free), and FP64 sums go
// through per-block partials reduced in a fixed order by a single block.

#include "gp_common.h"

#include <cmath>

namespace texo {

namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide sum of K values per thread; thread 0 writes out[0..K).  The
// reduction tree is fixed by the block shape, so results are reproducible.
template <int K>
__device__ void block_sum(double (&v)[K], double *out) {
  __shared__ double s[K][BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) s[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) / 64;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double acc = 0.0;
      for (int w = 0; w < nw; ++w) acc += s[k][w];
      out[k] = acc;
    }
  }
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *p, double v) {
  atomicMax(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}

// Block-wide maximum of a non-negative value and OR of a flag; thread 0
// writes out[0] = max, out[1] = flag (as 0.0 / 1.0).  Used instead of
// same-address atomics, which serialise across thousands of waves.
__device__ void block_max_flag(double v, bool flag, double *out) {
  __shared__ double s[BLOCK / 64];
  __shared__ int f[BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_max(v);
  const bool any = __any(flag);
  if (lane == 0) {
    s[wave] = v;
    f[wave] = any ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    int a = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      m = fmax(m, s[w]);
      a |= f[w];
    }
    out[0] = m;
    out[1] = a ? 1.0 : 0.0;
  }
}

// kkt_component (global_placement.rs:675-689) divided by sqrt(metric).
__device__ __forceinline__ double kkt_scaled(double g, double x, double lo, double hi, double m) {
  const bool equal = __double_as_longlong(lo) == __double_as_longlong(hi);
  if (equal || (x <= lo && g > 0.0) || (x >= hi && g < 0.0)) return 0.0;
  return fabs(g) / sqrt(m);
}

// ---------------------------------------------------------------------------
// Density deposit (electrostatic_placement.rs:296-332, 900-916).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_deposit_init(int total, const long long *__restrict__ base,
                                                        long long *__restrict__ acc_m,
                                                        long long *__restrict__ acc_f, EvalScalars *scalars) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < total) {
    acc_m[i] = base[i];
    acc_f[i] = 0;
  }
  if (i == 0) {
    for (int f = 0; f < MAXF; ++f) {
      scalars->over_fx[f] = 0;
      scalars->net_fx[f] = 0;
    }
    scalars->kkt_bits = 0;
    scalars->nonfinite = 0;
  }
}

__global__ void k_zero_eval(EvalScalars *scalars) {
  for (int f = 0; f < MAXF; ++f) {
    scalars->over_fx[f] = 0;
    scalars->net_fx[f] = 0;
  }
  scalars->kkt_bits = 0;
  scalars->nonfinite = 0;
}

__global__ void __launch_bounds__(BLOCK) k_deposit(DensityArgs a, const double *__restrict__ coords) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double x, y, q;
  int f;
  long long *acc;
  if (i < a.n_members) {
    const uint32_t u = a.m_unit[i];
    const double2 off = a.m_off[i];
    x = coords[2 * (size_t)u] + off.x;
    y = coords[2 * (size_t)u + 1] + off.y;
    q = a.m_charge[i];
    f = a.m_field[i];
    acc = a.acc_m;
  } else if (i < a.n_members + a.n_fillers) {
    const int j = i - a.n_members;
    const size_t c = 2 * ((size_t)a.n_units + j);
    x = coords[c];
    y = coords[c + 1];
    q = a.fl_charge[j];
    f = a.fl_field[j];
    acc = a.acc_f;
  } else {
    return;
  }
  Axis ax, ay;
  bspline_axis(a.W, x, ax);
  bspline_axis(a.H, y, ay);
  unsigned long long *g =
      reinterpret_cast<unsigned long long *>(acc + (size_t)f * (size_t)a.W * (size_t)a.H);
  const int ix[3] = {ax.i0, ax.i1, ax.i2};
  const int iy[3] = {ay.i0, ay.i1, ay.i2};
  const double wx[3] = {ax.w0, ax.w1, ax.w2};
  const double wy[3] = {ay.w0, ay.w1, ay.w2};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double v = q * wx[c] * wy[r];
      const long long fx = __double2ll_rn(v * FX_SCALE);
      atomicAdd(g + (size_t)iy[r] * a.W + ix[c], static_cast<unsigned long long>(fx));
    }
  }
}

// rho = members + fillers (fixed point) -> FP32 transform input; physical
// overflow (before fillers) and net charge as exact int64 sums.
__global__ void __launch_bounds__(BLOCK) k_density_convert(int G, int F, const long long *__restrict__ acc_m,
                                                           const long long *__restrict__ acc_f,
                                                           float *__restrict__ rho, double *__restrict__ density_out,
                                                           EvalScalars *scalars) {
  __shared__ unsigned long long s_over[MAXF], s_net[MAXF];
  if (threadIdx.x < MAXF) {
    s_over[threadIdx.x] = 0;
    s_net[threadIdx.x] = 0;
  }
  __syncthreads();
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < F * G) {
    const int f = i / G;
    const long long m = acc_m[i];
    const long long t = m + acc_f[i];
    const double v = static_cast<double>(t) * FX_INV;
    rho[i] = static_cast<float>(v);
    if (density_out) density_out[i] = v;
    if (m > 0) atomicAdd(&s_over[f], static_cast<unsigned long long>(m));
    atomicAdd(&s_net[f], static_cast<unsigned long long>(t));
  }
  __syncthreads();
  if (threadIdx.x < F) {
    atomicAdd(reinterpret_cast<unsigned long long *>(&scalars->over_fx[threadIdx.x]), s_over[threadIdx.x]);
    atomicAdd(reinterpret_cast<unsigned long long *>(&scalars->net_fx[threadIdx.x]), s_net[threadIdx.x]);
  }
}

// ---------------------------------------------------------------------------
// Field gradient gather (electrostatic_placement.rs:338-374, 919-941).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_coefficients(const GatherArgs &g, int F, double *coef) {
  // Each block re-reduces the energy partials of the transform in the same
  // fixed order, then forms lambda (1 + beta E/E0) (eplace.rs:1111-1124).
  if (threadIdx.x < F) {
    const int f = threadIdx.x;
    double e = 0.0;
    for (int b = 0; b < g.nb_energy; ++b) e += g.energy_partials[b * MAXF + f];
    const double normalized = e / g.scale[f];
    coef[f] = g.lambda[f] * (1.0 + AUGMENTED_DENSITY_BETA * normalized);
  }
  __syncthreads();
}

__device__ __forceinline__ void sample(const float *__restrict__ phi, int W, int H, double x, double y, double &gx,
                                       double &gy) {
  Axis ax, ay;
  bspline_axis(W, x, ax);
  bspline_axis(H, y, ay);
  const int ix[3] = {ax.i0, ax.i1, ax.i2};
  const int iy[3] = {ay.i0, ay.i1, ay.i2};
  const double wx[3] = {ax.w0, ax.w1, ax.w2};
  const double wy[3] = {ay.w0, ay.w1, ay.w2};
  const double dx[3] = {ax.d0, ax.d1, ax.d2};
  const double dy[3] = {ay.d0, ay.d1, ay.d2};
  gx = 0.0;
  gy = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double p = static_cast<double>(phi[(size_t)iy[r] * W + ix[c]]);
      gx += dx[c] * wy[r] * p;
      gy += wx[c] * dy[r] * p;
    }
  }
}

__global__ void __launch_bounds__(BLOCK) k_gather_units(DensityArgs a, GatherArgs g, const double *__restrict__ coords,
                                                        double2 *__restrict__ gfu, double *__restrict__ raw_units,
                                                        double *__restrict__ partials) {
  __shared__ double coef[MAXF];
  load_coefficients(g, a.F, coef);
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const size_t G = (size_t)a.W * a.H;
  double red[2 * MAXF];
#pragma unroll
  for (int k = 0; k < 2 * MAXF; ++k) red[k] = 0.0;
  if (i < a.n_density_units) {
    const uint32_t u = a.du_unit[i];
    const double ox = coords[2 * (size_t)u], oy = coords[2 * (size_t)u + 1];
    double gxf[MAXF], gyf[MAXF];
#pragma unroll
    for (int f = 0; f < MAXF; ++f) gxf[f] = gyf[f] = 0.0;
    for (uint32_t m = a.du_mstart[i]; m < a.du_mstart[i + 1]; ++m) {
      const int f = a.m_field[m];
      const double2 off = a.m_off[m];
      const double q = a.m_charge[m];
      double gx, gy;
      sample(g.phi + f * G, a.W, a.H, ox + off.x, oy + off.y, gx, gy);
#pragma unroll
      for (int k = 0; k < MAXF; ++k) {
        if (k == f) {
          red[k] += q * (fabs(gx) + fabs(gy));
          gxf[k] += q * gx;
          gyf[k] += q * gy;
        }
      }
    }
#pragma unroll
    for (int f = 0; f < MAXF; ++f) {
      if (f < a.F) {
        red[MAXF + f] = fabs(gxf[f]) + fabs(gyf[f]);
        gfu[(size_t)f * a.n_density_units + i] = make_double2(coef[f] * gxf[f], coef[f] * gyf[f]);
        if (raw_units) {
          raw_units[2 * ((size_t)f * a.n_density_units + i)] = gxf[f];
          raw_units[2 * ((size_t)f * a.n_density_units + i) + 1] = gyf[f];
        }
      }
    }
  }
  block_sum<2 * MAXF>(red, partials + (size_t)blockIdx.x * 2 * MAXF);
}

__global__ void __launch_bounds__(BLOCK) k_gather_fillers(DensityArgs a, GatherArgs g,
                                                          const double *__restrict__ coords,
                                                          double *__restrict__ raw_fillers, double *__restrict__ grad,
                                                          const double *__restrict__ metric,
                                                          const double *__restrict__ lower,
                                                          const double *__restrict__ upper,
                                                          double *__restrict__ partials, EvalScalars *scalars) {
  __shared__ double coef[MAXF];
  load_coefficients(g, a.F, coef);
  const int j = blockIdx.x * BLOCK + threadIdx.x;
  const size_t G = (size_t)a.W * a.H;
  double red[2 * MAXF];
#pragma unroll
  for (int k = 0; k < 2 * MAXF; ++k) red[k] = 0.0;
  double kkt = 0.0;
  bool bad = false;
  if (j < a.n_fillers) {
    const size_t c = 2 * ((size_t)a.n_units + j);
    const double x = coords[c], y = coords[c + 1];
    const int f = a.fl_field[j];
    const double q = a.fl_charge[j];
    double gx, gy;
    sample(g.phi + f * G, a.W, a.H, x, y, gx, gy);
    const double rx = q * gx, ry = q * gy;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
      if (k == f) {
        red[k] = q * (fabs(gx) + fabs(gy));
        red[MAXF + k] = fabs(rx) + fabs(ry);
      }
    }
    if (raw_fillers) {
      raw_fillers[2 * (size_t)j] = rx;
      raw_fillers[2 * (size_t)j + 1] = ry;
    }
    const double vx = 0.0 + coef[f] * rx, vy = 0.0 + coef[f] * ry;
    grad[c] = vx;
    grad[c + 1] = vy;
    bad = !isfinite(vx) || !isfinite(vy);
    if (metric) {
      kkt = fmax(kkt_scaled(vx, x, lower[c], upper[c], metric[c]),
                 kkt_scaled(vy, y, lower[c + 1], upper[c + 1], metric[c + 1]));
    }
  }
  block_sum<2 * MAXF>(red, partials + (size_t)blockIdx.x * GF_STRIDE);
  block_max_flag(kkt, bad, partials + (size_t)blockIdx.x * GF_STRIDE + 2 * MAXF);
}

// ---------------------------------------------------------------------------
// Weighted-average wirelength (analytical_placement.rs:576-703), FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 pin_coordinate(const WaArgs &a, const double *__restrict__ coords, uint32_t p) {
  const uint32_t u = a.pin_unit[p];
  const double2 v = a.pin_xy[p];
  if (u == NO_UNIT) return v;
  return make_double2(coords[2 * (size_t)u] + v.x, coords[2 * (size_t)u + 1] + v.y);
}

__device__ void wa_thread_nets(const WaArgs &a, const double *__restrict__ coords, double gamma,
                               double *__restrict__ partials, int block) {
  const int t = block * BLOCK + threadIdx.x;
  double value[1] = {0.0};
  if (t < a.n_small) {
    const uint32_t net = a.small_nets[t];
    const uint32_t b = a.net_start[net], e = a.net_start[net + 1];
    const double w = a.net_weight[net];
    double2 first = pin_coordinate(a, coords, b);
    double xmin = first.x, xmax = first.x, ymin = first.y, ymax = first.y;
    for (uint32_t p = b + 1; p < e; ++p) {
      const double2 c = pin_coordinate(a, coords, p);
      xmin = fmin(xmin, c.x);
      xmax = fmax(xmax, c.x);
      ymin = fmin(ymin, c.y);
      ymax = fmax(ymax, c.y);
    }
    double psx = 0, pdx = 0, nsx = 0, ndx = 0, psy = 0, pdy = 0, nsy = 0, ndy = 0;
    for (uint32_t p = b; p < e; ++p) {
      const double2 c = pin_coordinate(a, coords, p);
      const double dxm = c.x - xmax, dxn = c.x - xmin, dym = c.y - ymax, dyn = c.y - ymin;
      const double ex = dxm == 0.0 ? 1.0 : exp(dxm / gamma);
      const double enx = dxn == 0.0 ? 1.0 : exp(-dxn / gamma);
      const double ey = dym == 0.0 ? 1.0 : exp(dym / gamma);
      const double eny = dyn == 0.0 ? 1.0 : exp(-dyn / gamma);
      psx += ex;
      pdx += ex * dxm;
      nsx += enx;
      ndx += enx * dxn;
      psy += ey;
      pdy += ey * dym;
      nsy += eny;
      ndy += eny * dyn;
    }
    const double pmx = pdx / psx, nmx = ndx / nsx, pmy = pdy / psy, nmy = ndy / nsy;
    for (uint32_t p = b; p < e; ++p) {
      const double2 c = pin_coordinate(a, coords, p);
      const double dxm = c.x - xmax, dxn = c.x - xmin, dym = c.y - ymax, dyn = c.y - ymin;
      const double ex = dxm == 0.0 ? 1.0 : exp(dxm / gamma);
      const double enx = dxn == 0.0 ? 1.0 : exp(-dxn / gamma);
      const double ey = dym == 0.0 ? 1.0 : exp(dym / gamma);
      const double eny = dyn == 0.0 ? 1.0 : exp(-dyn / gamma);
      const double ddx = (ex / psx) * (1.0 + (dxm - pmx) / gamma) - (enx / nsx) * (1.0 - (dxn - nmx) / gamma);
      const double ddy = (ey / psy) * (1.0 + (dym - pmy) / gamma) - (eny / nsy) * (1.0 - (dyn - nmy) / gamma);
      a.dpin[p] = make_double2(w * ddx, w * ddy);
    }
    value[0] = w * ((xmax - xmin) + pmx - nmx) + w * ((ymax - ymin) + pmy - nmy);
  }
  block_sum<1>(value, partials + blockIdx.x);
}

// One wavefront per net for high-fanout nets.
__device__ void wa_wave_nets(const WaArgs &a, const double *__restrict__ coords, double gamma,
                             double *__restrict__ partials, int block) {
  const int lane = threadIdx.x & 63;
  const int t = block * (BLOCK / 64) + (threadIdx.x >> 6);
  double value[1] = {0.0};
  if (t < a.n_large) {
    const uint32_t net = a.large_nets[t];
    const uint32_t b = a.net_start[net], e = a.net_start[net + 1];
    const double w = a.net_weight[net];
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint32_t p = b + lane; p < e; p += 64) {
      const double2 c = pin_coordinate(a, coords, p);
      xmin = fmin(xmin, c.x);
      xmax = fmax(xmax, c.x);
      ymin = fmin(ymin, c.y);
      ymax = fmax(ymax, c.y);
    }
    xmin = wave_min(xmin);
    xmax = wave_max(xmax);
    ymin = wave_min(ymin);
    ymax = wave_max(ymax);
    double psx = 0, pdx = 0, nsx = 0, ndx = 0, psy = 0, pdy = 0, nsy = 0, ndy = 0;
    for (uint32_t p = b + lane; p < e; p += 64) {
      const double2 c = pin_coordinate(a, coords, p);
      const double dxm = c.x - xmax, dxn = c.x - xmin, dym = c.y - ymax, dyn = c.y - ymin;
      const double ex = dxm == 0.0 ? 1.0 : exp(dxm / gamma);
      const double enx = dxn == 0.0 ? 1.0 : exp(-dxn / gamma);
      const double ey = dym == 0.0 ? 1.0 : exp(dym / gamma);
      const double eny = dyn == 0.0 ? 1.0 : exp(-dyn / gamma);
      psx += ex;
      pdx += ex * dxm;
      nsx += enx;
      ndx += enx * dxn;
      psy += ey;
      pdy += ey * dym;
      nsy += eny;
      ndy += eny * dyn;
    }
    psx = wave_sum(psx);
    pdx = wave_sum(pdx);
    nsx = wave_sum(nsx);
    ndx = wave_sum(ndx);
    psy = wave_sum(psy);
    pdy = wave_sum(pdy);
    nsy = wave_sum(nsy);
    ndy = wave_sum(ndy);
    const double pmx = pdx / psx, nmx = ndx / nsx, pmy = pdy / psy, nmy = ndy / nsy;
    for (uint32_t p = b + lane; p < e; p += 64) {
      const double2 c = pin_coordinate(a, coords, p);
      const double dxm = c.x - xmax, dxn = c.x - xmin, dym = c.y - ymax, dyn = c.y - ymin;
      const double ex = dxm == 0.0 ? 1.0 : exp(dxm / gamma);
      const double enx = dxn == 0.0 ? 1.0 : exp(-dxn / gamma);
      const double ey = dym == 0.0 ? 1.0 : exp(dym / gamma);
      const double eny = dyn == 0.0 ? 1.0 : exp(-dyn / gamma);
      const double ddx = (ex / psx) * (1.0 + (dxm - pmx) / gamma) - (enx / nsx) * (1.0 - (dxn - nmx) / gamma);
      const double ddy = (ey / psy) * (1.0 + (dym - pmy) / gamma) - (eny / nsy) * (1.0 - (dyn - nmy) / gamma);
      a.dpin[p] = make_double2(w * ddx, w * ddy);
    }
    if (lane == 0) value[0] = w * ((xmax - xmin) + pmx - nmx) + w * ((ymax - ymin) + pmy - nmy);
  }
  block_sum<1>(value, partials + blockIdx.x);
}

// One launch for both shapes: the first `wave_blocks` blocks take the
// high-fanout nets (long serial chains, scheduled first), the rest take one
// small net per thread.  Block b writes partial b.
__global__ void __launch_bounds__(BLOCK) k_wa(WaArgs a, const double *__restrict__ coords, double gamma,
                                              double *__restrict__ partials, int wave_blocks) {
  if ((int)blockIdx.x < wave_blocks)
    wa_wave_nets(a, coords, gamma, partials, blockIdx.x);
  else
    wa_thread_nets(a, coords, gamma, partials, blockIdx.x - wave_blocks);
}

// Combined unit gradient (combined_objective, eplace.rs:1022-1044): WA
// baseline pins, then timed-arc pins, then each field in order.
__global__ void __launch_bounds__(BLOCK) k_unit_gradient(WaArgs a, int F, int DU, const double2 *__restrict__ gfu,
                                                         const double *__restrict__ coords,
                                                         const double *__restrict__ metric,
                                                         const double *__restrict__ lower,
                                                         const double *__restrict__ upper, double *__restrict__ grad,
                                                         double *__restrict__ wa_grad_out,
                                                         double *__restrict__ partials, EvalScalars *scalars) {
  const int u = blockIdx.x * BLOCK + threadIdx.x;
  double l1[1] = {0.0};
  double kkt = 0.0;
  bool bad = false;
  if (u < a.n_units) {
    double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
    for (uint32_t k = a.u_pstart[u]; k < a.u_pstart[u + 1]; ++k) {
      const uint32_t p = a.u_pins[k];
      const double2 d = a.dpin[p];
      if (p < a.first_arc_pin) {
        sx += d.x;
        sy += d.y;
      } else {
        tx += d.x;
        ty += d.y;
      }
    }
    double gx = sx + tx, gy = sy + ty;
    l1[0] = fabs(gx) + fabs(gy);
    if (wa_grad_out) {
      wa_grad_out[2 * (size_t)u] = gx;
      wa_grad_out[2 * (size_t)u + 1] = gy;
    }
    const int du = a.unit_du[u];
    if (du >= 0) {
      for (int f = 0; f < F; ++f) {
        const double2 v = gfu[(size_t)f * DU + du];
        gx += v.x;
        gy += v.y;
      }
    }
    const size_t c = 2 * (size_t)u;
    grad[c] = gx;
    grad[c + 1] = gy;
    bad = !isfinite(gx) || !isfinite(gy);
    if (metric) {
      kkt = fmax(kkt_scaled(gx, coords[c], lower[c], upper[c], metric[c]),
                 kkt_scaled(gy, coords[c + 1], lower[c + 1], upper[c + 1], metric[c + 1]));
    }
  }
  block_sum<1>(l1, partials + (size_t)blockIdx.x * UG_STRIDE);
  block_max_flag(kkt, bad, partials + (size_t)blockIdx.x * UG_STRIDE + 1);
}

// Single-block reductions of per-block partials.  Each thread accumulates a
// fixed strided subset, waves reduce by shuffles and four wave results are
// combined in wave order: a fixed order, so results are reproducible.
template <int K>
__device__ void sum_partials(const double *__restrict__ p, int count, double (&out)[K], int stride, int offset) {
  __shared__ double s[K][BLOCK / 64];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < count; b += BLOCK)
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += p[(size_t)b * stride + offset + k];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    acc[k] = wave_sum(acc[k]);
    if (lane == 0) s[k][wave] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = ((s[k][0] + s[k][1]) + s[k][2]) + s[k][3];
  __syncthreads();
}

template <int K>
__device__ void max_partials(const double *__restrict__ p, int count, double (&out)[K], int stride, int offset) {
  __shared__ double s[K][BLOCK / 64];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < count; b += BLOCK)
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = fmax(acc[k], p[(size_t)b * stride + offset + k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    acc[k] = wave_max(acc[k]);
    if (lane == 0) s[k][wave] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = fmax(fmax(s[k][0], s[k][1]), fmax(s[k][2], s[k][3]));
  __syncthreads();
}

__global__ void __launch_bounds__(BLOCK) k_finalize_eval(FinalizeArgs a, EvalScalars *scalars) {
  double v1[1];
  sum_partials<1>(a.wa_partials, a.nb_wa, v1, 1, 0);
  if (threadIdx.x == 0) scalars->wa_value = v1[0];
  sum_partials<1>(a.ug_partials, a.nb_ug, v1, UG_STRIDE, 0);
  if (threadIdx.x == 0) scalars->wa_l1 = v1[0];
  double gu[2 * MAXF], gf[2 * MAXF], mu[2], mf[2];
  sum_partials<2 * MAXF>(a.gu_partials, a.nb_gu, gu, 2 * MAXF, 0);
  sum_partials<2 * MAXF>(a.gf_partials, a.nb_gf, gf, GF_STRIDE, 0);
  max_partials<2>(a.ug_partials, a.nb_ug, mu, UG_STRIDE, 1);
  max_partials<2>(a.gf_partials, a.nb_gf, mf, GF_STRIDE, 2 * MAXF);
  if (threadIdx.x < MAXF) {
    // the energy uses the same serial order as the coefficient computed in
    // the gather kernels, so both agree bit for bit
    const int f = threadIdx.x;
    double e = 0.0;
    for (int b = 0; b < a.nb_energy; ++b) e += a.energy_partials[b * MAXF + f];
    scalars->energy[f] = e;
    scalars->force_l1[f] = gu[f] + gf[f];
    scalars->raw_l1[f] = gu[MAXF + f] + gf[MAXF + f];
  }
  if (threadIdx.x == 0) {
    const double kkt = fmax(mu[0], mf[0]);
    scalars->kkt_bits = static_cast<unsigned long long>(__double_as_longlong(kkt));
    scalars->nonfinite = (mu[1] > 0.0 || mf[1] > 0.0) ? 1u : 0u;
  }
}

// ---------------------------------------------------------------------------
// Preconditioner and Nesterov elementwise steps.
// ---------------------------------------------------------------------------
// diagonal_preconditioner (eplace.rs:1046-1100)
__global__ void __launch_bounds__(BLOCK) k_preconditioner(int n_units, int n_fillers, int F, int DU,
                                                          const double *__restrict__ wl_diag,
                                                          const int32_t *__restrict__ unit_du,
                                                          const double *__restrict__ du_fcharge,
                                                          const uint8_t *__restrict__ fl_field,
                                                          const double *__restrict__ fl_charge, double l0, double l1,
                                                          double l2, double l3, double *__restrict__ metric,
                                                          double *__restrict__ partials) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const double lambda[MAXF] = {l0, l1, l2, l3};
  double d = 0.0;
  bool valid = false;
  if (i < n_units) {
    d = wl_diag[i];
    const int du = unit_du[i];
    if (du >= 0) {
      for (int f = 0; f < F; ++f) {
        const double charge = du_fcharge[(size_t)f * DU + du];
        if (charge != 0.0) d += lambda[f] * charge;
      }
    }
    valid = true;
  } else if (i < n_units + n_fillers) {
    const int j = i - n_units;
    d = 0.0 + lambda[fl_field[j]] * fl_charge[j];
    valid = true;
  }
  double lo = INFINITY, hi = 0.0;
  bool bad = false;
  if (valid) {
    bad = !isfinite(d) || d < 0.0;
    d = fmax(d, 1.0);
    metric[2 * (size_t)i] = d;
    metric[2 * (size_t)i + 1] = d;
    lo = d;
    hi = d;
  }
  // block max / -min / flag, reduced by k_metric_finalize
  __shared__ double smax[BLOCK / 64], smin[BLOCK / 64];
  __shared__ int sbad[BLOCK / 64];
  lo = wave_min(lo);
  hi = wave_max(hi);
  const bool anybad = __any(bad);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smax[wave] = hi;
    smin[wave] = lo;
    sbad[wave] = anybad ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = 0.0, mn = INFINITY;
    int b = 0;
    for (int w = 0; w < BLOCK / 64; ++w) {
      mx = fmax(mx, smax[w]);
      mn = fmin(mn, smin[w]);
      b |= sbad[w];
    }
    partials[3 * (size_t)blockIdx.x] = mx;
    partials[3 * (size_t)blockIdx.x + 1] = mn;
    partials[3 * (size_t)blockIdx.x + 2] = b ? 1.0 : 0.0;
  }
}

__global__ void __launch_bounds__(BLOCK) k_metric_finalize(const double *__restrict__ partials, int nb,
                                                           MetricScalars *ms) {
  __shared__ double smax[BLOCK], smin[BLOCK], sbad[BLOCK];
  double mx = 0.0, mn = INFINITY, b = 0.0;
  for (int i = threadIdx.x; i < nb; i += BLOCK) {
    mx = fmax(mx, partials[3 * (size_t)i]);
    mn = fmin(mn, partials[3 * (size_t)i + 1]);
    b = fmax(b, partials[3 * (size_t)i + 2]);
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  sbad[threadIdx.x] = b;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
      sbad[threadIdx.x] = fmax(sbad[threadIdx.x], sbad[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ms->max_bits = static_cast<unsigned long long>(__double_as_longlong(smax[0]));
    ms->min_bits = static_cast<unsigned long long>(__double_as_longlong(smin[0]));
    ms->nonfinite = sbad[0] > 0.0 ? 1u : 0u;
  }
}

__global__ void __launch_bounds__(BLOCK) k_extrapolate(size_t n, const double *__restrict__ x,
                                                       const double *__restrict__ xp,
                                                       const double *__restrict__ lower,
                                                       const double *__restrict__ upper, double beta,
                                                       double *__restrict__ y, unsigned int *differs) {
  const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  bool changed = false;
  if (i < n) {
    const double cur = x[i];
    double v = cur + beta * (cur - xp[i]);
    const double lo = lower[i], hi = upper[i];
    v = v < lo ? lo : (v > hi ? hi : v);
    y[i] = v;
    changed = v != cur;
  }
  if (__syncthreads_or(changed) && threadIdx.x == 0) *differs = 1u;
}

__global__ void __launch_bounds__(BLOCK) k_candidate(size_t n, const double *__restrict__ y,
                                                     const double *__restrict__ gy,
                                                     const double *__restrict__ metric, double lipschitz,
                                                     const double *__restrict__ lower,
                                                     const double *__restrict__ upper,
                                                     const double *__restrict__ origin, double *__restrict__ cand,
                                                     double *__restrict__ partials) {
  const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  double red[2] = {0.0, 0.0};
  double change = 0.0;
  bool changed = false;
  if (i < n) {
    const double m = metric[i];
    const double yi = y[i], g = gy[i];
    double c = yi - g / (lipschitz * m);
    const double lo = lower[i], hi = upper[i];
    c = c < lo ? lo : (c > hi ? hi : c);
    cand[i] = c;
    changed = c != yi;
    const double d = c - yi;
    red[0] = g * d;
    red[1] = m * d * d;
    change = fabs(c - origin[i]);
  }
  block_sum<2>(red, partials + 4 * (size_t)blockIdx.x);
  block_max_flag(change, changed, partials + 4 * (size_t)blockIdx.x + 2);
}

__global__ void __launch_bounds__(BLOCK) k_step_finalize(const double *__restrict__ partials, int nb,
                                                         StepScalars *step) {
  double v[2], m[2];
  sum_partials<2>(partials, nb, v, 4, 0);
  max_partials<2>(partials, nb, m, 4, 2);
  if (threadIdx.x == 0) {
    step->gdelta = v[0];
    step->msq = v[1];
    step->change_bits = static_cast<unsigned long long>(__double_as_longlong(m[0]));
    step->changed = m[1] > 0.0 ? 1u : 0u;
  }
}

__global__ void k_zero_step(StepScalars *step, unsigned int *flag) {
  step->gdelta = 0.0;
  step->msq = 0.0;
  step->change_bits = 0;
  step->changed = 0;
  if (flag) *flag = 0;
}

// ---------------------------------------------------------------------------
// RUDY demand (eplace.rs:856-937).  One wavefront per net: lanes reduce the
// bounding box, then stride over the candidate tiles.  The reference scans
// every tile and keeps those whose overlap with [min-1/2, max+1/2] is
// positive; that set lies inside t in (min-1, max+1), so the kernel visits
// floor(min-1)..ceil(max+1) and applies the reference's own overlap formula,
// which keeps the tile ranges and overlap values bit-exact.  Demand is
// accumulated in 2^-40 fixed point so the sums are run-to-run identical.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tile_overlap(int t, double rlo, double rhi) {
  const double tmin = static_cast<double>(t) - 0.5, tmax = static_cast<double>(t) + 0.5;
  const double o = fmin(tmax, rhi) - fmax(tmin, rlo);
  return o > 0.0 ? o : 0.0;
}

__global__ void __launch_bounds__(BLOCK) k_rudy(int n_nets, const uint32_t *__restrict__ net_start,
                                                const uint32_t *__restrict__ pin_unit,
                                                const double2 *__restrict__ pin_xy,
                                                const double *__restrict__ coords, int W, int H,
                                                long long *__restrict__ hfx, long long *__restrict__ vfx,
                                                unsigned int *error) {
  const int lane = threadIdx.x & 63;
  const int net = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
  if (net >= n_nets) return;
  const uint32_t b = net_start[net], e = net_start[net + 1];
  if (b == e) return;
  const uint32_t first_unit = pin_unit[b];
  bool same = true;
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  for (uint32_t p = b + lane; p < e; p += 64) {
    const uint32_t u = pin_unit[p];
    same = same && (u == first_unit);
    const double2 v = pin_xy[p];
    double x = v.x, y = v.y;
    if (u != NO_UNIT) {
      x = coords[2 * (size_t)u] + v.x;
      y = coords[2 * (size_t)u + 1] + v.y;
    }
    xmin = fmin(xmin, x);
    xmax = fmax(xmax, x);
    ymin = fmin(ymin, y);
    ymax = fmax(ymax, y);
  }
  xmin = wave_min(xmin);
  xmax = wave_max(xmax);
  ymin = wave_min(ymin);
  ymax = wave_max(ymax);
  if (__all(same)) return;
  if (!isfinite(xmin) || !isfinite(xmax) || !isfinite(ymin) || !isfinite(ymax)) {
    if (lane == 0) *error = 1u;
    return;
  }
  const double dx = xmax - xmin, dy = ymax - ymin;
  const double area = (dx + 1.0) * (dy + 1.0);
  const double hden = dx / area, vden = dy / area;
  const double rx0 = xmin - 0.5, rx1 = xmax + 0.5, ry0 = ymin - 0.5, ry1 = ymax + 0.5;
  int x0 = static_cast<int>(fmax(floor(xmin - 1.0), 0.0));
  int x1 = static_cast<int>(fmin(ceil(xmax + 1.0), static_cast<double>(W - 1)));
  int y0 = static_cast<int>(fmax(floor(ymin - 1.0), 0.0));
  int y1 = static_cast<int>(fmin(ceil(ymax + 1.0), static_cast<double>(H - 1)));
  const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
  if (nx <= 0 || ny <= 0) return;
  const long long tiles = (long long)nx * ny;
  for (long long k = lane; k < tiles; k += 64) {
    const int tx = x0 + static_cast<int>(k % nx);
    const int ty = y0 + static_cast<int>(k / nx);
    const double ox = tile_overlap(tx, rx0, rx1);
    const double oy = tile_overlap(ty, ry0, ry1);
    if (ox > 0.0 && oy > 0.0) {
      const double overlap = ox * oy;
      const size_t idx = (size_t)ty * W + tx;
      atomicAdd(reinterpret_cast<unsigned long long *>(hfx + idx),
                static_cast<unsigned long long>(__double2ll_rn(hden * overlap * FX_SCALE)));
      atomicAdd(reinterpret_cast<unsigned long long *>(vfx + idx),
                static_cast<unsigned long long>(__double2ll_rn(vden * overlap * FX_SCALE)));
    }
  }
}

__global__ void __launch_bounds__(BLOCK) k_fill_ll(size_t n, long long *p) {
  const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void __launch_bounds__(BLOCK) k_fx_to_double(size_t n, const long long *__restrict__ fx,
                                                        double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = static_cast<double>(fx[i]) * FX_INV;
}

// rounded_targets_from_coordinates (eplace.rs:954-971, lib.rs:4566-4568)
__global__ void __launch_bounds__(BLOCK) k_round_targets(int n_units, const double *__restrict__ coords, int W, int H,
                                                         uint32_t *__restrict__ out) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= 2 * n_units) return;
  const double extent = static_cast<double>(((i & 1) ? H : W) - 1);
  double r = round(coords[i]);
  if (!(r >= 0.0)) r = 0.0;
  if (r > extent) r = extent;
  out[i] = static_cast<uint32_t>(r);
}

int blocks_for(size_t n) { return static_cast<int>((n + BLOCK - 1) / BLOCK); }

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
void launch_zero_eval(EvalScalars *scalars, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_eval, dim3(1), dim3(1), 0, s, scalars);
}

void launch_deposit(const DensityArgs &a, const double *coords, EvalScalars *scalars, hipStream_t s) {
  const int total = a.F * a.W * a.H;
  hipLaunchKernelGGL(k_deposit_init, dim3(blocks_for(total)), dim3(BLOCK), 0, s, total, a.base_fx, a.acc_m, a.acc_f,
                     scalars);
  const int n = a.n_members + a.n_fillers;
  if (n > 0) {
    timer_begin(g_timer, "deposit", s);
    hipLaunchKernelGGL(k_deposit, dim3(blocks_for(n)), dim3(BLOCK), 0, s, a, coords);
    timer_end(g_timer, "deposit", s);
  }
}

void launch_density_convert(const DensityArgs &a, float *rho, double *density_out, EvalScalars *scalars,
                            hipStream_t s) {
  const int G = a.W * a.H;
  hipLaunchKernelGGL(k_density_convert, dim3(blocks_for((size_t)a.F * G)), dim3(BLOCK), 0, s, G, a.F,
                     a.acc_m, a.acc_f, rho, density_out, scalars);
}

void launch_gather(const DensityArgs &a, const GatherArgs &g, const double *coords, double2 *gfu, double *raw_units,
                   double *raw_fillers, double *grad, const double *metric, const double *lower, const double *upper,
                   double *partials_units, double *partials_fillers, EvalScalars *scalars, int *nb_units,
                   int *nb_fillers, hipStream_t s) {
  *nb_units = std::max(1, blocks_for(a.n_density_units));
  *nb_fillers = std::max(1, blocks_for(a.n_fillers));
  timer_begin(g_timer, "gather", s);
  hipLaunchKernelGGL(k_gather_units, dim3(*nb_units), dim3(BLOCK), 0, s, a, g, coords, gfu, raw_units,
                     partials_units);
  timer_end(g_timer, "gather", s);
  hipLaunchKernelGGL(k_gather_fillers, dim3(*nb_fillers), dim3(BLOCK), 0, s, a, g, coords, raw_fillers, grad, metric,
                     lower, upper, partials_fillers, scalars);
}

void launch_wa(const WaArgs &a, const double *coords, double gamma, double *partials, int *nb, hipStream_t s) {
  const int nbs = a.n_small > 0 ? blocks_for(a.n_small) : 0;
  const int nbl = a.n_large > 0 ? (a.n_large + BLOCK / 64 - 1) / (BLOCK / 64) : 0;
  *nb = nbs + nbl;
  if (*nb == 0) return;
  timer_begin(g_timer, "wa", s);
  hipLaunchKernelGGL(k_wa, dim3(*nb), dim3(BLOCK), 0, s, a, coords, gamma, partials, nbl);
  timer_end(g_timer, "wa", s);
}

void launch_unit_gradient(const WaArgs &a, int F, int DU, const double2 *gfu, const double *coords,
                          const double *metric, const double *lower, const double *upper, double *grad,
                          double *wa_grad_out, double *partials, EvalScalars *scalars, int *nb, hipStream_t s) {
  *nb = std::max(1, blocks_for(a.n_units));
  hipLaunchKernelGGL(k_unit_gradient, dim3(*nb), dim3(BLOCK), 0, s, a, F, DU, gfu, coords, metric, lower, upper,
                     grad, wa_grad_out, partials, scalars);
}

void launch_finalize_eval(const FinalizeArgs &a, EvalScalars *scalars, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_eval, dim3(1), dim3(BLOCK), 0, s, a, scalars);
}

void launch_preconditioner(int n_units, int n_fillers, int F, int DU, const double *wl_diag, const int32_t *unit_du,
                           const double *du_fcharge, const uint8_t *fl_field, const double *fl_charge,
                           const double *lambda, double *metric, double *partials, MetricScalars *ms, hipStream_t s) {
  const int n = n_units + n_fillers;
  const int nb = std::max(1, blocks_for(n));
  hipLaunchKernelGGL(k_preconditioner, dim3(nb), dim3(BLOCK), 0, s, n_units, n_fillers, F, DU, wl_diag, unit_du,
                     du_fcharge, fl_field, fl_charge, lambda[0], lambda[1], lambda[2], lambda[3], metric, partials);
  hipLaunchKernelGGL(k_metric_finalize, dim3(1), dim3(BLOCK), 0, s, partials, nb, ms);
}

void launch_extrapolate(size_t n, const double *x, const double *xp, const double *lower, const double *upper,
                        double beta, double *y, unsigned int *differs, hipStream_t s) {
  hipLaunchKernelGGL(k_extrapolate, dim3(std::max(1, blocks_for(n))), dim3(BLOCK), 0, s, n, x, xp, lower, upper, beta,
                     y, differs);
}

void launch_candidate(size_t n, const double *y, const double *gy, const double *metric, double lipschitz,
                      const double *lower, const double *upper, const double *origin, double *cand, double *partials,
                      StepScalars *step, hipStream_t s) {
  const int nb = std::max(1, blocks_for(n));
  hipLaunchKernelGGL(k_candidate, dim3(nb), dim3(BLOCK), 0, s, n, y, gy, metric, lipschitz, lower, upper, origin,
                     cand, partials);
  hipLaunchKernelGGL(k_step_finalize, dim3(1), dim3(BLOCK), 0, s, partials, nb, step);
}

void launch_zero_step(StepScalars *step, unsigned int *flag, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_step, dim3(1), dim3(1), 0, s, step, flag);
}

void launch_rudy(int n_nets, const uint32_t *net_start, const uint32_t *pin_unit, const double2 *pin_xy,
                 const double *coords, int W, int H, long long *fx, double *out, unsigned int *error, hipStream_t s) {
  const size_t G = (size_t)W * H;
  hipLaunchKernelGGL(k_fill_ll, dim3(std::max(1, blocks_for(2 * G))), dim3(BLOCK), 0, s, 2 * G, fx);
  const int nb = std::max(1, (n_nets + BLOCK / 64 - 1) / (BLOCK / 64));
  timer_begin(g_timer, "rudy", s);
  hipLaunchKernelGGL(k_rudy, dim3(nb), dim3(BLOCK), 0, s, n_nets, net_start, pin_unit, pin_xy, coords, W, H, fx,
                     fx + G, error);
  timer_end(g_timer, "rudy", s);
  hipLaunchKernelGGL(k_fx_to_double, dim3(std::max(1, blocks_for(2 * G))), dim3(BLOCK), 0, s, 2 * G, fx, out);
}

void launch_round_targets(int n_units, const double *coords, int W, int H, uint32_t *out, hipStream_t s) {
  if (n_units == 0) return;
  hipLaunchKernelGGL(k_round_targets, dim3(blocks_for(2 * (size_t)n_units)), dim3(BLOCK), 0, s, n_units, coords, W, H,
                     out);
}

}  // namespace texo
