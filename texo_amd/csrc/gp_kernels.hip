// gp_kernels.hip — scatter/gather and elementwise kernels of the electrostatic
// global-placement loop for gfx950 (64-lane wavefronts).
//
// Every kernel covers a table of tasks through a block plan; each block works
// for exactly one task.  All are HBM/latency bound.  Reductions are
// deterministic: integer sums use int64 atomics (exact, order-free), FP64 sums
// and maxima go through per-block partials that one block per task reduces
// in a fixed order.  No kernel uses same-address atomics across blocks.

#include "gp_common.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace texo {

namespace {

__device__ __forceinline__ double wave_sum(double v) { return lane_reduce<64>(v, OpAdd{}); }
__device__ __forceinline__ double wave_max(double v) { return lane_reduce<64>(v, OpMaxF{}); }
__device__ __forceinline__ double wave_min(double v) { return lane_reduce<64>(v, OpMinF{}); }

// Block-wide sum of K values per thread; thread 0 writes out[0..K).
template <int K>
__device__ void block_sum(double (&v)[K], double *out, uint32_t active = ~0u) {
  // entries outside `active` (a block-uniform mask) are zero in every lane:
  // their sums are written as 0.0 without reducing
  __shared__ double s[K][BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (active >> k & 1u) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) s[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = (active >> k & 1u) ? ((s[k][0] + s[k][1]) + s[k][2]) + s[k][3] : 0.0;
  }
}

// Mask of the per-field entries [first, first + F) of a partial record.
__device__ __forceinline__ uint32_t field_mask(int F, int first) { return ((1u << F) - 1u) << first; }

// Block-wide maximum of a non-negative value and OR of a flag; thread 0
// writes out[0] = max and out[1] = flag as 0.0 / 1.0.
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
    out[0] = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
    out[1] = (f[0] | f[1] | f[2] | f[3]) ? 1.0 : 0.0;
  }
}

// Single-block reductions of per-block partials: each thread accumulates a
// fixed strided subset, waves reduce by shuffles, four wave results combine
// in wave order.  The order is fixed, so results are reproducible.
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

// kkt_component (global_placement.rs:675-689) divided by sqrt(metric).
__device__ __forceinline__ double kkt_scaled(double g, double x, double lo, double hi, double m) {
  const bool equal = __double_as_longlong(lo) == __double_as_longlong(hi);
  if (equal || (x <= lo && g > 0.0) || (x >= hi && g < 0.0)) return 0.0;
  return fabs(g) / sqrt(m);
}

// ---------------------------------------------------------------------------
// Density deposit (electrostatic_placement.rs:296-332, 900-916).
// ---------------------------------------------------------------------------
// Deposit, one block per sort tile (DEP_TILE x DEP_TILE bins) of the last
// charge sort.  The block walks the tile's charges (contiguous in the sorted arrays) and adds each
// stencil cell's int64 quantum into an LDS window covering the tile plus
// DEP_HALO bins, one layer per (field, members | fillers); it then stores the
// whole window to the tile's partial with plain stores.  A charge that moved
// so far since the sort that its stencil leaves the window adds its quanta to
// the HBM accumulators with atomics instead.  k_convert adds, for every bin,
// the accumulators and the windows that cover it.  The quanta are the same
// as a sequential deposit's and int64 addition is exact, so the density does
// not depend on the order, the tiling or which path a charge took.
__global__ void __launch_bounds__(BLOCK) k_deposit(const Task *__restrict__ tasks, BlockPlan plan) {
  __shared__ unsigned long long win[2 * MAXF * DEP_WIN2];
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const Task &t = tasks[bm.x];
  const JobStatic &a = t.j;
  const int key = bm.y;
  const uint32_t begin = key > 0 ? a.key_end[key - 1] : 0u, end = a.key_end[key];
  if (begin == end) {
    if (threadIdx.x == 0) t.tile_flag[key] = 0u;
    return;
  }
  const int cells = 2 * a.F * DEP_WIN2;
  for (int idx = threadIdx.x; idx < cells; idx += BLOCK) win[idx] = 0ull;
  __syncthreads();
  const int x0 = (key % a.tiles_x) * DEP_TILE - DEP_HALO, y0 = (key / a.tiles_x) * DEP_TILE - DEP_HALO;
  const size_t G = (size_t)a.W * (size_t)a.H;
  unsigned int moved = 0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += BLOCK) {
    // a filler's offset is +0.0, which leaves its coordinate unchanged
    const uint32_t ci = a.s_ci[i];
    const double2 off = a.s_off[i];
    const int layer = a.s_layer[i];
    const double q = a.s_q[i];
    Axis ax, ay;
    bspline_axis(a.W, t.coords[2 * (size_t)ci] + off.x, ax);
    bspline_axis(a.H, t.coords[2 * (size_t)ci + 1] + off.y, ay);
    const int ix[3] = {ax.i0, ax.i1, ax.i2};
    const int iy[3] = {ay.i0, ay.i1, ay.i2};
    const double wx[3] = {ax.w0, ax.w1, ax.w2};
    const double wy[3] = {ay.w0, ay.w1, ay.w2};
    const bool inside = ax.i0 >= x0 && ax.i2 < x0 + DEP_WIN && ay.i0 >= y0 && ay.i2 < y0 + DEP_WIN;
    if (inside) {
      unsigned long long *w = win + (size_t)layer * DEP_WIN2;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          // charge * column.weight * row.weight, quantised before the add
          const double v = q * wx[c] * wy[r];
          atomicAdd(w + (iy[r] - y0) * DEP_WIN + (ix[c] - x0),
                    static_cast<unsigned long long>(__double2ll_rn(v * FX_SCALE)));
        }
      }
    } else {
      moved += 1;
      unsigned long long *g =
          reinterpret_cast<unsigned long long *>(((layer & 1) ? t.acc_f : t.acc_m) + (size_t)(layer >> 1) * G);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double v = q * wx[c] * wy[r];
          atomicAdd(g + (size_t)iy[r] * a.W + ix[c], static_cast<unsigned long long>(__double2ll_rn(v * FX_SCALE)));
        }
      }
    }
  }
  // charges outside the window (moved since the sort): one count per block
  __shared__ unsigned int s_moved[BLOCK / 64];
  moved = lane_reduce<64>(moved, OpAdd{});
  if ((threadIdx.x & 63) == 0) s_moved[threadIdx.x >> 6] = moved;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int n = s_moved[0] + s_moved[1] + s_moved[2] + s_moved[3];
    if (n) atomicAdd(&t.scalars->moved_acc, n);
  }
  long long *out = t.partial + (size_t)key * cells;
  for (int idx = threadIdx.x; idx < cells; idx += BLOCK) out[idx] = static_cast<long long>(win[idx]);
  if (threadIdx.x == 0) t.tile_flag[key] = 1u;
}

// ---------------------------------------------------------------------------
// Counting sort of charges by sort tile (clear, count, scan, scatter).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int charge_key(const JobStatic &a, const double *__restrict__ coords, int i) {
  double x, y;
  if (i < a.n_members) {
    const uint32_t u = a.m_unit[i];
    const double2 off = a.m_off[i];
    x = coords[2 * (size_t)u] + off.x;
    y = coords[2 * (size_t)u + 1] + off.y;
  } else {
    const size_t o = 2 * ((size_t)a.n_units + (i - a.n_members));
    x = coords[o];
    y = coords[o + 1];
  }
  const int cx = bin_center(a.W, x), cy = bin_center(a.H, y);
  return (cy >> SORT_TILE_SHIFT) * a.tiles_x + (cx >> SORT_TILE_SHIFT);
}

__global__ void __launch_bounds__(BLOCK) k_sort_clear(const SortTask *__restrict__ tasks,
                                                      BlockPlan plan) {
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const SortTask &t = tasks[bm.x];
  const int i = bm.y * BLOCK + threadIdx.x;
  if (i <= t.j.n_keys) t.count[i] = 0u;
}

// Charges are walked in the previous deposit order, so the charges of one
// block mostly share a few tiles: the block counts them in an LDS table
// spanning its key range and adds each nonzero count to HBM once (one
// charge at a time when the range is too wide).
constexpr int SORT_LDS_KEYS = 1024;

__device__ __forceinline__ void block_key_range(int key, bool valid, int &kmin, int &kmax) {
  __shared__ int s_lo[BLOCK / 64], s_hi[BLOCK / 64];
  int lo = valid ? key : 0x7fffffff, hi = valid ? key : -1;
  lo = lane_reduce<64>(lo, OpMinI{});
  hi = lane_reduce<64>(hi, OpMaxI{});
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_lo[wave] = lo;
    s_hi[wave] = hi;
  }
  __syncthreads();
  kmin = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
  kmax = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
}

__global__ void __launch_bounds__(BLOCK) k_sort_count(const SortTask *__restrict__ tasks, BlockPlan plan) {
  __shared__ uint32_t cnt[SORT_LDS_KEYS];
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const SortTask &t = tasks[bm.x];
  const int i = bm.y * BLOCK + threadIdx.x;
  const bool valid = i < t.j.n_members + t.j.n_fillers;
  const int key = valid ? charge_key(t.j, t.coords, (int)t.order[i]) : 0;
  int kmin, kmax;
  block_key_range(key, valid, kmin, kmax);
  if (kmax < 0) return;
  if (kmax - kmin >= SORT_LDS_KEYS) {
    if (valid) atomicAdd(&t.count[key], 1u);
    return;
  }
  const int span = kmax - kmin + 1;
  for (int k = threadIdx.x; k < span; k += BLOCK) cnt[k] = 0u;
  __syncthreads();
  if (valid) atomicAdd(&cnt[key - kmin], 1u);
  __syncthreads();
  for (int k = threadIdx.x; k < span; k += BLOCK)
    if (cnt[k]) atomicAdd(&t.count[kmin + k], cnt[k]);
}

// Exclusive scan of the key counts; one block per task.
__global__ void __launch_bounds__(BLOCK) k_sort_scan(const SortTask *__restrict__ tasks,
                                                     const int *__restrict__ list) {
  __shared__ uint32_t part[BLOCK];
  const SortTask &t = tasks[list[blockIdx.x]];
  const int n = t.j.n_keys;
  const int per = (n + BLOCK - 1) / BLOCK;
  const int b = threadIdx.x * per, e = min(n, b + per);
  uint32_t sum = 0;
  for (int i = b; i < e; ++i) sum += t.count[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k < BLOCK; ++k) {
      const uint32_t v = part[k];
      part[k] = run;
      run += v;
    }
  }
  __syncthreads();
  uint32_t run = part[threadIdx.x];
  for (int i = b; i < e; ++i) {
    const uint32_t v = t.count[i];
    t.count[i] = run;
    run += v;
  }
}

// Each block reserves one range per key it holds (one HBM atomic per key),
// then places its charges at LDS-assigned ranks inside the range.
__global__ void __launch_bounds__(BLOCK) k_sort_scatter(const SortTask *__restrict__ tasks, BlockPlan plan) {
  __shared__ uint32_t cnt[SORT_LDS_KEYS], base[SORT_LDS_KEYS];
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const SortTask &t = tasks[bm.x];
  const JobStatic &a = t.j;
  const int i = bm.y * BLOCK + threadIdx.x;
  const bool valid = i < a.n_members + a.n_fillers;
  const int c = valid ? (int)t.order[i] : 0;
  const int key = valid ? charge_key(a, t.coords, c) : 0;
  int kmin, kmax;
  block_key_range(key, valid, kmin, kmax);
  if (kmax < 0) return;
  uint32_t pos = 0;
  if (kmax - kmin >= SORT_LDS_KEYS) {
    if (valid) pos = atomicAdd(&t.count[key], 1u);
  } else {
    const int span = kmax - kmin + 1;
    for (int k = threadIdx.x; k < span; k += BLOCK) cnt[k] = 0u;
    __syncthreads();
    const uint32_t rank = valid ? atomicAdd(&cnt[key - kmin], 1u) : 0u;
    __syncthreads();
    for (int k = threadIdx.x; k < span; k += BLOCK)
      if (cnt[k]) base[k] = atomicAdd(&t.count[kmin + k], cnt[k]);
    __syncthreads();
    pos = valid ? base[key - kmin] + rank : 0u;
  }
  if (!valid) return;
  t.perm[pos] = static_cast<uint32_t>(c);
  if (c < a.n_members) {
    t.s_ci[pos] = a.m_unit[c];
    if (t.s_off32)
      t.s_off32[pos] = a.m_off.f[c];
    else
      t.s_off[pos] = a.m_off.d[c];
    t.s_q[pos] = a.m_charge[c];
    t.s_layer[pos] = static_cast<uint8_t>(2 * a.m_field[c]);
  } else {
    const int k = c - a.n_members;
    t.s_ci[pos] = static_cast<uint32_t>(a.n_units + k);
    if (t.s_off32)
      t.s_off32[pos] = make_float2(0.f, 0.f);
    else
      t.s_off[pos] = make_double2(0.0, 0.0);
    t.s_q[pos] = a.fl_charge[k];
    t.s_layer[pos] = static_cast<uint8_t>(2 * a.fl_field[k] + 1);
  }
}

// rho = members + fillers -> FP32 transform input; physical overflow (before
// fillers) and net charge as exact int64 sums.  Each bin adds the background
// (-smoothed capacity, members only), its HBM accumulators and the deposit
// windows of the (at most 2 x 2) tiles whose window covers it, and clears
// the accumulators it read, which are zero again for the next deposit.
__device__ __forceinline__ void tile_range(int v, int tiles, int &lo, int &hi) {
  // tiles t with t*DEP_TILE - DEP_HALO <= v < t*DEP_TILE - DEP_HALO + DEP_WIN
  const int span = DEP_WIN - DEP_HALO - 1;  // v - span <= t*DEP_TILE <= v + DEP_HALO
  lo = v > span ? (v - span + DEP_TILE - 1) / DEP_TILE : 0;
  hi = min((v + DEP_HALO) / DEP_TILE, tiles - 1);
}

__device__ __forceinline__ long long wave_sum_ll(long long v) { return lane_reduce<64>(v, OpAdd{}); }

__global__ void __launch_bounds__(BLOCK) k_convert(const Task *__restrict__ tasks, BlockPlan plan) {
  __shared__ long long s_over[BLOCK / 64][MAXF], s_net[BLOCK / 64][MAXF];
  __shared__ unsigned int s_max[BLOCK / 64][MAXF];
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const Task &t = tasks[bm.x];
  const JobStatic &a = t.j;
  const int G = a.W * a.H;
  const int i = bm.y * BLOCK + threadIdx.x;
  int f = -1;
  long long over = 0, net = 0;
  unsigned int amax = 0u;  // |rho| as float bits: the FP16-scaled transform's input range
  if (i < a.F * G) {
    f = i / G;
    const int cell = i - f * G, y = cell / a.W, x = cell - y * a.W;
    const long long am = t.acc_m[i], af = t.acc_f[i];
    if (am != 0) t.acc_m[i] = 0;
    if (af != 0) t.acc_f[i] = 0;
    long long m = a.base_fx[i] + am, fl = af;
    int tx0, tx1, ty0, ty1;
    tile_range(x, a.tiles_x, tx0, tx1);
    tile_range(y, a.n_keys / a.tiles_x, ty0, ty1);
    const int cells = 2 * a.F * DEP_WIN2;
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tx = tx0; tx <= tx1; ++tx) {
        const int key = ty * a.tiles_x + tx;
        if (!t.tile_flag[key]) continue;
        const long long *w = t.partial + (size_t)key * cells;
        const int o = (y - (ty * DEP_TILE - DEP_HALO)) * DEP_WIN + (x - (tx * DEP_TILE - DEP_HALO));
        m += w[(2 * f) * DEP_WIN2 + o];
        fl += w[(2 * f + 1) * DEP_WIN2 + o];
      }
    const long long v = m + fl;
    const double d = static_cast<double>(v) * FX_INV;
    const float r = static_cast<float>(d);
    t.rho[i] = r;
    if (t.density_out) t.density_out[i] = d;
    over = m > 0 ? m : 0;
    net = v;
    amax = __float_as_uint(fabsf(r));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < MAXF; ++k) {
    const long long o = wave_sum_ll(f == k ? over : 0);
    const long long n = wave_sum_ll(f == k ? net : 0);
    unsigned int mx = f == k ? amax : 0u;
    mx = lane_reduce<64>(mx, OpMaxI{});
    if (lane == 0) {
      s_over[wave][k] = o;
      s_net[wave][k] = n;
      s_max[wave][k] = mx;
    }
  }
  __syncthreads();
  // plain per-block partials (k_finalize sums them; the transform's first
  // stage takes the maximum of its field's blocks)
  if (threadIdx.x < MAXF) {
    const int k = threadIdx.x;
    long long *pc = t.part_conv + (size_t)bm.y * 2 * MAXF;
    pc[k] = s_over[0][k] + s_over[1][k] + s_over[2][k] + s_over[3][k];
    pc[MAXF + k] = s_net[0][k] + s_net[1][k] + s_net[2][k] + s_net[3][k];
    t.amax_part[(size_t)bm.y * MAXF + k] = max(max(s_max[0][k], s_max[1][k]), max(s_max[2][k], s_max[3][k]));
  }
}

// ---------------------------------------------------------------------------
// Field gradient gather (electrostatic_placement.rs:338-374, 919-941).
// ---------------------------------------------------------------------------
// One block per task: E_f = 1/2 <rho_bar, phi> from the transform's
// partials (fixed order) and the coefficient lambda (1 + beta E/E0)
// (eplace.rs:1111-1124).
// The WA value of an evaluation from its block partials, summed as
// finalize_task sums them (each thread its partials in order, then the
// wavefronts, then the four waves), so both give the same bits.  Block-wide.
__device__ double wa_value_of(const Task &t) {
  __shared__ double s[2][BLOCK / 64];
  double live = 0.0, fixed = 0.0;
  for (int b = threadIdx.x; b < t.nb_wa; b += BLOCK) {
    if (b < t.nb_wa_live)
      live += t.part_wa[b];
    else
      fixed += t.part_wa[b];
  }
  live = wave_sum(live);
  fixed = wave_sum(fixed);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s[0][wave] = live;
    s[1][wave] = fixed;
  }
  __syncthreads();
  const double v = ((s[0][0] + s[0][1]) + s[0][2]) + s[0][3];
  double fw = 0.0;
  if (t.fixed_wa_mode == 1)
    fw = ((s[1][0] + s[1][1]) + s[1][2]) + s[1][3];
  else if (t.fixed_wa_mode == 2)
    fw = *t.fixed_wa;
  __syncthreads();
  return v + fw;
}

// The acceptance test of a line-search trial (global_placement.rs:274-290)
// with a margin four times the reference's, so that a candidate the host
// accepts always passes it; and whether every gradient entry is bounded
// (then the reference's gradient check passes).  Sets gradient_skipped.
__device__ void trial_test(const Task &t, const double (&e)[MAXF]) {
  const double wa = wa_value_of(t);
  // max |phi| per field from the last transform stage's tile partials
  __shared__ unsigned int sm[MAXF][BLOCK / 64];
  unsigned int m[MAXF];
#pragma unroll
  for (int f = 0; f < MAXF; ++f) {
    m[f] = 0u;
    if (t.phi_amax && f < t.j.F)
      for (int i = threadIdx.x; i < t.phi_tiles; i += BLOCK) m[f] = max(m[f], t.phi_amax[(size_t)f * t.phi_tiles + i]);
    m[f] = lane_reduce<64>(m[f], OpMaxI{});
    if ((threadIdx.x & 63) == 0) sm[f][threadIdx.x >> 6] = m[f];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  // combined_objective (eplace.rs:1022-1044), as the host computes it
  double value = wa;
  double bound = t.trial_wa_bound;
  bool bounded = t.phi_amax != nullptr;
  for (int f = 0; f < t.j.F; ++f) {
    const double ef = e[f] / t.scale[f];
    value += t.lambda[f] * t.scale[f] * (ef + 0.5 * AUGMENTED_DENSITY_BETA * ef * ef);
    const unsigned int mf = max(max(sm[f][0], sm[f][1]), max(sm[f][2], sm[f][3]));
    bounded = bounded && mf < 0x7f800000u;
    const double coef = t.lambda[f] * (1.0 + AUGMENTED_DENSITY_BETA * (e[f] / t.scale[f]));
    // a raw entry is the sum of its charges times a field sample, and a
    // sample is at most 2 max |phi| (the stencil's derivative weights sum to
    // at most 2 in magnitude, its value weights to 1)
    bound += fabs(coef) * 4.0 * t.trial_charge[f] * static_cast<double>(__uint_as_float(mf));
  }
  bounded = bounded && isfinite(bound) && bound <= 1e300;
  const StepScalars &st = *t.trial_step;
  const double upper = t.trial_y_value + st.gdelta + 0.5 * t.trial_L * st.msq;
  const double sm_ = 4.0 * DBL_EPSILON * (1.0 + fabs(value) + fabs(upper));
  const double mo_ = 4.0 * DBL_EPSILON * (1.0 + fabs(value) + fabs(t.trial_origin_value));
  // NaN compares false: a non-finite value or bound is not "maybe accepted"
  const bool maybe = !(value > upper + sm_) && !(value > t.trial_origin_value + mo_);
  // candidate == y: the reference does not evaluate it at all
  const bool skip = st.changed == 0u || (!maybe && bounded && isfinite(value) && isfinite(upper));
  t.scalars->gradient_skipped = skip ? 1u : 0u;
}

__device__ void task_coefficients(const Task &t) {
  double e[MAXF];
  sum_partials<MAXF>(t.part_energy, t.nb_energy, e, MAXF, 0);
  if (threadIdx.x < MAXF) {
    const int f = threadIdx.x;
    const double ef = f < t.j.F ? e[f] : 0.0;
    t.field_coef[f] = t.lambda[f] * (1.0 + AUGMENTED_DENSITY_BETA * (ef / t.scale[f]));
    t.field_coef[MAXF + f] = ef;
  }
  if (t.trial)
    trial_test(t, e);
  else if (threadIdx.x == 0)
    t.scalars->gradient_skipped = 0u;
}

// d/dx, d/dy of the quadratic B-spline interpolant of phi at (x, y)
// (sample_quadratic_bspline_with_gradient, electrostatic_placement.rs:919-941).
// `at(ix, iy)` returns phi at a bin.
template <class At>
__device__ __forceinline__ void sample_field(const Axis &ax, const Axis &ay, At at, double &gx, double &gy) {
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
      const double p = static_cast<double>(at(ix[c], iy[r]));
      gx += dx[c] * wy[r] * p;
      gy += wx[c] * dy[r] * p;
    }
  }
}

// One block per task: the field coefficients of the evaluation.
__global__ void __launch_bounds__(BLOCK) k_coefficients(const Task *__restrict__ tasks) {
  task_coefficients(tasks[blockIdx.x]);
}

// Thread per filler, input order: combined filler gradient, force and raw
// l1 sums, KKT term.
__device__ void filler_gradient_block(const Task &t, int lb) {
  const JobStatic &a = t.j;
  const int k = lb * BLOCK + threadIdx.x;
  double red[2 * MAXF];
#pragma unroll
  for (int q = 0; q < 2 * MAXF; ++q) red[q] = 0.0;
  double kkt = 0.0;
  bool bad = false;
  if (k < a.n_fillers) {
    const size_t c = 2 * ((size_t)a.n_units + k);
    const int f = a.fl_field[k];
    const double q = a.fl_charge[k];
    // the field at the filler (a filler's offset is +0.0)
    Axis ax, ay;
    bspline_axis(a.W, t.coords[c] + 0.0, ax);
    bspline_axis(a.H, t.coords[c + 1] + 0.0, ay);
    const float *phi = t.phi + (size_t)f * ((size_t)a.W * a.H);
    double gx, gy;
    sample_field(ax, ay, [&](int x, int y) { return phi[(size_t)y * a.W + x]; }, gx, gy);
    const double rx = q * gx, ry = q * gy;
#pragma unroll
    for (int p = 0; p < MAXF; ++p) {
      if (p == f) {
        red[p] = q * (fabs(gx) + fabs(gy));
        red[MAXF + p] = fabs(rx) + fabs(ry);
      }
    }
    if (t.raw_fillers) {
      t.raw_fillers[2 * (size_t)k] = rx;
      t.raw_fillers[2 * (size_t)k + 1] = ry;
    }
    // the combined gradient of a filler starts at 0.0 (eplace.rs:1031-1041)
    const double coef = t.field_coef[f];
    const double vx = 0.0 + coef * rx, vy = 0.0 + coef * ry;
    t.grad[c] = vx;
    t.grad[c + 1] = vy;
    bad = !isfinite(vx) || !isfinite(vy);
    if (t.metric) {
      kkt = fmax(kkt_scaled(vx, t.coords[c], a.lower[c], a.upper[c], t.metric[c]),
                 kkt_scaled(vy, t.coords[c + 1], a.lower[c + 1], a.upper[c + 1], t.metric[c + 1]));
    }
  }
  block_sum<2 * MAXF>(red, t.part_gf + (size_t)lb * GF_STRIDE, field_mask(a.F, 0) | field_mask(a.F, MAXF));
  block_max_flag(kkt, bad, t.part_gf + (size_t)lb * GF_STRIDE + 2 * MAXF);
}

// ---------------------------------------------------------------------------
// Weighted-average wirelength (analytical_placement.rs:576-703), FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 pin_coordinate(const JobStatic &a, const double *__restrict__ coords, uint32_t p) {
  const uint32_t u = a.pin_unit[p];
  const double2 v = a.pin_xy[p];
  if (u == NO_UNIT) return v;
  return make_double2(coords[2 * (size_t)u] + v.x, coords[2 * (size_t)u + 1] + v.y);
}

struct WaSums {
  double psx, pdx, nsx, ndx, psy, pdy, nsy, ndy;
};

// The reference divides by gamma; these multiply by its reciprocal `ig`
// (a relative difference of one rounding in each exponent's argument).
__device__ __forceinline__ void wa_accumulate(double2 c, double xmin, double xmax, double ymin, double ymax,
                                              double ig, WaSums &s) {
  const double dxm = c.x - xmax, dxn = c.x - xmin, dym = c.y - ymax, dyn = c.y - ymin;
  // exactly 1.0 where the delta is zero (analytical_placement.rs:666-678)
  const double ex = dxm == 0.0 ? 1.0 : exp(dxm * ig);
  const double enx = dxn == 0.0 ? 1.0 : exp(-dxn * ig);
  const double ey = dym == 0.0 ? 1.0 : exp(dym * ig);
  const double eny = dyn == 0.0 ? 1.0 : exp(-dyn * ig);
  s.psx += ex;
  s.pdx += ex * dxm;
  s.nsx += enx;
  s.ndx += enx * dxn;
  s.psy += ey;
  s.pdy += ey * dym;
  s.nsy += eny;
  s.ndy += eny * dyn;
}

// r = reciprocals of the four exponential sums (psx, nsx, psy, nsy)
__device__ __forceinline__ double2 wa_derivative(double2 c, double xmin, double xmax, double ymin, double ymax,
                                                 double ig, const double4 &r, double pmx, double nmx, double pmy,
                                                 double nmy) {
  const double dxm = c.x - xmax, dxn = c.x - xmin, dym = c.y - ymax, dyn = c.y - ymin;
  const double ex = dxm == 0.0 ? 1.0 : exp(dxm * ig);
  const double enx = dxn == 0.0 ? 1.0 : exp(-dxn * ig);
  const double ey = dym == 0.0 ? 1.0 : exp(dym * ig);
  const double eny = dyn == 0.0 ? 1.0 : exp(-dyn * ig);
  const double ddx = (ex * r.x) * (1.0 + (dxm - pmx) * ig) - (enx * r.y) * (1.0 - (dxn - nmx) * ig);
  const double ddy = (ey * r.z) * (1.0 + (dym - pmy) * ig) - (eny * r.w) * (1.0 - (dyn - nmy) * ig);
  return make_double2(ddx, ddy);
}

// Reductions inside aligned groups of G lanes (every lane of the group gets
// the same bits).
template <int G>
__device__ __forceinline__ double group_min(double v) {
  return lane_reduce<G>(v, OpMinF{});
}
template <int G>
__device__ __forceinline__ double group_max(double v) {
  return lane_reduce<G>(v, OpMaxF{});
}
template <int G>
__device__ __forceinline__ double group_sum(double v) {
  return lane_reduce<G>(v, OpAdd{});
}

// One net per aligned group of G lanes (a WA segment of G-lane nets).  Lane
// r of the group holds pins b + r + G k (k < WA_PINS_PER_LANE) in
// registers: every pin coordinate is loaded once and its four exponentials
// are computed once; the group combines extremes and sums by lane reductions.
template <int G>
__device__ double wa_group_net(const JobStatic &a, const Task &t, const WaSeg &seg, int wave) {
  constexpr int K = WA_PINS_PER_LANE;
  const int lane = threadIdx.x & 63, r = lane & (G - 1);
  const int local = (wave - seg.first_wave) * (64 / G) + lane / G;
  const bool valid = local < seg.nets;
  uint32_t b = 0, e = 0;
  double w = 0.0;
  if (valid) {
    const uint32_t net = (uint32_t)(seg.first_net + local);
    if (seg.degree > 0) {  // one degree: the pins follow from the segment
      b = (uint32_t)seg.first_pin + (uint32_t)local * (uint32_t)seg.degree;
      e = b + (uint32_t)seg.degree;
    } else {
      b = a.net_start[net];
      e = a.net_start[net + 1];
    }
    w = a.net_weight[net];
  }
  const double ig = 1.0 / t.gamma;
  double2 c[K];
  bool in[K];
  uint32_t slot[K];
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t p = b + (uint32_t)(r + G * k);
    in[k] = p < e;
    c[k] = make_double2(0.0, 0.0);
    slot[k] = NO_UNIT;
    if (in[k]) {
      c[k] = pin_coordinate(a, t.coords, p);
      if (!t.value_only) slot[k] = a.pin_slot[p];
      xmin = fmin(xmin, c[k].x);
      xmax = fmax(xmax, c[k].x);
      ymin = fmin(ymin, c[k].y);
      ymax = fmax(ymax, c[k].y);
    }
  }
  xmin = group_min<G>(xmin);
  xmax = group_max<G>(xmax);
  ymin = group_min<G>(ymin);
  ymax = group_max<G>(ymax);
  double ex[K], enx[K], ey[K], eny[K];
  // A two-pin net has one exponential per axis: its lower pin's exp(dxm/g)
  // and its upper pin's exp(-dxn/g) take the same argument, bit for bit
  // ((xmin - xmax) * ig, the subtraction and the negation being exact), and
  // the other two are 1.0.
  const bool pair = G == 1 && seg.degree == 2;
  double pair_x = 1.0, pair_y = 1.0;
  if (pair) {
    if (xmax != xmin) pair_x = exp((xmin - xmax) * ig);
    if (ymax != ymin) pair_y = exp((ymin - ymax) * ig);
  }
  WaSums s{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < K; ++k) {
    ex[k] = enx[k] = ey[k] = eny[k] = 0.0;
    if (in[k]) {
      const double dxm = c[k].x - xmax, dxn = c[k].x - xmin, dym = c[k].y - ymax, dyn = c[k].y - ymin;
      // exactly 1.0 where the delta is zero (analytical_placement.rs:666-678)
      if (pair) {
        ex[k] = dxm == 0.0 ? 1.0 : pair_x;
        enx[k] = dxn == 0.0 ? 1.0 : pair_x;
        ey[k] = dym == 0.0 ? 1.0 : pair_y;
        eny[k] = dyn == 0.0 ? 1.0 : pair_y;
      } else {
        ex[k] = dxm == 0.0 ? 1.0 : exp(dxm * ig);
        enx[k] = dxn == 0.0 ? 1.0 : exp(-dxn * ig);
        ey[k] = dym == 0.0 ? 1.0 : exp(dym * ig);
        eny[k] = dyn == 0.0 ? 1.0 : exp(-dyn * ig);
      }
      s.psx += ex[k];
      s.pdx += ex[k] * dxm;
      s.nsx += enx[k];
      s.ndx += enx[k] * dxn;
      s.psy += ey[k];
      s.pdy += ey[k] * dym;
      s.nsy += eny[k];
      s.ndy += eny[k] * dyn;
    }
  }
  s.psx = group_sum<G>(s.psx);
  s.pdx = group_sum<G>(s.pdx);
  s.nsx = group_sum<G>(s.nsx);
  s.ndx = group_sum<G>(s.ndx);
  s.psy = group_sum<G>(s.psy);
  s.pdy = group_sum<G>(s.pdy);
  s.nsy = group_sum<G>(s.nsy);
  s.ndy = group_sum<G>(s.ndy);
  const double4 rs = make_double4(1.0 / s.psx, 1.0 / s.nsx, 1.0 / s.psy, 1.0 / s.nsy);
  const double pmx = s.pdx * rs.x, nmx = s.ndx * rs.y, pmy = s.pdy * rs.z, nmy = s.ndy * rs.w;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!in[k] || t.value_only) continue;
    const double dxm = c[k].x - xmax, dxn = c[k].x - xmin, dym = c[k].y - ymax, dyn = c[k].y - ymin;
    const double ddx = (ex[k] * rs.x) * (1.0 + (dxm - pmx) * ig) - (enx[k] * rs.y) * (1.0 - (dxn - nmx) * ig);
    const double ddy = (ey[k] * rs.z) * (1.0 + (dym - pmy) * ig) - (eny[k] * rs.w) * (1.0 - (dyn - nmy) * ig);
    if (slot[k] != NO_UNIT) t.dpin[slot[k]] = make_double2(w * ddx, w * ddy);
  }
  return valid && r == 0 ? w * ((xmax - xmin) + pmx - nmx) + w * ((ymax - ymin) + pmy - nmy) : 0.0;
}

// One wavefront per net of more than WA_LOOP_PINS pins, looping over its
// pins in each pass; lane 0 returns the value.
__device__ double wa_loop_net(const JobStatic &a, const Task &t, uint32_t net) {
  const int lane = threadIdx.x & 63;
  const uint32_t b = a.net_start[net], e = a.net_start[net + 1];
  const double w = a.net_weight[net];
  const double ig = 1.0 / t.gamma;
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  for (uint32_t p = b + lane; p < e; p += 64) {
    const double2 c = pin_coordinate(a, t.coords, p);
    xmin = fmin(xmin, c.x);
    xmax = fmax(xmax, c.x);
    ymin = fmin(ymin, c.y);
    ymax = fmax(ymax, c.y);
  }
  xmin = wave_min(xmin);
  xmax = wave_max(xmax);
  ymin = wave_min(ymin);
  ymax = wave_max(ymax);
  WaSums s{0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t p = b + lane; p < e; p += 64)
    wa_accumulate(pin_coordinate(a, t.coords, p), xmin, xmax, ymin, ymax, ig, s);
  s.psx = wave_sum(s.psx);
  s.pdx = wave_sum(s.pdx);
  s.nsx = wave_sum(s.nsx);
  s.ndx = wave_sum(s.ndx);
  s.psy = wave_sum(s.psy);
  s.pdy = wave_sum(s.pdy);
  s.nsy = wave_sum(s.nsy);
  s.ndy = wave_sum(s.ndy);
  const double4 r = make_double4(1.0 / s.psx, 1.0 / s.nsx, 1.0 / s.psy, 1.0 / s.nsy);
  const double pmx = s.pdx * r.x, nmx = s.ndx * r.y, pmy = s.pdy * r.z, nmy = s.ndy * r.w;
  for (uint32_t p = b + lane; p < e && !t.value_only; p += 64) {
    const double2 d = wa_derivative(pin_coordinate(a, t.coords, p), xmin, xmax, ymin, ymax, ig, r, pmx, nmx, pmy, nmy);
    const uint32_t slot = a.pin_slot[p];
    if (slot != NO_UNIT) t.dpin[slot] = make_double2(w * d.x, w * d.y);
  }
  return lane == 0 ? w * ((xmax - xmin) + pmx - nmx) + w * ((ymax - ymin) + pmy - nmy) : 0.0;
}

// Per task: each wavefront takes part of one WA segment (looping nets first).
__global__ void __launch_bounds__(BLOCK) k_wa(const Task *__restrict__ tasks, BlockPlan plan) {
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const Task &t = tasks[bm.x];
  const JobStatic &a = t.j;
  const int wave = bm.y * (BLOCK / 64) + (threadIdx.x >> 6);
  // the last segment starting at or before this wavefront (segments are in
  // wavefront order); past its end the wavefront has no net
  int lo = 0, hi = a.n_wa_seg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.wa_seg[mid].first_wave <= wave)
      lo = mid;
    else
      hi = mid - 1;
  }
  WaSeg seg{0, 0, -1, 0, 0, 0, 0, 0};
  if (a.n_wa_seg > 0) {
    seg = a.wa_seg[lo];
    const int n = seg.lanes == 0 ? seg.nets : (seg.nets * seg.lanes + 63) / 64;
    if (wave < seg.first_wave || wave >= seg.first_wave + n) seg.lanes = -1;
  }
  double value[1] = {0.0};
  switch (seg.lanes) {
    case 0:
      value[0] = wa_loop_net(a, t, (uint32_t)(seg.first_net + (wave - seg.first_wave)));
      break;
    case 1:
      value[0] = wa_group_net<1>(a, t, seg, wave);
      break;
    case 2:
      value[0] = wa_group_net<2>(a, t, seg, wave);
      break;
    case 4:
      value[0] = wa_group_net<4>(a, t, seg, wave);
      break;
    case 8:
      value[0] = wa_group_net<8>(a, t, seg, wave);
      break;
    case 16:
      value[0] = wa_group_net<16>(a, t, seg, wave);
      break;
    case 32:
      value[0] = wa_group_net<32>(a, t, seg, wave);
      break;
    case 64:
      value[0] = wa_group_net<64>(a, t, seg, wave);
      break;
    default:
      break;
  }
  block_sum<1>(value, t.part_wa + bm.y);
}

// Combined unit gradient (combined_objective, eplace.rs:1022-1044): WA
// baseline pins, then timed-arc pins, then each field in order, the field
// term being coefficient x the sum of the unit's member samples x charge in
// member order (electrostatic_placement.rs:338-374).
// A wave's 64 units own one contiguous run of the unit-ordered derivative
// array; the wave copies it into LDS with coalesced loads (when it fits) and
// each lane then sums its unit's entries from there in the same order.
constexpr int DPIN_STAGE = 384;

__device__ void unit_gradient_block(const Task &t, int lb) {
  const JobStatic &a = t.j;
  const int u = lb * BLOCK + threadIdx.x;
  __shared__ double2 s_dpin[BLOCK / 64][DPIN_STAGE];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int u0 = lb * BLOCK + wave * 64;
  uint32_t w_begin = 0, w_end = 0;
  if (u0 < a.n_units) {
    w_begin = a.u_pstart[u0];
    w_end = a.u_pstart[min(u0 + 64, a.n_units)];
  }
  const bool staged = w_end - w_begin <= (uint32_t)DPIN_STAGE;
  if (staged)
    for (uint32_t k = w_begin + lane; k < w_end; k += 64) s_dpin[wave][k - w_begin] = t.dpin[k];
  __syncthreads();
  double red[1 + 2 * MAXF];
#pragma unroll
  for (int k = 0; k < 1 + 2 * MAXF; ++k) red[k] = 0.0;
  double kkt = 0.0;
  bool bad = false;
  if (u < a.n_units) {
    double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
    const uint32_t pb = a.u_pstart[u], pe = a.u_pstart[u + 1], pm = pb + a.u_nbase[u];
    auto pin_sums = [&](auto at) {
      for (uint32_t k = pb; k < pm; ++k) {  // net pins
        const double2 d = at(k);
        sx += d.x;
        sy += d.y;
      }
      for (uint32_t k = pm; k < pe; ++k) {  // timed-arc pins
        const double2 d = at(k);
        tx += d.x;
        ty += d.y;
      }
    };
    if (staged)
      pin_sums([&](uint32_t k) { return s_dpin[wave][k - w_begin]; });
    else
      pin_sums([&](uint32_t k) { return t.dpin[k]; });
    double gx = sx + tx, gy = sy + ty;
    red[0] = fabs(gx) + fabs(gy);
    if (t.wa_grad_out) {
      t.wa_grad_out[2 * (size_t)u] = gx;
      t.wa_grad_out[2 * (size_t)u + 1] = gy;
    }
    const uint32_t mb = a.u_mstart[u], me = a.u_mstart[u + 1];
    if (mb < me) {
      double gxf[MAXF], gyf[MAXF];
#pragma unroll
      for (int f = 0; f < MAXF; ++f) gxf[f] = gyf[f] = 0.0;
      const double ux = t.coords[2 * (size_t)u], uy = t.coords[2 * (size_t)u + 1];
      const size_t G = (size_t)a.W * a.H;
      for (uint32_t m = mb; m < me; ++m) {
        const int f = a.m_field[m];
        const double q = a.m_charge[m];
        // the field at the member (sample_quadratic_bspline_with_gradient,
        // electrostatic_placement.rs:919-941)
        const double2 off = a.m_off[m];
        Axis ax, ay;
        bspline_axis(a.W, ux + off.x, ax);
        bspline_axis(a.H, uy + off.y, ay);
        const float *phi = t.phi + (size_t)f * G;
        double2 g;
        sample_field(ax, ay, [&](int x, int y) { return phi[(size_t)y * a.W + x]; }, g.x, g.y);
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
          if (k == f) {
            red[1 + k] += q * (fabs(g.x) + fabs(g.y));
            gxf[k] += q * g.x;
            gyf[k] += q * g.y;
          }
        }
      }
#pragma unroll
      for (int f = 0; f < MAXF; ++f) {
        if (f < a.F) {
          red[1 + MAXF + f] = fabs(gxf[f]) + fabs(gyf[f]);
          const double coef = t.field_coef[f];
          gx += coef * gxf[f];
          gy += coef * gyf[f];
          if (t.raw_units) {
            const int du = a.unit_du[u];
            t.raw_units[2 * ((size_t)f * a.n_density_units + du)] = gxf[f];
            t.raw_units[2 * ((size_t)f * a.n_density_units + du) + 1] = gyf[f];
          }
        }
      }
    }
    const size_t c = 2 * (size_t)u;
    t.grad[c] = gx;
    t.grad[c + 1] = gy;
    bad = !isfinite(gx) || !isfinite(gy);
    if (t.metric) {
      kkt = fmax(kkt_scaled(gx, t.coords[c], a.lower[c], a.upper[c], t.metric[c]),
                 kkt_scaled(gy, t.coords[c + 1], a.lower[c + 1], a.upper[c + 1], t.metric[c + 1]));
    }
  }
  block_sum<1 + 2 * MAXF>(red, t.part_ug + (size_t)lb * UG_STRIDE, 1u | field_mask(a.F, 1) | field_mask(a.F, 1 + MAXF));
  block_max_flag(kkt, bad, t.part_ug + (size_t)lb * UG_STRIDE + 1 + 2 * MAXF);
}

// Every FP64 scalar of one evaluation from its partials (one block).
__device__ void finalize_task(const Task &t) {
  // one pass over the three partial arrays; each sum and maximum combines
  // in the order of sum_partials / max_partials
  constexpr int NS = 1 + (1 + 2 * MAXF) + 2 * MAXF, NM = 4;
  __shared__ double sh[NS + NM][BLOCK / 64];
  double acc[NS + NM];
#pragma unroll
  for (int k = 0; k < NS + NM; ++k) acc[k] = 0.0;
  // fixed-point density sums of the k_convert blocks (exact in any order)
  long long cv[2 * MAXF];
#pragma unroll
  for (int k = 0; k < 2 * MAXF; ++k) cv[k] = 0;
  // one pass over the four partial arrays, so that their loads are in
  // flight together
  // WA: the nets with a movable pin, then (when evaluated) those of fixed pins
  double fixed_wa = 0.0;
  // the gradient blocks of a skipped trial wrote no partials
  const bool skipped = t.trial && t.scalars->gradient_skipped;
  const int nb_ug = skipped ? 0 : t.nb_ug, nb_gf = skipped ? 0 : t.nb_gf;
  const int n = max(max(t.nb_wa, t.nb_conv), max(nb_ug, nb_gf));
  for (int b = threadIdx.x; b < n; b += BLOCK) {
    if (b < t.nb_conv) {
      const long long *pc = t.part_conv + (size_t)b * 2 * MAXF;
#pragma unroll
      for (int k = 0; k < 2 * MAXF; ++k) cv[k] += pc[k];
    }
    if (b < t.nb_wa_live)
      acc[0] += t.part_wa[b];
    else if (b < t.nb_wa)
      fixed_wa += t.part_wa[b];
    if (b < nb_ug) {
      const double *p = t.part_ug + (size_t)b * UG_STRIDE;
#pragma unroll
      for (int k = 0; k < 1 + 2 * MAXF; ++k) acc[1 + k] += p[k];
      acc[NS] = fmax(acc[NS], p[1 + 2 * MAXF]);
      acc[NS + 1] = fmax(acc[NS + 1], p[2 + 2 * MAXF]);
    }
    if (b < nb_gf) {
      const double *p = t.part_gf + (size_t)b * GF_STRIDE;
#pragma unroll
      for (int k = 0; k < 2 * MAXF; ++k) acc[2 + 2 * MAXF + k] += p[k];
      acc[NS + 2] = fmax(acc[NS + 2], p[2 * MAXF]);
      acc[NS + 3] = fmax(acc[NS + 3], p[2 * MAXF + 1]);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // sums of fields past F are zero in every lane: not reduced
  const int F = t.j.F;
  const uint32_t live = 3u | field_mask(F, 2) | field_mask(F, 2 + MAXF) | field_mask(F, 2 + 2 * MAXF) |
                        field_mask(F, 2 + 3 * MAXF) | (((1u << NM) - 1u) << NS);
#pragma unroll
  for (int k = 0; k < NS + NM; ++k) {
    if (live >> k & 1u) acc[k] = k < NS ? wave_sum(acc[k]) : wave_max(acc[k]);
    if (lane == 0) sh[k][wave] = acc[k];
  }
  __shared__ double sfw[BLOCK / 64];
  fixed_wa = wave_sum(fixed_wa);
  if (lane == 0) sfw[wave] = fixed_wa;
  __syncthreads();
  double r[NS + NM];
#pragma unroll
  for (int k = 0; k < NS + NM; ++k)
    r[k] = k < NS ? ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3]
                  : fmax(fmax(sh[k][0], sh[k][1]), fmax(sh[k][2], sh[k][3]));
  __shared__ long long scv[2 * MAXF][BLOCK / 64];
#pragma unroll
  for (int k = 0; k < 2 * MAXF; ++k) {
    if ((k & (MAXF - 1)) < F) cv[k] = lane_reduce<64>(cv[k], OpAdd{});
    if (lane == 0) scv[k][wave] = cv[k];
  }
  __syncthreads();
  // r: wa value, unit sums (wa l1, force l1 x MAXF, raw l1 x MAXF), filler
  // sums (force l1 x MAXF, raw l1 x MAXF), then unit / filler kkt and flag
  if (threadIdx.x == 0) {
    EvalScalars *sc = t.scalars;
    double fw = 0.0;
    if (t.fixed_wa_mode == 1) {
      fw = ((sfw[0] + sfw[1]) + sfw[2]) + sfw[3];
      *t.fixed_wa = fw;
    } else if (t.fixed_wa_mode == 2) {
      fw = *t.fixed_wa;
    }
    sc->wa_value = r[0] + fw;
    sc->wa_l1 = r[1];
#pragma unroll
    for (int f = 0; f < MAXF; ++f) {
      sc->energy[f] = f < t.j.F ? t.field_coef[MAXF + f] : 0.0;
      sc->force_l1[f] = r[2 + f] + r[2 + 2 * MAXF + f];
      sc->raw_l1[f] = r[2 + MAXF + f] + r[2 + 3 * MAXF + f];
    }
    sc->kkt_bits = static_cast<unsigned long long>(__double_as_longlong(fmax(r[NS], r[NS + 2])));
    sc->nonfinite = (r[NS + 1] > 0.0 || r[NS + 3] > 0.0) ? 1u : 0u;
    sc->moved = sc->moved_acc;
    sc->moved_acc = 0u;
    // with no density pass this round the slot's partials are those of the
    // same point (the sums then repeat)
    if (t.nb_conv > 0) {
#pragma unroll
      for (int k = 0; k < MAXF; ++k) {
        sc->over_fx[k] = scv[k][0] + scv[k][1] + scv[k][2] + scv[k][3];
        sc->net_fx[k] = scv[MAXF + k][0] + scv[MAXF + k][1] + scv[MAXF + k][2] + scv[MAXF + k][3];
      }
    }
  }
}

// Gradient combination of one evaluation in one launch: per task the unit
// blocks, then the filler blocks.
__global__ void __launch_bounds__(BLOCK) k_gradient(const Task *__restrict__ tasks, BlockPlan plan) {
  int2 bm;
  if (!plan_block(plan, bm)) return;
  const Task &t = tasks[bm.x];
  // a rejected trial with a bounded gradient (the same flag for the block)
  if (t.trial && t.scalars->gradient_skipped) return;
  if (bm.y < t.nb_ug)
    unit_gradient_block(t, bm.y);
  else
    filler_gradient_block(t, bm.y - t.nb_ug);
}

// One block per task: every FP64 scalar of the evaluation.  (A last-block
// reduction inside k_gradient would need a device-scope fence per block,
// which writes back the XCD's L2.)
__global__ void __launch_bounds__(BLOCK) k_finalize(const Task *__restrict__ tasks) { finalize_task(tasks[blockIdx.x]); }

// ---------------------------------------------------------------------------
// Optimizer elementwise kernels.
// ---------------------------------------------------------------------------
// Optimizer elementwise work of one round in one launch: preconditioner
// blocks, then extrapolation blocks, then candidate blocks (different jobs,
// or different vectors of one job, so they do not depend on each other).
// diagonal_preconditioner (eplace.rs:1046-1100); per-block max/min/flag.
__device__ void opt_preconditioner(const OptTask &t, int lb) {
  const int i = lb * BLOCK + threadIdx.x;
  double d = 0.0;
  bool valid = false;
  if (i < t.n_units) {
    d = t.wl_diag[i];
    const int du = t.unit_du[i];
    if (du >= 0) {
      for (int f = 0; f < t.F; ++f) {
        const double charge = t.du_fcharge[(size_t)f * t.DU + du];
        if (charge != 0.0) d += t.lambda[f] * charge;
      }
    }
    valid = true;
  } else if (i < t.n_units + t.n_fillers) {
    const int k = i - t.n_units;
    d = 0.0 + t.lambda[t.fl_field[k]] * t.fl_charge[k];
    valid = true;
  }
  double lo = INFINITY, hi = 0.0;
  bool bad = false;
  if (valid) {
    bad = !isfinite(d) || d < 0.0;
    d = fmax(d, 1.0);  // elfPlace eq. (16) floor
    t.metric[2 * (size_t)i] = d;
    t.metric[2 * (size_t)i + 1] = d;
    lo = d;
    hi = d;
  }
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
    double *p = t.part_metric + 3 * (size_t)lb;
    p[0] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
    p[1] = fmin(fmin(smin[0], smin[1]), fmin(smin[2], smin[3]));
    p[2] = (sbad[0] | sbad[1] | sbad[2] | sbad[3]) ? 1.0 : 0.0;
  }
}

// y = clamp(x + beta (x - xp)); per-block flag "y differs from x".
__device__ void opt_extrapolate(const OptTask &t, int lb) {
  const size_t i = (size_t)lb * BLOCK + threadIdx.x;
  bool changed = false;
  if (i < t.n) {
    const double cur = t.x[i];
    double v = cur + t.beta * (cur - t.xp[i]);
    const double lo = t.lower[i], hi = t.upper[i];
    v = v < lo ? lo : (v > hi ? hi : v);
    t.y[i] = v;
    changed = v != cur;
  }
  const bool any = __syncthreads_or(changed);
  if (threadIdx.x == 0) t.part_step[4 * (size_t)lb + 3] = any ? 1.0 : 0.0;
}

// c = clamp(yp - gy / (L m)); per-block g.(c - y), sum m (c - y)^2, max |c - x|, changed.
__device__ void opt_candidate(const OptTask &t, int lb) {
  const size_t i = (size_t)lb * BLOCK + threadIdx.x;
  double red[2] = {0.0, 0.0};
  double change = 0.0;
  bool changed = false;
  if (i < t.n) {
    const double m = t.metric[i];
    const double yi = t.yp[i], g = t.gy[i];
    double c = yi - g / (t.lipschitz * m);
    const double lo = t.lower[i], hi = t.upper[i];
    c = c < lo ? lo : (c > hi ? hi : c);
    t.cand[i] = c;
    changed = c != yi;
    const double d = c - yi;
    red[0] = g * d;
    red[1] = m * d * d;
    change = fabs(c - t.origin[i]);
  }
  block_sum<2>(red, t.part_step + 4 * (size_t)lb);
  block_max_flag(change, changed, t.part_step + 4 * (size_t)lb + 2);
}

__global__ void __launch_bounds__(BLOCK) k_optimizer(const OptTask *__restrict__ tasks, OptMaps m) {
  const int g0 = plan_grid(m.precond), g1 = plan_grid(m.extrapolate);
  int b = blockIdx.x;
  int2 bm;
  if (b < g0) {
    if (plan_block_at(m.precond, b, bm)) opt_preconditioner(tasks[bm.x], bm.y);
    return;
  }
  b -= g0;
  if (b < g1) {
    if (plan_block_at(m.extrapolate, b, bm)) opt_extrapolate(tasks[bm.x], bm.y);
    return;
  }
  b -= g1;
  if (plan_block_at(m.candidate, b, bm)) opt_candidate(tasks[bm.x], bm.y);
}

// One block per task: the metric extremes (line search denominator checks,
// global_placement.rs:215-232), the extrapolation flag, and the candidate
// step's reductions, each in a fixed order.
__global__ void __launch_bounds__(BLOCK) k_optimizer_finalize(const OptTask *__restrict__ tasks) {
  const OptTask &t = tasks[blockIdx.x];
  if (t.do_precond) {
    const int nb = (t.n_units + t.n_fillers + BLOCK - 1) / BLOCK;
    __shared__ double smax[BLOCK / 64], smin[BLOCK / 64], sbad[BLOCK / 64];
    double mx = 0.0, mn = INFINITY, bad = 0.0;
    for (int i = threadIdx.x; i < nb; i += BLOCK) {
      mx = fmax(mx, t.part_metric[3 * (size_t)i]);
      mn = fmin(mn, t.part_metric[3 * (size_t)i + 1]);
      bad = fmax(bad, t.part_metric[3 * (size_t)i + 2]);
    }
    mx = wave_max(mx);
    mn = wave_min(mn);
    bad = wave_max(bad);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
      smax[wave] = mx;
      smin[wave] = mn;
      sbad[wave] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      t.metric_scalars->max_bits = static_cast<unsigned long long>(
          __double_as_longlong(fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]))));
      t.metric_scalars->min_bits = static_cast<unsigned long long>(
          __double_as_longlong(fmin(fmin(smin[0], smin[1]), fmin(smin[2], smin[3]))));
      t.metric_scalars->nonfinite = fmax(fmax(sbad[0], sbad[1]), fmax(sbad[2], sbad[3])) > 0.0 ? 1u : 0u;
    }
  }
  const int nb = (int)((t.n + BLOCK - 1) / BLOCK);
  if (t.do_extrapolate) {
    double f[1];
    max_partials<1>(t.part_step, nb, f, 4, 3);
    if (threadIdx.x == 0) t.step->differs = f[0] > 0.0 ? 1u : 0u;
  }
  if (t.do_candidate) {
    double v[2], mm[2];
    sum_partials<2>(t.part_step, nb, v, 4, 0);
    max_partials<2>(t.part_step, nb, mm, 4, 2);
    if (threadIdx.x == 0) {
      t.step->gdelta = v[0];
      t.step->msq = v[1];
      t.step->change_bits = static_cast<unsigned long long>(__double_as_longlong(mm[0]));
      t.step->changed = mm[1] > 0.0 ? 1u : 0u;
    }
  }
}

// ---------------------------------------------------------------------------
// RUDY demand (eplace.rs:856-937).  One wavefront per net: lanes reduce the
// bounding box, then stride over the candidate tiles.  The reference scans
// every tile and keeps those whose overlap with [min-1/2, max+1/2] is
// positive; that set lies inside t in (min-1, max+1), so the kernel visits
// floor(min-1)..ceil(max+1) and applies the reference's overlap formula,
// which keeps tile ranges and overlaps bit-exact.  Demand accumulates in
// 2^-40 fixed point so sums are run-to-run identical.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tile_overlap(int t, double rlo, double rhi) {
  const double tmin = static_cast<double>(t) - 0.5, tmax = static_cast<double>(t) + 0.5;
  const double o = fmin(tmax, rhi) - fmax(tmin, rlo);
  return o > 0.0 ? o : 0.0;
}

__global__ void __launch_bounds__(BLOCK) k_rudy(int n_nets, const uint32_t *__restrict__ net_start,
                                                const uint32_t *__restrict__ pin_unit,
                                                OffArr pin_xy,
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
  if (__all(same)) return;  // a net inside one unit is not routing demand
  if (!isfinite(xmin) || !isfinite(xmax) || !isfinite(ymin) || !isfinite(ymax)) {
    if (lane == 0) *error = 1u;
    return;
  }
  const double dx = xmax - xmin, dy = ymax - ymin;
  const double area = (dx + 1.0) * (dy + 1.0);
  const double hden = dx / area, vden = dy / area;
  const double rx0 = xmin - 0.5, rx1 = xmax + 0.5, ry0 = ymin - 0.5, ry1 = ymax + 0.5;
  const int x0 = static_cast<int>(fmax(floor(xmin - 1.0), 0.0));
  const int x1 = static_cast<int>(fmin(ceil(xmax + 1.0), static_cast<double>(W - 1)));
  const int y0 = static_cast<int>(fmax(floor(ymin - 1.0), 0.0));
  const int y1 = static_cast<int>(fmin(ceil(ymax + 1.0), static_cast<double>(H - 1)));
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

// Continuous bounding-box HPWL of a job's nets at `coords` (the reported
// final HPWL, over the pin coordinate map of analytical_placement.rs:589-597),
// FP64: the first nb_small blocks take one net per thread (nets of more than
// HPWL_THREAD_PINS pins contribute nothing there), the blocks after them one
// listed large net per wavefront; each block writes the sum of its nets, and
// one block adds the block sums in block order.  The same bits for a job
// alone or among others.
constexpr uint32_t HPWL_THREAD_PINS = 32;

__device__ __forceinline__ double2 hpwl_pin(const uint32_t *__restrict__ pin_unit, const OffArr &pin_xy,
                                            const double *__restrict__ coords, uint32_t p) {
  const uint32_t u = pin_unit[p];
  const double2 o = pin_xy[p];
  if (u == NO_UNIT) return o;
  return make_double2(o.x + coords[2 * (size_t)u], o.y + coords[2 * (size_t)u + 1]);
}

__global__ void __launch_bounds__(BLOCK) k_hpwl_nets(int n_nets, int nb_small, const uint32_t *__restrict__ net_start,
                                                     const uint32_t *__restrict__ pin_unit, OffArr pin_xy,
                                                     const double *__restrict__ coords,
                                                     const uint32_t *__restrict__ large, int n_large,
                                                     double *__restrict__ part) {
  double v[1] = {0.0};
  if ((int)blockIdx.x < nb_small) {
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    if (n < n_nets) {
      const uint32_t b = net_start[n], e = net_start[n + 1];
      if (e - b <= HPWL_THREAD_PINS) {
        double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
        for (uint32_t p = b; p < e; ++p) {
          const double2 c = hpwl_pin(pin_unit, pin_xy, coords, p);
          xmin = fmin(xmin, c.x);
          xmax = fmax(xmax, c.x);
          ymin = fmin(ymin, c.y);
          ymax = fmax(ymax, c.y);
        }
        v[0] = (xmax - xmin) + (ymax - ymin);
      }
    }
  } else {
    const int w = ((int)blockIdx.x - nb_small) * (BLOCK / 64) + (threadIdx.x >> 6);
    if (w < n_large) {  // wavefront-uniform
      const uint32_t n = large[w], b = net_start[n], e = net_start[n + 1];
      double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
      for (uint32_t p = b + (threadIdx.x & 63); p < e; p += 64) {
        const double2 c = hpwl_pin(pin_unit, pin_xy, coords, p);
        xmin = fmin(xmin, c.x);
        xmax = fmax(xmax, c.x);
        ymin = fmin(ymin, c.y);
        ymax = fmax(ymax, c.y);
      }
      xmin = wave_min(xmin);
      xmax = wave_max(xmax);
      ymin = wave_min(ymin);
      ymax = wave_max(ymax);
      if ((threadIdx.x & 63) == 0) v[0] = (xmax - xmin) + (ymax - ymin);
    }
  }
  block_sum<1>(v, part + blockIdx.x);
}

__global__ void __launch_bounds__(BLOCK) k_hpwl_total(int n_blocks, const double *__restrict__ part,
                                                      double *__restrict__ out) {
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < n_blocks; b += BLOCK) v[0] += part[b];
  block_sum<1>(v, out);
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

int blocks_for(size_t n) { return static_cast<int>((n + BLOCK - 1) / BLOCK); }

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
void launch_density(const Task *tasks, const Maps &m, hipStream_t s) {
  if (m.deposit.total > 0) {
    timer_begin(g_timer, "deposit", s);
    hipLaunchKernelGGL(k_deposit, dim3(plan_grid(m.deposit)), dim3(BLOCK), 0, s, tasks, m.deposit);
    timer_end(g_timer, "deposit", s, m.n_density);
  }
  if (m.convert.total > 0)
    hipLaunchKernelGGL(k_convert, dim3(plan_grid(m.convert)), dim3(BLOCK), 0, s, tasks, m.convert);
}

void launch_sort(const SortTask *tasks, const SortMaps &m, hipStream_t s) {
  if (m.charges.total <= 0) return;
  hipLaunchKernelGGL(k_sort_clear, dim3(plan_grid(m.keys)), dim3(BLOCK), 0, s, tasks, m.keys);
  hipLaunchKernelGGL(k_sort_count, dim3(plan_grid(m.charges)), dim3(BLOCK), 0, s, tasks, m.charges);
  hipLaunchKernelGGL(k_sort_scan, dim3(m.n_list), dim3(BLOCK), 0, s, tasks, m.list);
  hipLaunchKernelGGL(k_sort_scatter, dim3(plan_grid(m.charges)), dim3(BLOCK), 0, s, tasks, m.charges);
}

void launch_gather(const Task *tasks, const Maps &m, hipStream_t s) {
  if (m.n_tasks > 0) hipLaunchKernelGGL(k_coefficients, dim3(m.n_tasks), dim3(BLOCK), 0, s, tasks);
}

void launch_wa(const Task *tasks, const Maps &m, hipStream_t s) {
  if (m.wa.total <= 0) return;
  timer_begin(g_timer, "wa", s);
  hipLaunchKernelGGL(k_wa, dim3(plan_grid(m.wa)), dim3(BLOCK), 0, s, tasks, m.wa);
  timer_end(g_timer, "wa", s, m.wa_evaluations);
}

void launch_combine(const Task *tasks, const Maps &m, hipStream_t s) {
  if (m.gradient.total > 0) {
    timer_begin(g_timer, "gather", s);
    hipLaunchKernelGGL(k_gradient, dim3(plan_grid(m.gradient)), dim3(BLOCK), 0, s, tasks, m.gradient);
    timer_end(g_timer, "gather", s, m.n_gradient);
  }
  if (m.n_tasks > 0) hipLaunchKernelGGL(k_finalize, dim3(m.n_tasks), dim3(BLOCK), 0, s, tasks);
}

void launch_optimizer(const OptTask *tasks, const OptMaps &m, hipStream_t s, int passes) {
  const int grid = plan_grid(m.precond) + plan_grid(m.extrapolate) + plan_grid(m.candidate);
  if (grid <= 0) return;
  timer_begin(g_timer, "optimizer", s);
  hipLaunchKernelGGL(k_optimizer, dim3(grid), dim3(BLOCK), 0, s, tasks, m);
  timer_end(g_timer, "optimizer", s, passes);
}

void launch_optimizer_finalize(const OptTask *tasks, const OptMaps &m, hipStream_t s) {
  if (m.n_tasks > 0) hipLaunchKernelGGL(k_optimizer_finalize, dim3(m.n_tasks), dim3(BLOCK), 0, s, tasks);
}

void launch_rudy(int n_nets, const uint32_t *net_start, const uint32_t *pin_unit, OffArr pin_xy,
                 const double *coords, int W, int H, long long *fx, double *out, unsigned int *error, hipStream_t s) {
  const size_t G = (size_t)W * H;
  hipLaunchKernelGGL(k_fill_ll, dim3(std::max(1, blocks_for(2 * G))), dim3(BLOCK), 0, s, 2 * G, fx);
  const int nb = std::max(1, (n_nets + BLOCK / 64 - 1) / (BLOCK / 64));
  timer_begin(g_timer, "rudy", s);
  hipLaunchKernelGGL(k_rudy, dim3(nb), dim3(BLOCK), 0, s, n_nets, net_start, pin_unit, pin_xy, coords, W, H, fx,
                     fx + G, error);
  timer_end(g_timer, "rudy", s, 1);
  hipLaunchKernelGGL(k_fx_to_double, dim3(std::max(1, blocks_for(2 * G))), dim3(BLOCK), 0, s, 2 * G, fx, out);
}

int hpwl_blocks(int n_nets, int n_large) {
  return std::max(1, blocks_for((size_t)n_nets)) + (n_large + BLOCK / 64 - 1) / (BLOCK / 64);
}

void launch_hpwl(int n_nets, const uint32_t *net_start, const uint32_t *pin_unit, OffArr pin_xy,
                 const double *coords, const uint32_t *large, int n_large, double *part, double *out, hipStream_t s) {
  const int nb_small = std::max(1, blocks_for((size_t)n_nets)), nb = hpwl_blocks(n_nets, n_large);
  hipLaunchKernelGGL(k_hpwl_nets, dim3(nb), dim3(BLOCK), 0, s, n_nets, nb_small, net_start, pin_unit, pin_xy, coords,
                     large, n_large, part);
  hipLaunchKernelGGL(k_hpwl_total, dim3(1), dim3(BLOCK), 0, s, nb, part, out);
}

}  // namespace texo
