// gp_seed.hip — the coarse quadratic seed that produces eplace::place's
// initial coordinates (analytical_place, crates/texo-pnr/src/lib.rs:4042-4066):
// per axis, bound-to-bound edges at the device centre
// (linearize_axis, crates/texo-pnr/src/analytical_placement.rs:283-356), the
// axis equations (analytical_axis_equations, lib.rs:4464-4527) with one
// gauge unit per floating component (finalize_component_gauges,
// lib.rs:4419-4461), a Jacobi-preconditioned CG solve (solve_quadratic,
// lib.rs:4586-4692) and the re-centring of floating components
// (solve_analytical_axis, lib.rs:4529-4556).
//
// The graph work (edges, CSR rows in insertion order, components) is
// integer and order-dependent bookkeeping done once on the host; the solve
// runs on the GPU: both axes' systems in the same launches, one thread per
// row for the product (the row's terms folded in the reference's order, so
// each product is the reference's bit for bit), block partials of the dot
// products written to mapped host memory and summed on the host in a fixed
// order, and the CG scalars and stop decisions on the host from those sums.

#include "gp_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

namespace texo {
namespace {

constexpr double WEIGHT_SCALE = 64.0;  // PLACEMENT_WEIGHT_SCALE, analytical_placement.rs:13
constexpr double RESIDUAL_TOLERANCE = 1e-8;  // squared relative residual, lib.rs:4594

struct SeedPin {
  uint32_t unit, cell;
};

struct SeedEdge {
  uint32_t left, right, left_cell, right_cell;
  double weight;
};

// One axis system in device memory (CSR rows keep the reference's
// insertion order).
struct AxisSystem {
  int n;
  const uint32_t *row_start, *col;
  const double *weight, *diagonal, *rhs;
  double *x, *r, *z, *p, *ap;
  double *partial;  // [blocks][2] in mapped host memory
  int active;
};
struct SeedSystems {
  AxisSystem s[2];
  int blocks[2];  // blocks of each system (grid = blocks[0] + blocks[1])
};

__device__ __forceinline__ bool seed_row(const SeedSystems &ss, int &sys, int &row, int &lb) {
  const int b = blockIdx.x;
  sys = b < ss.blocks[0] ? 0 : 1;
  lb = sys ? b - ss.blocks[0] : b;
  row = lb * BLOCK + threadIdx.x;
  return true;
}

// multiply_quadratic (lib.rs:4767-4782) of one row: the diagonal term, then
// minus each neighbour term in insertion order.
__device__ __forceinline__ double row_product(const AxisSystem &a, int i, const double *v) {
  double sum = a.diagonal[i] * v[i];
  for (uint32_t k = a.row_start[i]; k < a.row_start[i + 1]; ++k) sum = sum - a.weight[k] * v[a.col[k]];
  return sum;
}

__device__ __forceinline__ void block_sums2(double v0, double v1, double *out) {
  __shared__ double sh[2][BLOCK / 64];
  v0 = lane_reduce<64>(v0, OpAdd{});
  v1 = lane_reduce<64>(v1, OpAdd{});
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wave] = v0;
    sh[1][wave] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
    out[1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
  }
}

// mode 0: ap = A p, partials (p . ap, 0).
// mode 1: r = rhs - A x, z = r / d, partials (r . r, r . z).
__global__ void __launch_bounds__(BLOCK) k_seed_product(SeedSystems ss, int mode) {
  int sys, i, lb;
  seed_row(ss, sys, i, lb);
  const AxisSystem &a = ss.s[sys];
  if (!a.active) return;  // whole block
  double v0 = 0.0, v1 = 0.0;
  if (i < a.n) {
    if (mode == 0) {
      const double q = row_product(a, i, a.p);
      a.ap[i] = q;
      v0 = a.p[i] * q;
    } else {
      const double res = a.rhs[i] - row_product(a, i, a.x);
      const double z = res / a.diagonal[i];
      a.r[i] = res;
      a.z[i] = z;
      v0 = res * res;
      v1 = res * z;
    }
  }
  block_sums2(v0, v1, a.partial + 2 * (size_t)lb);
}

// x += alpha p, r -= alpha ap, z = r / d; partials (r . r, r . z).
__global__ void __launch_bounds__(BLOCK) k_seed_update(SeedSystems ss, double alpha0, double alpha1) {
  int sys, i, lb;
  seed_row(ss, sys, i, lb);
  const AxisSystem &a = ss.s[sys];
  if (!a.active) return;
  const double alpha = sys ? alpha1 : alpha0;
  double v0 = 0.0, v1 = 0.0;
  if (i < a.n) {
    a.x[i] += alpha * a.p[i];
    const double res = a.r[i] - alpha * a.ap[i];
    const double z = res / a.diagonal[i];
    a.r[i] = res;
    a.z[i] = z;
    v0 = res * res;
    v1 = res * z;
  }
  block_sums2(v0, v1, a.partial + 2 * (size_t)lb);
}

// p = z + beta p (beta < 0: p = z, a restart from the preconditioned residual).
__global__ void __launch_bounds__(BLOCK) k_seed_direction(SeedSystems ss, double beta0, double beta1) {
  int sys, i, lb;
  seed_row(ss, sys, i, lb);
  const AxisSystem &a = ss.s[sys];
  if (!a.active || i >= a.n) return;
  const double beta = sys ? beta1 : beta0;
  a.p[i] = beta < 0.0 ? a.z[i] : a.z[i] + beta * a.p[i];
}

// Host side of one axis: the equations, then the gauge.
// The off-diagonal terms are a CSR whose rows keep insertion order.
struct HostAxis {
  std::vector<double> diagonal, rhs, initial;
  std::vector<uint32_t> start, col;
  std::vector<double> weight;
  std::vector<uint8_t> anchored;
  std::vector<std::vector<uint32_t>> floating;
  uint32_t center = 0;
};

// u64_to_f64 (analytical_placement.rs:705-709)
double u64_to_f64(uint64_t v) {
  return (double)(uint32_t)(v >> 32) * 4294967296.0 + (double)(uint32_t)(v & 0xffffffffu);
}

// f64::total_cmp's order on finite values: -0.0 sorts below +0.0
bool total_less(double a, double b) {
  if (a == b) return std::signbit(a) && !std::signbit(b);
  return a < b;
}

// axis_bounds (analytical_placement.rs:793-815): the lowest and highest pin
// by coordinate (ties: the first lowest, the last highest); a collapsed pair
// on one unit takes the last pin of another unit as the upper bound.
bool axis_bounds(const SeedPin *pins, size_t n, const double *pos, size_t &lower, size_t &upper) {
  if (n == 0) return false;
  bool same = true;
  for (size_t i = 1; i < n; ++i) same = same && pins[i].unit == pins[0].unit;
  if (same) return false;
  lower = upper = 0;
  for (size_t i = 1; i < n; ++i) {
    if (total_less(pos[i], pos[lower])) lower = i;
    if (!total_less(pos[i], pos[upper])) upper = i;
  }
  if (pos[lower] == pos[upper] && pins[lower].unit == pins[upper].unit) {
    for (size_t i = n; i-- > 0;)
      if (pins[i].unit != pins[lower].unit) {
        upper = i;
        return true;
      }
    return false;
  }
  return true;
}

// finalize_component_gauges (lib.rs:4419-4461): one gauge row per component
// with no anchored unit; returns the floating components (sorted units).
void gauge(HostAxis &q) {
  const size_t n = q.diagonal.size();
  std::vector<uint8_t> visited(n, 0), g(n, 0);
  std::vector<uint32_t> pending, component;
  for (size_t start = 0; start < n; ++start) {
    if (visited[start]) continue;
    visited[start] = 1;
    pending.assign(1, (uint32_t)start);
    component.clear();
    while (!pending.empty()) {
      const uint32_t u = pending.back();
      pending.pop_back();
      component.push_back(u);
      for (uint32_t k = q.start[u]; k < q.start[u + 1]; ++k)
        if (!visited[q.col[k]]) {
          visited[q.col[k]] = 1;
          pending.push_back(q.col[k]);
        }
    }
    bool anchored = false;
    for (uint32_t u : component) anchored = anchored || q.anchored[u];
    if (anchored) continue;
    std::sort(component.begin(), component.end());
    g[component[0]] = 1;
    q.floating.push_back(component);
  }
  // the gauge rows become x = 0; every other row drops its terms on them
  size_t out = 0;
  uint32_t row_begin = 0;
  for (size_t u = 0; u < n; ++u) {
    const uint32_t b = row_begin, e = q.start[u + 1];
    row_begin = e;
    if (g[u]) {
      q.diagonal[u] = 1.0;
      q.rhs[u] = 0.0;
    } else {
      for (uint32_t k = b; k < e; ++k)
        if (!g[q.col[k]]) {
          q.col[out] = q.col[k];
          q.weight[out] = q.weight[k];
          ++out;
        }
    }
    q.start[u + 1] = (uint32_t)out;
  }
  q.col.resize(out);
  q.weight.resize(out);
  for (auto &c : q.floating) {
    const double shift = q.initial[c[0]];
    for (uint32_t u : c) q.initial[u] -= shift;
  }
}

}  // namespace

// solve_quadratic of both axes on `device`; on return each axis holds its
// solution in `initial` (re-centred), or the call fails with the reference's
// non-convergence (DID_NOT_CONVERGE: iterations, axis, residual ratio).
uint64_t solve_axes(std::vector<HostAxis *> axes, int device) {
  HIP(hipSetDevice(device));
  hipStream_t s = nullptr;
  HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  std::vector<void *> device_blocks;
  double *mapped = nullptr;
  auto cleanup = [&] {
    if (s) (void)hipStreamSynchronize(s);
    for (void *p : device_blocks) (void)hipFree(p);
    if (mapped) (void)hipHostFree(mapped);
    if (s) (void)hipStreamDestroy(s);
  };
  uint64_t total_iterations = 0;
  try {
    SeedSystems ss{};
    const int na = (int)axes.size();
    size_t partial_total = 0;
    for (int a = 0; a < na; ++a) partial_total += 2 * (size_t)std::max(1, (int)((axes[a]->diagonal.size() + BLOCK - 1) / BLOCK));
    HIP(hipHostMalloc(reinterpret_cast<void **>(&mapped), partial_total * sizeof(double),
                      hipHostMallocMapped | hipHostMallocCoherent));
    double *mapped_dev = nullptr;
    HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&mapped_dev), mapped, 0));
    size_t partial_at = 0;
    for (int a = 0; a < 2; ++a) {
      AxisSystem &d = ss.s[a];
      if (a >= na) {
        d.active = 0;
        ss.blocks[a] = 0;
        continue;
      }
      HostAxis &q = *axes[a];
      const size_t n = q.diagonal.size();
      const std::vector<uint32_t> &start = q.start, &col = q.col;
      const std::vector<double> &w = q.weight;
      const size_t nnz = std::max<size_t>(col.size(), 1);
      char *blk = nullptr;
      const size_t o_start = 0, o_col = (4 * (n + 1) + 255) & ~size_t(255);
      const size_t o_w = (o_col + 4 * nnz + 255) & ~size_t(255);
      const size_t o_vec = (o_w + 8 * nnz + 255) & ~size_t(255);
      const size_t vec = (8 * n + 255) & ~size_t(255);
      HIP(hipMalloc(&blk, o_vec + 7 * vec));
      device_blocks.push_back(blk);
      HIP(hipMemcpyAsync(blk + o_start, start.data(), 4 * (n + 1), hipMemcpyHostToDevice, s));
      if (!col.empty()) {
        HIP(hipMemcpyAsync(blk + o_col, col.data(), 4 * col.size(), hipMemcpyHostToDevice, s));
        HIP(hipMemcpyAsync(blk + o_w, w.data(), 8 * w.size(), hipMemcpyHostToDevice, s));
      }
      double *v[7];
      for (int k = 0; k < 7; ++k) v[k] = reinterpret_cast<double *>(blk + o_vec + k * vec);
      HIP(hipMemcpyAsync(v[0], q.diagonal.data(), 8 * n, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(v[1], q.rhs.data(), 8 * n, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(v[2], q.initial.data(), 8 * n, hipMemcpyHostToDevice, s));
      d.n = (int)n;
      d.row_start = reinterpret_cast<const uint32_t *>(blk + o_start);
      d.col = reinterpret_cast<const uint32_t *>(blk + o_col);
      d.weight = reinterpret_cast<const double *>(blk + o_w);
      d.diagonal = v[0];
      d.rhs = v[1];
      d.x = v[2];
      d.r = v[3];
      d.z = v[4];
      d.p = v[5];
      d.ap = v[6];
      ss.blocks[a] = std::max(1, (int)((n + BLOCK - 1) / BLOCK));
      d.partial = mapped_dev + partial_at;
      partial_at += 2 * (size_t)ss.blocks[a];
      d.active = 1;
    }
    const int grid = ss.blocks[0] + ss.blocks[1];
    // fixed-order sums of the block partials (component k of each pair)
    auto sums = [&](int a, int k) {
      const double *p = mapped + (ss.s[a].partial - mapped_dev);
      double t = 0.0;
      for (int b = 0; b < ss.blocks[a]; ++b) t += p[2 * b + k];
      return t;
    };
    auto run = [&](auto launch) {
      launch();
      HIP(hipGetLastError());
      HIP(hipStreamSynchronize(s));
    };
    struct State {
      double rr = 0, target = 0, rho = 0, r0 = 0;
      uint64_t iterations = 0;
      bool done = false, breakdown = false;
    } st[2];
    // r = b - A x, z = D^-1 r, rho = r . z, p = z
    run([&] { hipLaunchKernelGGL(k_seed_product, dim3(grid), dim3(BLOCK), 0, s, ss, 1); });
    for (int a = 0; a < na; ++a) {
      st[a].rr = sums(a, 0);
      st[a].r0 = st[a].rr;
      st[a].target = std::max(st[a].r0, DBL_EPSILON) * RESIDUAL_TOLERANCE;
      st[a].rho = sums(a, 1);
    }
    hipLaunchKernelGGL(k_seed_direction, dim3(grid), dim3(BLOCK), 0, s, ss, -1.0, -1.0);
    auto live = [&] {
      bool any = false;
      for (int a = 0; a < na; ++a) {
        ss.s[a].active = st[a].done ? 0 : 1;
        any = any || !st[a].done;
      }
      return any;
    };
    for (int a = 0; a < na; ++a)
      if (st[a].rr <= st[a].target || ss.s[a].n == 0) st[a].done = true;
    while (live()) {
      // q = A p, denominator p . q (lib.rs:4618-4630)
      run([&] { hipLaunchKernelGGL(k_seed_product, dim3(grid), dim3(BLOCK), 0, s, ss, 0); });
      double alpha[2] = {0.0, 0.0};
      for (int a = 0; a < na; ++a) {
        if (st[a].done) continue;
        const double den = sums(a, 0);
        if (!std::isfinite(den) || den <= 0.0 || !std::isfinite(st[a].rho) || st[a].rho <= 0.0) {
          st[a].done = st[a].breakdown = true;
          continue;
        }
        alpha[a] = st[a].rho / den;
        if (!std::isfinite(alpha[a])) st[a].done = st[a].breakdown = true;
      }
      if (!live()) break;
      run([&] { hipLaunchKernelGGL(k_seed_update, dim3(grid), dim3(BLOCK), 0, s, ss, alpha[0], alpha[1]); });
      double beta[2] = {-1.0, -1.0};
      bool recheck = false;
      bool recheck_axis[2] = {false, false};
      for (int a = 0; a < na; ++a) {
        if (st[a].done) continue;
        st[a].iterations += 1;
        st[a].rr = sums(a, 0);
        if (st[a].rr <= st[a].target) {
          recheck = recheck_axis[a] = true;
          continue;
        }
        const double next = sums(a, 1);
        if (!std::isfinite(next) || next <= 0.0) {
          st[a].done = st[a].breakdown = true;
          continue;
        }
        beta[a] = next / st[a].rho;
        st[a].rho = next;
        if (st[a].iterations >= (uint64_t)ss.s[a].n) st[a].done = true;  // Krylov dimension bound
      }
      if (recheck) {
        // the true residual before accepting convergence (lib.rs:4641-4660):
        // only the axes that reached the target recompute theirs
        for (int a = 0; a < na; ++a) ss.s[a].active = recheck_axis[a] ? 1 : 0;
        run([&] { hipLaunchKernelGGL(k_seed_product, dim3(grid), dim3(BLOCK), 0, s, ss, 1); });
        for (int a = 0; a < na; ++a) {
          if (!recheck_axis[a]) continue;
          st[a].rr = sums(a, 0);
          if (st[a].rr <= st[a].target) {
            st[a].done = true;
            continue;
          }
          st[a].rho = sums(a, 1);
          if (!std::isfinite(st[a].rho) || st[a].rho <= 0.0) {
            st[a].done = st[a].breakdown = true;
            continue;
          }
          beta[a] = -1.0;  // restart the recurrence from the new residual
          if (st[a].iterations >= (uint64_t)ss.s[a].n) st[a].done = true;
        }
      }
      if (!live()) break;
      hipLaunchKernelGGL(k_seed_direction, dim3(grid), dim3(BLOCK), 0, s, ss, beta[0], beta[1]);
    }
    // finish_quadratic_solve (lib.rs:4702-4752): the true residual decides
    for (int a = 0; a < na; ++a) ss.s[a].active = 1;
    run([&] { hipLaunchKernelGGL(k_seed_product, dim3(grid), dim3(BLOCK), 0, s, ss, 1); });
    for (int a = 0; a < na; ++a) {
      const double fr = sums(a, 0);
      total_iterations += st[a].iterations;
      if (!(fr <= st[a].target)) {
        Failure f{TEXO_GP_ERR_DID_NOT_CONVERGE};
        f.index = (int64_t)st[a].iterations;
        f.index2 = a;
        f.kind = st[a].breakdown ? 1u : 0u;  // the reference's stop: breakdown, else dimension
        f.v0 = st[a].r0 > 0.0 ? std::sqrt(fr / st[a].r0) : 0.0;
        throw f;
      }
      HostAxis &q = *axes[a];
      HIP(hipMemcpyAsync(q.initial.data(), ss.s[a].x, 8 * q.initial.size(), hipMemcpyDeviceToHost, s));
    }
    HIP(hipStreamSynchronize(s));
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  // re-centre each floating component on the device centre
  for (HostAxis *q : axes)
    for (const auto &c : q->floating) {
      double mean = 0.0;
      for (uint32_t u : c) mean += q->initial[u];
      mean /= (double)c.size();
      const double t = (double)q->center - mean;
      for (uint32_t u : c) q->initial[u] += t;
    }
  return total_iterations;
}

}  // namespace texo

extern "C" {

int texo_gp_solve_axis(int32_t device, uint32_t n, const double *diagonal, const uint32_t *adjacency_start,
                       const uint32_t *adjacency_other, const double *adjacency_weight, const double *rhs,
                       const uint8_t *anchored, const double *initial, uint32_t center, double *solution,
                       uint64_t *iterations) {
  if (!diagonal || !adjacency_start || !rhs || !anchored || !initial || !solution) return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    texo::HostAxis q;
    q.diagonal.assign(diagonal, diagonal + n);
    q.rhs.assign(rhs, rhs + n);
    q.anchored.assign(anchored, anchored + n);
    q.initial.assign(initial, initial + n);
    q.center = center;
    if (adjacency_start[0] != 0) return TEXO_GP_ERR_INVALID_ARGUMENT;
    for (uint32_t u = 0; u < n; ++u)
      if (adjacency_start[u + 1] < adjacency_start[u]) return TEXO_GP_ERR_INVALID_ARGUMENT;
    const uint32_t nnz = adjacency_start[n];
    if (nnz && (!adjacency_other || !adjacency_weight)) return TEXO_GP_ERR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < nnz; ++k)
      if (adjacency_other[k] >= n) return TEXO_GP_ERR_INVALID_ARGUMENT;
    q.start.assign(adjacency_start, adjacency_start + n + 1);
    q.col.assign(adjacency_other, adjacency_other + nnz);
    q.weight.assign(adjacency_weight, adjacency_weight + nnz);
    texo::gauge(q);
    const uint64_t it = texo::solve_axes({&q}, device);
    if (iterations) *iterations = it;
    std::copy(q.initial.begin(), q.initial.end(), solution);
    return 0;
  } catch (...) {
    const texo::Failure f = texo::current_failure();
    if (iterations && f.status == TEXO_GP_ERR_DID_NOT_CONVERGE) *iterations = (uint64_t)f.index;
    return f.status;
  }
}

int texo_gp_quadratic_seed(const texo_gp_problem *p, int32_t device, double *out_xy, texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!p || !out_xy || p->abi_version != TEXO_GP_ABI_VERSION || p->width == 0 || p->height == 0)
    return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    const size_t U = p->n_units, NC = p->n_cells;
    for (size_t c = 0; c < NC; ++c)
      if (p->cell_unit[c] >= U) return TEXO_GP_ERR_INVALID_ARGUMENT;
    const uint32_t n_pins = p->n_nets ? p->net_pin_start[p->n_nets] : 0;
    for (uint32_t k = 0; k < n_pins; ++k)
      if (p->pin_cell[k] >= NC) return TEXO_GP_ERR_INVALID_ARGUMENT;
    for (uint32_t a = 0; a < p->n_arcs; ++a)
      if (p->arc_cells[2 * a] >= NC || p->arc_cells[2 * a + 1] >= NC) return TEXO_GP_ERR_INVALID_ARGUMENT;
    texo::HostAxis axes[2];
    std::vector<texo::SeedPin> pins(n_pins);
    for (uint32_t k = 0; k < n_pins; ++k) pins[k] = {p->cell_unit[p->pin_cell[k]], p->pin_cell[k]};
    auto assemble = [&](int axis) {
      texo::HostAxis &q = axes[axis];
      q.center = axis ? p->height / 2 : p->width / 2;
      const double center = (double)q.center;
      auto coord = [&](const texo::SeedPin &pin) {
        return p->cell_fixed[pin.cell] ? p->cell_fixed_xy[2 * pin.cell + axis]
                                       : center + p->cell_offset_xy[2 * pin.cell + axis];
      };
      auto offset = [&](uint32_t cell) { return p->cell_offset_xy[2 * cell + axis]; };
      // linearize_axis at the centre (analytical_placement.rs:283-356)
      std::vector<texo::SeedEdge> edges;
      edges.reserve(2 * (size_t)n_pins + p->n_arcs);
      std::vector<double> pp;
      for (uint32_t n = 0; n < p->n_nets; ++n) {
        const uint32_t b = p->net_pin_start[n], e = p->net_pin_start[n + 1];
        const size_t m = e - b;
        pp.resize(m);
        for (size_t i = 0; i < m; ++i) pp[i] = coord(pins[b + i]);
        size_t lower = 0, upper = 0;
        if (!texo::axis_bounds(pins.data() + b, m, pp.data(), lower, upper)) continue;
        const double fanout = (double)(uint32_t)std::max<size_t>(m - 1, 1);
        for (size_t i = 0; i < m; ++i)
          for (size_t bound : {lower, upper}) {
            if (i == bound || (i == upper && bound == lower)) continue;
            const texo::SeedPin l = pins[b + i], r = pins[b + bound];
            if (l.unit == r.unit) continue;
            const double separation = std::max(std::fabs(pp[i] - pp[bound]), 1.0);
            edges.push_back({l.unit, r.unit, l.cell, r.cell, texo::WEIGHT_SCALE / (fanout * separation)});
          }
      }
      for (uint32_t a = 0; a < p->n_arcs; ++a) {
        const texo::SeedPin d{p->cell_unit[p->arc_cells[2 * a]], p->arc_cells[2 * a]};
        const texo::SeedPin t{p->cell_unit[p->arc_cells[2 * a + 1]], p->arc_cells[2 * a + 1]};
        if (d.unit == t.unit) continue;
        const double separation = std::max(std::fabs(coord(d) - coord(t)), 1.0);
        edges.push_back({d.unit, t.unit, d.cell, t.cell,
                         texo::WEIGHT_SCALE * texo::u64_to_f64(p->arc_extra_weight[a]) / separation});
      }
      // analytical_axis_equations (lib.rs:4464-4527); each row's terms in
      // edge order
      q.diagonal.assign(U, 0.0);
      q.rhs.assign(U, 0.0);
      q.anchored.assign(U, 0);
      q.initial.assign(U, center);
      q.start.assign(U + 1, 0);
      for (size_t u = 0; u < U; ++u)
        if (p->unit_fixed[u]) {
          q.diagonal[u] = 1.0;
          q.rhs[u] = p->unit_bounds[4 * u + axis];
          q.anchored[u] = 1;
        }
      for (const texo::SeedEdge &e : edges) {
        const double lo = offset(e.left_cell), ro = offset(e.right_cell), w = e.weight;
        const bool lf = p->unit_fixed[e.left] != 0, rf = p->unit_fixed[e.right] != 0;
        if (lf && rf) continue;
        if (lf) {
          q.diagonal[e.right] += w;
          q.rhs[e.right] += w * (p->cell_fixed_xy[2 * e.left_cell + axis] - ro);
          q.anchored[e.right] = 1;
        } else if (rf) {
          q.diagonal[e.left] += w;
          q.rhs[e.left] += w * (p->cell_fixed_xy[2 * e.right_cell + axis] - lo);
          q.anchored[e.left] = 1;
        } else {
          q.diagonal[e.left] += w;
          q.diagonal[e.right] += w;
          q.start[e.left + 1] += 1;
          q.start[e.right + 1] += 1;
          q.rhs[e.left] += w * (ro - lo);
          q.rhs[e.right] += w * (lo - ro);
        }
      }
      for (size_t u = 0; u < U; ++u) q.start[u + 1] += q.start[u];
      q.col.resize(q.start[U]);
      q.weight.resize(q.start[U]);
      std::vector<uint32_t> at(q.start.begin(), q.start.end() - 1);
      for (const texo::SeedEdge &e : edges) {
        if (p->unit_fixed[e.left] || p->unit_fixed[e.right]) continue;
        q.col[at[e.left]] = e.right;
        q.weight[at[e.left]++] = e.weight;
        q.col[at[e.right]] = e.left;
        q.weight[at[e.right]++] = e.weight;
      }
      texo::gauge(q);
    };
    // the two axes are independent: y on a second host thread
    std::exception_ptr y_error;
    std::thread y_thread([&] {
      try {
        assemble(1);
      } catch (...) {
        y_error = std::current_exception();
      }
    });
    try {
      assemble(0);
    } catch (...) {
      y_thread.join();
      throw;
    }
    y_thread.join();
    if (y_error) std::rethrow_exception(y_error);
    const uint64_t it = texo::solve_axes({&axes[0], &axes[1]}, device);
    for (size_t u = 0; u < U; ++u) {
      out_xy[2 * u] = axes[0].initial[u];
      out_xy[2 * u + 1] = axes[1].initial[u];
    }
    if (report) report->iterations = it;
    return 0;
  } catch (...) {
    const texo::Failure f = texo::current_failure();
    if (report) {
      report->status = f.status;
      report->err_index = f.index;
      report->err_index2 = f.index2;
      report->err_kind = f.kind;
      report->err_value[0] = f.v0;
    }
    return f.status;
  }
}

}  // extern "C"
