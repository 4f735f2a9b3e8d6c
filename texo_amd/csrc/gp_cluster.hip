// gp_cluster.hip — register-clustering areas of an area round on the GPU
// (clustering_areas, crates/texo-pnr/src/register_clustering.rs:17-81).
//
// For every movable register r, the expected number of registers of its
// (clock, LSR) group that fall in a +-2.5 sigma window around it, split by
// clock-enable: E_ce(r) = sum over the group's registers o (in input order)
// of P_x(o) * P_y(o), with P the Gaussian interval probability of o's
// coordinate (erf, register_clustering.rs:83-86).  The area follows from the
// per-CE slice demands (soft_division_ceiling, :88-96).
//
// The reference evaluates O(R^2) pairs per group on the host (rayon over
// registers).  Here one thread owns one focal register: a block takes up to
// BLOCK focal registers of one group, stages the group's registers through
// LDS in tiles and every thread accumulates its per-CE sums in LDS, adding
// the pairs in the group's order exactly as the reference does; the erf
// values are the device library's, so areas agree with the reference to its
// last bits only where erf does.

#include "gp_host.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

namespace texo {
namespace {

constexpr int CLUSTER_KMAX = 16;  // clock-enable classes per (clock, LSR) group held in LDS

// One block: focal registers [begin, end) of the group whose registers are
// [gbegin, gend) (group-major arrays).
struct ClusterBlock {
  int begin, end, gbegin, gend, k;
};

__device__ __forceinline__ double interval_probability(double low, double high, double mean, double scale) {
  const double p = 0.5 * (erf((high - mean) / scale) - erf((low - mean) / scale));
  return p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
}

__device__ __forceinline__ double soft_ceiling(double value, double divisor) {
  const double q = value / divisor;
  const double fl = floor(q);
  if (q - fl < 1.0 / divisor) return value + (1.0 - divisor) * fl;
  return ceil(q);
}

__global__ void __launch_bounds__(BLOCK) k_cluster(const ClusterBlock *__restrict__ blocks,
                                                   const double2 *__restrict__ xy, const uint8_t *__restrict__ ce,
                                                   double half_window, double scale, double *__restrict__ area,
                                                   unsigned int *__restrict__ invalid) {
  const ClusterBlock b = blocks[blockIdx.x];
  __shared__ double acc[CLUSTER_KMAX][BLOCK];
  __shared__ double2 t_xy[BLOCK];
  __shared__ uint8_t t_ce[BLOCK];
  const int f = b.begin + (int)threadIdx.x;
  const bool active = f < b.end;
  for (int k = 0; k < b.k; ++k) acc[k][threadIdx.x] = 0.0;
  double2 me = make_double2(0.0, 0.0);
  if (active) me = xy[f];
  const double lox = me.x - half_window, hix = me.x + half_window;
  const double loy = me.y - half_window, hiy = me.y + half_window;
  for (int t0 = b.gbegin; t0 < b.gend; t0 += BLOCK) {
    const int n = min(BLOCK, b.gend - t0);
    __syncthreads();
    if ((int)threadIdx.x < n) {
      t_xy[threadIdx.x] = xy[t0 + threadIdx.x];
      t_ce[threadIdx.x] = ce[t0 + threadIdx.x];
    }
    __syncthreads();
    if (active) {
      for (int j = 0; j < n; ++j) {
        const double2 o = t_xy[j];
        const double px = interval_probability(lox, hix, o.x, scale);
        const double py = interval_probability(loy, hiy, o.y, scale);
        acc[t_ce[j]][threadIdx.x] += px * py;
      }
    }
  }
  if (!active) return;
  const double focal = acc[ce[f]][threadIdx.x];
  if (!(isfinite(focal) && focal > 0.0)) {
    atomicOr(invalid, 1u);
    return;
  }
  double total = 0.0;  // clock-enable classes in ascending CE order
  for (int k = 0; k < b.k; ++k) total += soft_ceiling(acc[k][threadIdx.x], 2.0);
  const double a = 8.0 / focal * soft_ceiling(focal, 2.0) / total * soft_ceiling(total, 4.0);
  if (!(isfinite(a) && a >= 0.0)) {
    atomicOr(invalid, 1u);
    return;
  }
  area[f] = a;
}

}  // namespace

int register_clustering_device(const std::vector<std::pair<uint32_t, std::pair<double, double>>> &regs,
                               const std::map<uint32_t, Control> &controls, uint64_t movable,
                               std::vector<double> &areas, BlockPool *pool, hipStream_t s) {
  areas.assign(regs.size(), 0.0);
  if (regs.empty() || movable == 0) return 1;
  if (movable > 0xffffffffull) return 0;
  const double sigma = std::sqrt(1.0e-5 * (double)(uint32_t)movable);
  if (!std::isfinite(sigma) || sigma <= 0.0) return 0;
  // registers grouped by (clock, LSR), each group in input order; CE values
  // numbered within the group in ascending order (the reference's BTreeMap)
  std::map<std::pair<uint64_t, uint64_t>, std::vector<size_t>> groups;
  for (size_t i = 0; i < regs.size(); ++i) {
    auto it = controls.find(regs[i].first);
    if (it == controls.end()) return 0;
    groups[{it->second.clock, it->second.lsr}].push_back(i);
  }
  std::vector<double2> xy(regs.size());
  std::vector<uint8_t> ce(regs.size());
  std::vector<size_t> order(regs.size());
  std::vector<ClusterBlock> blocks;
  size_t at = 0;
  for (auto &kv : groups) {
    std::map<uint64_t, int> ce_index;
    for (size_t i : kv.second) ce_index[controls.at(regs[i].first).ce] = 0;
    if (ce_index.size() > (size_t)CLUSTER_KMAX) return -1;  // caller falls back to the host
    int k = 0;
    for (auto &c : ce_index) c.second = k++;
    const int gbegin = (int)at;
    for (size_t i : kv.second) {
      xy[at] = make_double2(regs[i].second.first, regs[i].second.second);
      ce[at] = (uint8_t)ce_index[controls.at(regs[i].first).ce];
      order[at] = i;
      ++at;
    }
    const int gend = (int)at;
    for (int b = gbegin; b < gend; b += BLOCK) blocks.push_back({b, std::min(b + BLOCK, gend), gbegin, gend, k});
  }
  const size_t n = regs.size();
  const size_t o_blocks = 0, o_xy = (blocks.size() * sizeof(ClusterBlock) + 255) & ~size_t(255);
  const size_t o_ce = (o_xy + n * sizeof(double2) + 255) & ~size_t(255);
  const size_t o_area = (o_ce + n + 255) & ~size_t(255);
  const size_t o_flag = (o_area + n * sizeof(double) + 255) & ~size_t(255);
  const size_t bytes = o_flag + 256;
  size_t cap = 0;
  char *ws = nullptr;
  if (pool) {
    ws = static_cast<char *>(pool->acquire(bytes, &cap));
  } else {
    HIP(hipMalloc(&ws, bytes));
  }
  std::vector<double> out(n);
  unsigned int invalid = 0;
  try {
    HIP(hipMemcpyAsync(ws + o_blocks, blocks.data(), blocks.size() * sizeof(ClusterBlock), hipMemcpyHostToDevice, s));
    HIP(hipMemcpyAsync(ws + o_xy, xy.data(), n * sizeof(double2), hipMemcpyHostToDevice, s));
    HIP(hipMemcpyAsync(ws + o_ce, ce.data(), n, hipMemcpyHostToDevice, s));
    HIP(hipMemcpyAsync(ws + o_flag, &invalid, sizeof(invalid), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cluster, dim3((unsigned)blocks.size()), dim3(BLOCK), 0, s,
                       reinterpret_cast<const ClusterBlock *>(ws + o_blocks), reinterpret_cast<const double2 *>(ws + o_xy),
                       reinterpret_cast<const uint8_t *>(ws + o_ce), 2.5 * sigma, sigma * 1.41421356237309504880,
                       reinterpret_cast<double *>(ws + o_area), reinterpret_cast<unsigned int *>(ws + o_flag));
    HIP(hipGetLastError());
    HIP(hipMemcpyAsync(out.data(), ws + o_area, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP(hipMemcpyAsync(&invalid, ws + o_flag, sizeof(invalid), hipMemcpyDeviceToHost, s));
    HIP(hipStreamSynchronize(s));
  } catch (...) {
    if (pool)
      pool->release(ws, cap);
    else
      (void)hipFree(ws);
    throw;
  }
  if (pool)
    pool->release(ws, cap);
  else
    (void)hipFree(ws);
  if (invalid) return 0;
  for (size_t a = 0; a < n; ++a) areas[order[a]] = out[a];
  return 1;
}

}  // namespace texo

extern "C" int texo_gp_clustering_areas(int32_t device, uint64_t n_registers, const uint32_t *register_cell,
                                        const double *register_xy, uint64_t n_controls, const uint32_t *control_cell,
                                        const uint64_t *control_clock_lsr, const uint64_t *control_ce,
                                        uint64_t movable_instance_count, double *areas) {
  if ((n_registers && (!register_cell || !register_xy || !areas)) ||
      (n_controls && (!control_cell || !control_clock_lsr || !control_ce)))
    return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    std::vector<std::pair<uint32_t, std::pair<double, double>>> regs(n_registers);
    for (uint64_t i = 0; i < n_registers; ++i)
      regs[i] = {register_cell[i], {register_xy[2 * i], register_xy[2 * i + 1]}};
    std::map<uint32_t, texo::Control> controls;
    for (uint64_t i = 0; i < n_controls; ++i)
      controls[control_cell[i]] = {control_clock_lsr[2 * i], control_clock_lsr[2 * i + 1], control_ce[i]};
    HIP(hipSetDevice(device));
    hipStream_t s = nullptr;
    HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<double> out;
    int ok = 0;
    try {
      ok = texo::register_clustering_device(regs, controls, movable_instance_count, out, nullptr, s);
    } catch (...) {
      (void)hipStreamDestroy(s);
      throw;
    }
    (void)hipStreamDestroy(s);
    if (ok < 0) {
      // more clock-enable classes in a group than the kernel holds
      if (!texo::register_clustering(regs, controls, movable_instance_count, out)) return TEXO_GP_ERR_INVALID_NORMALIZATION;
      ok = 1;
    }
    if (!ok) return TEXO_GP_ERR_INVALID_NORMALIZATION;
    std::copy(out.begin(), out.end(), areas);
    return 0;
  } catch (...) {
    return texo::current_failure().status;
  }
}
