// gp_builders.hip — the placement-side input builders that feed the global
// placement boundary (texo-pnr routing_demand.rs and texo-flow's timing
// weights) and the provenance record of a placement run.
//
//  * texo_gp_routing_capacity_map: usable general-channel capacity per tile
//    (routing_capacity_map, crates/texo-pnr/src/routing_demand.rs:201-239):
//    one pass marks the wires incident to an unblocked pip, one pass adds each
//    usable H/V wire's capacity to its tile (64-bit sums, clamped to u32 at
//    the end: the reference's saturating adds of non-negative terms give the
//    same value).  Bit-exact.
//  * texo_gp_routing_demand_map: bbox-uniform directional demand of a legal
//    placement (routing_demand_map_with_capacity, routing_demand.rs:262-313):
//    one wavefront per net reduces the bounding box of its placed pins, then
//    strides over the box's tiles adding dx/area and dy/area as 2^-40 fixed
//    point (int64 atomics; run-to-run identical, within nets * 2^-41 of the
//    reference's FP64 sums in net order).
//  * texo_gp_timing_placement_weights: ecp5_timing_placement_weights
//    (crates/texo-flow/src/lib.rs:3185-3201) — 128-bit integer work per
//    timing edge on the host, bit-exact.
//  * texo_gp_placement_model: the reference's checkpoint placement-model
//    record (crates/texo-cli/src/checkpoint.rs:87-100) extended with this
//    library's global-placement engine, precision and options.

#include "gp_host.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace texo {
namespace {

constexpr uint32_t UNPLACED = 0xffffffffu;
constexpr int WAVES = BLOCK / 64;

__global__ void __launch_bounds__(BLOCK) k_builder_clear(uint64_t n, unsigned long long *p, uint64_t m,
                                                         uint32_t *q) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = 0ull;
  if (i < m) q[i] = 0u;
}

// ---- routing capacity -------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_capacity_pips(uint32_t n_pips, const uint32_t *__restrict__ from,
                                                         const uint32_t *__restrict__ to,
                                                         const uint8_t *__restrict__ blocked,
                                                         uint32_t *__restrict__ usable) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n_pips || blocked[i]) return;
  usable[from[i]] = 1u;  // every writer stores the same value
  usable[to[i]] = 1u;
}

__global__ void __launch_bounds__(BLOCK) k_capacity_wires(uint32_t n_wires, uint32_t width, uint64_t tiles,
                                                          const uint8_t *__restrict__ orientation,
                                                          const uint32_t *__restrict__ xy,
                                                          const uint16_t *__restrict__ capacity,
                                                          const uint32_t *__restrict__ usable,
                                                          unsigned long long *__restrict__ sums) {
  const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
  if (w >= n_wires || !usable[w] || orientation[w] == 0) return;
  const uint64_t tile = (uint64_t)xy[2 * w + 1] * width + xy[2 * w];
  atomicAdd(sums + (orientation[w] == 1 ? 0 : tiles) + tile, (unsigned long long)capacity[w]);
}

__global__ void __launch_bounds__(BLOCK) k_capacity_out(uint64_t n, const unsigned long long *__restrict__ sums,
                                                        uint32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = sums[i] > 0xffffffffull ? 0xffffffffu : (uint32_t)sums[i];
}

// ---- routing demand ---------------------------------------------------------
// One wavefront per net.  counts[0] included nets, counts[1] clock-driven.
__global__ void __launch_bounds__(BLOCK) k_demand_nets(uint32_t n_nets, uint32_t width, uint64_t tiles,
                                                       const uint32_t *__restrict__ net_start,
                                                       const uint32_t *__restrict__ pin_xy,
                                                       const uint8_t *__restrict__ net_clock,
                                                       long long *__restrict__ acc,
                                                       unsigned long long *__restrict__ counts) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t n = blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (n >= n_nets) return;  // whole wavefronts leave together
  if (net_clock[n]) {
    if (lane == 0) atomicAdd(counts + 1, 1ull);
    return;
  }
  const uint32_t b = net_start[n], e = net_start[n + 1];
  if (b == e || pin_xy[2 * b] == UNPLACED) return;  // no placed driver
  uint32_t x0 = UNPLACED, y0 = UNPLACED, x1 = 0, y1 = 0;
  for (uint32_t p = b + lane; p < e; p += 64) {
    const uint32_t x = pin_xy[2 * p], y = pin_xy[2 * p + 1];
    if (x == UNPLACED) continue;
    x0 = min(x0, x);
    y0 = min(y0, y);
    x1 = max(x1, x);
    y1 = max(y1, y);
  }
  x0 = lane_reduce<64>(x0, OpMinI());
  y0 = lane_reduce<64>(y0, OpMinI());
  x1 = lane_reduce<64>(x1, OpMaxI());
  y1 = lane_reduce<64>(y1, OpMaxI());
  if (lane == 0) atomicAdd(counts, 1ull);
  const uint32_t dx = x1 - x0, dy = y1 - y0;
  const double area = static_cast<double>(dx + 1) * static_cast<double>(dy + 1);
  const long long hq = __double2ll_rn(static_cast<double>(dx) / area * FX_SCALE);
  const long long vq = __double2ll_rn(static_cast<double>(dy) / area * FX_SCALE);
  const uint64_t cols = dx + 1, box = cols * (dy + 1);
  for (uint64_t t = lane; t < box; t += 64) {
    const uint64_t tile = (uint64_t)(y0 + t / cols) * width + (x0 + t % cols);
    if (hq) atomicAdd(reinterpret_cast<unsigned long long *>(acc + tile), static_cast<unsigned long long>(hq));
    if (vq)
      atomicAdd(reinterpret_cast<unsigned long long *>(acc + tiles + tile), static_cast<unsigned long long>(vq));
  }
}

__global__ void __launch_bounds__(BLOCK) k_demand_out(uint64_t n, const long long *__restrict__ acc,
                                                      double *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = static_cast<double>(acc[i]) * FX_INV;
}

unsigned grid_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + BLOCK - 1) / BLOCK); }

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

// Device layout of one demand-map call: inputs staged from the host, the
// fixed-point sums, the FP64 result and the two net counters.
struct DemandLayout {
  size_t start, pins, clock, acc, out, counts, total;
  DemandLayout(uint32_t n_nets, uint64_t n_pins, uint64_t tiles) {
    start = 0;
    pins = align256(start + 4 * ((size_t)n_nets + 1));
    clock = align256(pins + 8 * n_pins);
    acc = align256(clock + std::max<size_t>(n_nets, 1));
    out = align256(acc + 16 * tiles);
    counts = align256(out + 16 * tiles);
    total = align256(counts + 16);
  }
};

// Host-side checks of a demand-map call (the reference indexes the tile grid
// with every placed pin).
int validate_demand(uint32_t width, uint32_t height, uint32_t n_nets, const uint32_t *net_pin_start,
                    const uint32_t *pin_xy, const uint8_t *net_clock) {
  if (width == 0 || height == 0 || !net_pin_start || (!pin_xy && n_nets && net_pin_start[n_nets]) || !net_clock)
    return TEXO_GP_ERR_INVALID_ARGUMENT;
  for (uint32_t n = 0; n < n_nets; ++n)
    if (net_pin_start[n + 1] < net_pin_start[n]) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const uint64_t pins = net_pin_start[n_nets];
  for (uint64_t p = 0; p < pins; ++p) {
    const uint32_t x = pin_xy[2 * p], y = pin_xy[2 * p + 1];
    if (x == UNPLACED) continue;
    if (x >= width || y >= height) return TEXO_GP_ERR_POINT_OUTSIDE_GRID;
  }
  return 0;
}

// Enqueues one demand-map computation in the workspace `ws` (laid out by
// DemandLayout) on `s`: input copies, clear, nets, conversion, result copies.
void enqueue_demand(char *ws, const DemandLayout &L, uint32_t width, uint32_t height, uint32_t n_nets,
                    const uint32_t *net_pin_start, const uint32_t *pin_xy, const uint8_t *net_clock, double *horizontal,
                    double *vertical, unsigned long long *counts_host, hipStream_t s) {
  const uint64_t tiles = (uint64_t)width * height, pins = net_pin_start[n_nets];
  uint32_t *d_start = reinterpret_cast<uint32_t *>(ws + L.start);
  uint32_t *d_pins = reinterpret_cast<uint32_t *>(ws + L.pins);
  uint8_t *d_clock = reinterpret_cast<uint8_t *>(ws + L.clock);
  long long *d_acc = reinterpret_cast<long long *>(ws + L.acc);
  double *d_out = reinterpret_cast<double *>(ws + L.out);
  unsigned long long *d_counts = reinterpret_cast<unsigned long long *>(ws + L.counts);
  HIP(hipMemcpyAsync(d_start, net_pin_start, 4 * ((size_t)n_nets + 1), hipMemcpyHostToDevice, s));
  if (pins) HIP(hipMemcpyAsync(d_pins, pin_xy, 8 * pins, hipMemcpyHostToDevice, s));
  if (n_nets) HIP(hipMemcpyAsync(d_clock, net_clock, n_nets, hipMemcpyHostToDevice, s));
  // the sums and counters start at zero on every call (a fill kernel, not a
  // memset, so the sequence can sit in a captured graph)
  hipLaunchKernelGGL(k_builder_clear, dim3(grid_for(2 * tiles)), dim3(BLOCK), 0, s, 2 * tiles,
                     reinterpret_cast<unsigned long long *>(d_acc), (uint64_t)0, (uint32_t *)nullptr);
  hipLaunchKernelGGL(k_builder_clear, dim3(1), dim3(BLOCK), 0, s, (uint64_t)2, d_counts, (uint64_t)0,
                     (uint32_t *)nullptr);
  if (n_nets)
    hipLaunchKernelGGL(k_demand_nets, dim3((n_nets + WAVES - 1) / WAVES), dim3(BLOCK), 0, s, n_nets, width, tiles,
                       d_start, d_pins, d_clock, d_acc, d_counts);
  hipLaunchKernelGGL(k_demand_out, dim3(grid_for(2 * tiles)), dim3(BLOCK), 0, s, 2 * tiles, d_acc, d_out);
  HIP(hipGetLastError());
  HIP(hipMemcpyAsync(horizontal, d_out, 8 * tiles, hipMemcpyDeviceToHost, s));
  HIP(hipMemcpyAsync(vertical, d_out + tiles, 8 * tiles, hipMemcpyDeviceToHost, s));
  HIP(hipMemcpyAsync(counts_host, d_counts, 16, hipMemcpyDeviceToHost, s));
}

// Scratch of the allocation-free demand map (texo_gp_routing_demand_map_noalloc):
// staged inputs, fixed-point sums, the FP64 result and the net counters of
// one call, laid out by DemandLayout.  Every GPU the library runs on holds
// its own copy for the life of the process.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
// The device address of `scratch` on each device, looked up once per device
// with that device current (the host symbol is only a stand-in); one lock
// per device serialises this process's calls on it.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu[MAX_DEVICES];
char *g_scratch_ptr[MAX_DEVICES] = {};

char *scratch_on(int32_t device) {
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<char *>(ptr);
  }
  return g_scratch_ptr[device];
}
}  // namespace

}  // namespace texo

using texo::Failure;

namespace {

// A call-scoped stream on `device` (default priority).
struct CallStream {
  hipStream_t s = nullptr;
  explicit CallStream(int32_t device) {
    HIP(hipSetDevice(device));
    HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~CallStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

using u128 = unsigned __int128;

// scaled_binary_fraction (texo-flow lib.rs:3413-3433): floor(n / d * 2^bits)
// by restoring division, saturating at 2^bits.
uint64_t binary_fraction(u128 n, u128 d, uint32_t bits) {
  if (n >= d) return 1ull << bits;
  u128 r = n;
  uint64_t q = 0;
  for (uint32_t i = 0; i < bits; ++i) {
    q <<= 1;
    const u128 c = d - r;  // compare 2r with d as r >= d - r
    if (r >= c) {
      r -= c;
      q |= 1;
    } else {
      r *= 2;
    }
  }
  return q;
}

// u128::checked_pow
bool pow_u128(u128 base, uint32_t e, u128 &out) {
  u128 r = 1;
  while (e--) {
    if (base && r > ~(u128)0 / base) return false;
    r *= base;
  }
  out = r;
  return true;
}

// scaled_fraction_power (texo-flow lib.rs:3395-3411): rounded fixed-point
// binary exponentiation within `scale`.
uint64_t fixed_power(uint64_t scaled, uint32_t e, uint64_t scale) {
  const u128 s = scale;
  u128 result = s, factor = scaled;
  while (e) {
    if (e & 1) result = (result * factor + s / 2) / s;
    e >>= 1;
    if (e) factor = (factor * factor + s / 2) / s;
  }
  return (uint64_t)result;
}

// ecp5_placement_criticality_weight (texo-flow lib.rs:3355-3390):
// 1 + round(10 c^k) on a 1/1024 grid, c = min(path, worst) / worst.
uint64_t criticality_weight(u128 path, u128 worst, uint32_t exponent) {
  constexpr uint64_t SCALE = 1u << 10, WEIGHT = 10;
  exponent = std::max(exponent, 1u);
  const uint64_t c = worst == 0 ? 0 : binary_fraction(path < worst ? path : worst, worst, 10);
  u128 num = 0, den = 0, extra = 0;
  if (pow_u128(c, exponent, num) && pow_u128(SCALE, exponent, den) && num <= ~(u128)0 / WEIGHT) {
    extra = (WEIGHT * num + den / 2) / den;
  } else {
    const uint64_t p = fixed_power(c, exponent, SCALE);
    uint64_t m = p > UINT64_MAX / WEIGHT ? UINT64_MAX : WEIGHT * p;
    m = m > UINT64_MAX - SCALE / 2 ? UINT64_MAX : m + SCALE / 2;
    extra = m / SCALE;
  }
  const uint64_t e = extra > (u128)UINT64_MAX ? WEIGHT : (uint64_t)extra;
  return 1 + std::min(e, WEIGHT);
}

const char *precision_name(uint32_t p) {
  switch (p) {
    case TEXO_GP_DCT_FP32:
      return "fp32";
    case TEXO_GP_DCT_BF16X3:
      return "bf16x3";
    case TEXO_GP_DCT_FP16_SCALED:
      return "fp16_scaled";
  }
  return nullptr;
}

}  // namespace

extern "C" {

int texo_gp_channel_orientation(const char *name) {
  if (!name) return 0;
  const char *local = std::strrchr(name, '/');
  local = local ? local + 1 : name;
  if (std::strlen(local) < 3 || !std::isdigit((unsigned char)local[1]) || !std::isdigit((unsigned char)local[2]))
    return 0;
  return local[0] == 'H' ? 1 : local[0] == 'V' ? 2 : 0;
}

int texo_gp_routing_capacity_map(int32_t device, uint32_t width, uint32_t height, uint32_t n_wires,
                                 const uint8_t *wire_orientation, const uint32_t *wire_xy,
                                 const uint16_t *wire_capacity, uint32_t n_pips, const uint32_t *pip_from,
                                 const uint32_t *pip_to, const uint8_t *pip_blocked, uint32_t *horizontal,
                                 uint32_t *vertical) {
  if (width == 0 || height == 0 || !horizontal || !vertical) return TEXO_GP_ERR_INVALID_ARGUMENT;
  if (n_wires && (!wire_orientation || !wire_xy || !wire_capacity)) return TEXO_GP_ERR_INVALID_ARGUMENT;
  if (n_pips && (!pip_from || !pip_to)) return TEXO_GP_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n_pips; ++i)
    if (pip_from[i] >= n_wires || pip_to[i] >= n_wires) return TEXO_GP_ERR_INVALID_ARGUMENT;
  for (uint32_t w = 0; w < n_wires; ++w)
    if (wire_orientation[w] > 2 || wire_xy[2 * w] >= width || wire_xy[2 * w + 1] >= height)
      return TEXO_GP_ERR_INVALID_ARGUMENT;
  const uint64_t tiles = (uint64_t)width * height;
  char *ws = nullptr;
  int rc = 0;
  try {
    CallStream cs(device);
    const size_t o_or = 0, o_xy = texo::align256(std::max<size_t>(n_wires, 1));
    const size_t o_cap = texo::align256(o_xy + 8 * (size_t)n_wires);
    const size_t o_from = texo::align256(o_cap + 2 * (size_t)n_wires);
    const size_t o_to = texo::align256(o_from + 4 * (size_t)n_pips);
    const size_t o_bl = texo::align256(o_to + 4 * (size_t)n_pips);
    const size_t o_use = texo::align256(o_bl + std::max<size_t>(n_pips, 1));
    const size_t o_sum = texo::align256(o_use + 4 * (size_t)std::max<uint32_t>(n_wires, 1));
    const size_t o_out = texo::align256(o_sum + 16 * tiles);
    const size_t total = texo::align256(o_out + 8 * tiles);
    HIP(hipMalloc(&ws, total));
    const hipStream_t s = cs.s;
    if (n_wires) {
      HIP(hipMemcpyAsync(ws + o_or, wire_orientation, n_wires, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(ws + o_xy, wire_xy, 8 * (size_t)n_wires, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(ws + o_cap, wire_capacity, 2 * (size_t)n_wires, hipMemcpyHostToDevice, s));
    }
    std::vector<uint8_t> blocked;
    if (n_pips) {
      blocked.assign(n_pips, 0);
      if (pip_blocked) std::copy(pip_blocked, pip_blocked + n_pips, blocked.begin());
      HIP(hipMemcpyAsync(ws + o_from, pip_from, 4 * (size_t)n_pips, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(ws + o_to, pip_to, 4 * (size_t)n_pips, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(ws + o_bl, blocked.data(), n_pips, hipMemcpyHostToDevice, s));
    }
    auto *sums = reinterpret_cast<unsigned long long *>(ws + o_sum);
    auto *usable = reinterpret_cast<uint32_t *>(ws + o_use);
    auto *out = reinterpret_cast<uint32_t *>(ws + o_out);
    const uint64_t clear = std::max<uint64_t>(2 * tiles, n_wires);
    hipLaunchKernelGGL(texo::k_builder_clear, dim3(texo::grid_for(clear)), dim3(texo::BLOCK), 0, s, 2 * tiles, sums,
                       (uint64_t)n_wires, usable);
    if (n_pips)
      hipLaunchKernelGGL(texo::k_capacity_pips, dim3(texo::grid_for(n_pips)), dim3(texo::BLOCK), 0, s, n_pips,
                         reinterpret_cast<const uint32_t *>(ws + o_from), reinterpret_cast<const uint32_t *>(ws + o_to),
                         reinterpret_cast<const uint8_t *>(ws + o_bl), usable);
    if (n_wires)
      hipLaunchKernelGGL(texo::k_capacity_wires, dim3(texo::grid_for(n_wires)), dim3(texo::BLOCK), 0, s, n_wires,
                         width, tiles, reinterpret_cast<const uint8_t *>(ws + o_or),
                         reinterpret_cast<const uint32_t *>(ws + o_xy), reinterpret_cast<const uint16_t *>(ws + o_cap),
                         usable, sums);
    hipLaunchKernelGGL(texo::k_capacity_out, dim3(texo::grid_for(2 * tiles)), dim3(texo::BLOCK), 0, s, 2 * tiles,
                       sums, out);
    HIP(hipGetLastError());
    HIP(hipMemcpyAsync(horizontal, out, 4 * tiles, hipMemcpyDeviceToHost, s));
    HIP(hipMemcpyAsync(vertical, out + tiles, 4 * tiles, hipMemcpyDeviceToHost, s));
    HIP(hipStreamSynchronize(s));
  } catch (...) {
    rc = texo::current_failure().status;
  }
  if (ws) (void)hipFree(ws);
  return rc;
}

int texo_gp_routing_demand_map(int32_t device, uint32_t width, uint32_t height, uint32_t n_nets,
                               const uint32_t *net_pin_start, const uint32_t *pin_xy, const uint8_t *net_clock,
                               double *horizontal, double *vertical, uint64_t *included_nets,
                               uint64_t *excluded_clock_nets) {
  if (!horizontal || !vertical) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const int bad = texo::validate_demand(width, height, n_nets, net_pin_start, pin_xy, net_clock);
  if (bad) return bad;
  const uint64_t tiles = (uint64_t)width * height;
  const texo::DemandLayout L(n_nets, net_pin_start[n_nets], tiles);
  char *ws = nullptr;
  int rc = 0;
  try {
    CallStream cs(device);
    HIP(hipMalloc(&ws, L.total));
    unsigned long long counts[2] = {0, 0};
    texo::enqueue_demand(ws, L, width, height, n_nets, net_pin_start, pin_xy, net_clock, horizontal, vertical, counts,
                         cs.s);
    HIP(hipStreamSynchronize(cs.s));
    if (included_nets) *included_nets = counts[0];
    if (excluded_clock_nets) *excluded_clock_nets = counts[1];
  } catch (...) {
    rc = texo::current_failure().status;
  }
  if (ws) (void)hipFree(ws);
  return rc;
}

int texo_gp_routing_demand_map_noalloc(int32_t device, uint32_t width, uint32_t height, uint32_t n_nets,
                                       const uint32_t *net_pin_start, const uint32_t *pin_xy,
                                       const uint8_t *net_clock, double *horizontal, double *vertical,
                                       uint64_t *included_nets, uint64_t *excluded_clock_nets) {
  if (!horizontal || !vertical || device < 0 || device >= texo::MAX_DEVICES) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const int bad = texo::validate_demand(width, height, n_nets, net_pin_start, pin_xy, net_clock);
  if (bad) return bad;
  const uint64_t tiles = (uint64_t)width * height;
  const texo::DemandLayout L(n_nets, net_pin_start[n_nets], tiles);
  if (L.total > texo::NOALLOC_SCRATCH_BYTES) {
    texo::g_last_error = "routing demand map needs " + std::to_string(L.total) + " bytes of device scratch; the " +
                         "allocation-free entry point holds " + std::to_string(texo::NOALLOC_SCRATCH_BYTES);
    return TEXO_GP_ERR_OUT_OF_MEMORY;
  }
  std::lock_guard<std::mutex> lk(texo::g_scratch_mu[device]);
  try {
    HIP(hipSetDevice(device));
    char *ws = texo::scratch_on(device);
    // the null stream: no stream object is created
    unsigned long long counts[2] = {0, 0};
    texo::enqueue_demand(ws, L, width, height, n_nets, net_pin_start, pin_xy, net_clock, horizontal, vertical, counts,
                         nullptr);
    HIP(hipStreamSynchronize(nullptr));
    if (included_nets) *included_nets = counts[0];
    if (excluded_clock_nets) *excluded_clock_nets = counts[1];
    return 0;
  } catch (...) {
    return texo::current_failure().status;
  }
}

uint64_t texo_gp_timing_placement_weights(uint64_t n_edges, const uint32_t *net, const uint32_t *sink,
                                          const uint64_t *path_delay_ps, const uint64_t *worst_path_delay_ps,
                                          uint32_t exponent, uint32_t *out_net, uint32_t *out_sink,
                                          uint64_t *out_weight) {
  if (n_edges && (!net || !sink || !path_delay_ps || !worst_path_delay_ps || !out_net || !out_sink || !out_weight))
    return 0;
  // (net, sink) -> maximum weight, iterated in key order like the
  // reference's BTreeMap
  std::map<std::pair<uint32_t, uint32_t>, uint64_t> weights;
  for (uint64_t i = 0; i < n_edges; ++i) {
    const u128 path = ((u128)path_delay_ps[2 * i + 1] << 64) | path_delay_ps[2 * i];
    const u128 worst = ((u128)worst_path_delay_ps[2 * i + 1] << 64) | worst_path_delay_ps[2 * i];
    const uint64_t w = criticality_weight(path, worst, exponent);
    auto ins = weights.emplace(std::make_pair(net[i], sink[i]), w);
    if (!ins.second) ins.first->second = std::max(ins.first->second, w);
  }
  uint64_t k = 0;
  for (const auto &kv : weights) {
    out_net[k] = kv.first.first;
    out_sink[k] = kv.first.second;
    out_weight[k] = kv.second;
    ++k;
  }
  return k;
}

int texo_gp_placement_model(uint32_t algorithm, uint32_t weight_exponent, const texo_gp_options *options, char *out,
                            size_t capacity) {
  // checkpoint_placement_model (crates/texo-cli/src/checkpoint.rs:87-100)
  // and Ecp5InitialPlacementAlgorithm::checkpoint_name
  // (crates/texo-flow/src/lib.rs:316-324), then the engine record
  const char *name = nullptr;
  bool timing = false, routability = false;
  switch (algorithm) {
    case TEXO_GP_ALGORITHM_IMPORTED:
      name = "imported_v1";
      break;
    case TEXO_GP_ALGORITHM_ROUTABILITY_ELECTROSTATIC:
      name = "routability_electrostatic_v1";
      routability = true;
      break;
    case TEXO_GP_ALGORITHM_TIMING_ROUTABILITY_ELECTROSTATIC:
      name = "ecp5_timing_routability_electrostatic_v1";
      routability = timing = true;
      break;
    default:
      return TEXO_GP_ERR_INVALID_ARGUMENT;
  }
  texo_gp_options o{};
  if (options) o = *options;
  const char *prec = precision_name(o.dct_precision);
  if (!prec) return TEXO_GP_ERR_INVALID_ARGUMENT;
  std::string s = "{\"initial_algorithm\":\"";
  s += name;
  s += "\",\"timing_weight_model\":";
  s += timing ? "\"ecp5_1_plus_10_criticality_power_v1\"" : "null";
  s += ",\"criticality_exponent\":";
  s += timing ? std::to_string(weight_exponent) : "null";
  s += ",\"routability_model\":";
  s += routability ? "\"directional_rudy_area_adjustment_v1\"" : "null";
  s += ",\"initial_predicted_detail\":false";
  if (algorithm != TEXO_GP_ALGORITHM_IMPORTED) {
    char gamma[64];
    std::snprintf(gamma, sizeof gamma, "%.17g", o.gamma_base > 0.0 ? o.gamma_base : 8.0);
    s += ",\"global_placement\":{\"engine\":\"texo_amd_mi355x_electrostatic_v1\",\"abi_version\":";
    s += std::to_string(TEXO_GP_ABI_VERSION);
    s += ",\"dct_precision\":\"";
    s += prec;
    s += "\",\"density_accumulation\":\"int64_fixed_point_2^-40\",\"wirelength_arithmetic\":\"f64\"";
    s += ",\"fixed_iterations\":";
    s += o.fixed_iterations ? std::to_string(o.fixed_iterations) : "null";
    s += ",\"max_iterations\":";
    s += std::to_string(o.max_iterations ? o.max_iterations : 3000u);
    s += ",\"gamma_base\":";
    s += gamma;
    s += "}";
  }
  s += "}";
  if (!out || capacity < s.size() + 1) return -(int)(s.size() + 1);
  std::memcpy(out, s.c_str(), s.size() + 1);
  return 0;
}

}  // extern "C"
