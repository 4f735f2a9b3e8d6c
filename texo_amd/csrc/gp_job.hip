// gp_job.hip — one placement job on the host: problem set-up
// (eplace.rs:255-336, electrostatic_placement.rs:139-240,
// analytical_placement.rs:116-231), its device arena, and the schedule
// arithmetic of optimize_density_once (eplace.rs:1111-1415) and the area
// rounds (instance_area.rs, register_clustering.rs).

#include "gp_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <set>
#include <type_traits>

namespace texo {

thread_local std::string g_last_error;
thread_local KernelTimer *g_timer = nullptr;

namespace {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
constexpr double MULTIPLIER_ALPHA_LOW = 1.05;   // eplace.rs:128
constexpr double MULTIPLIER_ALPHA_HIGH = 1.06;  // eplace.rs:129
constexpr double AREA_ADJUSTMENT_OVERFLOW_TARGET = 0.15;  // eplace.rs:130
constexpr uint32_t LARGE_NET_PINS = 16;  // nets above this use one wavefront
}  // namespace

[[noreturn]] void fail(int status, int64_t index, int64_t index2) {
  Failure f;
  f.status = status;
  f.index = index;
  f.index2 = index2;
  throw f;
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    fail(e == hipErrorOutOfMemory ? TEXO_GP_ERR_OUT_OF_MEMORY : TEXO_GP_ERR_DEVICE);
  }
}

bool is_lut(uint32_t k) { return k >= 0x100u && k < 0x200u; }
static bool exchangeable(uint32_t k) { return is_lut(k) || k == TEXO_KIND_REGISTER || k == TEXO_KIND_MEMORY; }

// eplace.rs:1300-1308
double overflow_target(uint32_t kind) {
  if (kind == TEXO_KIND_MEMORY || kind == TEXO_KIND_DSP) return 0.20;
  if (is_lut(kind) || kind == TEXO_KIND_REGISTER) return 0.10;
  return 0.0;
}

// crates/texo-pnr/src/lib.rs:4566-4568 (f64::round: half away from zero)
uint32_t rounded_coordinate(double v, uint32_t extent) {
  double r = std::round(v);
  const double hi = static_cast<double>(extent - 1);
  if (!(r >= 0.0)) r = 0.0;
  if (r > hi) r = hi;
  return static_cast<uint32_t>(r);
}

double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

KernelTimer::~KernelTimer() {
  for (auto &e : events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
}

void timer_begin(KernelTimer *t, const char *name, hipStream_t s) {
  if (!t || !t->enabled || t->name != name) return;
  if (t->used == t->events.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    t->events.push_back({a, b});
  }
  (void)hipEventRecord(t->events[t->used].first, s);
}

void timer_end(KernelTimer *t, const char *name, hipStream_t s, double tasks) {
  if (!t || !t->enabled || t->name != name || t->used >= t->events.size()) return;
  (void)hipEventRecord(t->events[t->used].second, s);
  t->used += 1;
  t->tasks += tasks;
}

BlockPool::~BlockPool() {
  for (auto &b : free_) free_block(b.second);
}

void BlockPool::free_block(void *p) {
  if (host_)
    (void)hipHostFree(p);
  else
    (void)hipFree(p);
}

// Blocks come in four size classes per power of two (at most 25% unused),
// and a request takes the smallest free block of at least its class that is
// no more than twice as large.
void *BlockPool::acquire(size_t bytes, size_t *capacity) {
  size_t cls = std::max<size_t>(bytes, 1u << 20);
  const int top = 63 - __builtin_clzll(cls);
  const size_t step = size_t(1) << (top - 2);
  cls = (cls + step - 1) & ~(step - 1);
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.lower_bound(cls);
    if (it != free_.end() && it->first <= 2 * cls) {
      void *p = it->second;
      *capacity = it->first;
      kept_ -= it->first;
      free_.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  if (host_)
    HIP(hipHostMalloc(&p, cls, hipHostMallocDefault));
  else
    HIP(hipMalloc(&p, cls));
  *capacity = cls;
  return p;
}

void BlockPool::release(void *p, size_t capacity) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (kept_ + capacity <= keep_) {
      free_.emplace(capacity, p);
      kept_ += capacity;
      return;
    }
  }
  free_block(p);
}

void Arena::release() {
  if (!base_) return;
  if (pool_)
    pool_->release(base_, capacity_);
  else
    (void)hipFree(base_);
  base_ = nullptr;
  pool_.reset();
}

void Arena::commit(std::shared_ptr<BlockPool> pool) {
  if (pool) {
    base_ = pool->acquire(planned_, &capacity_);
    pool_ = std::move(pool);
  } else {
    HIP(hipMalloc(&base_, planned_));
    capacity_ = planned_;
  }
  for (auto &e : plan_) *e.slot = static_cast<char *>(base_) + e.offset;
}

// Poisson2d::new (electrostatic_placement.rs:675-714) and the orthonormal DCT
// bases (:616-672) as dense matrices: FP32, and BF16 hi/lo splits transposed.
void Basis::build(int w, int h) {
  W = w;
  H = h;
  const size_t dims[4] = {(size_t)W, (size_t)H, (size_t)H, (size_t)W};
  for (int j = 0; j < 4; ++j) {
    arena.reserve(b32[j], dims[j] * dims[j]);
    arena.reserve(bthi[j], dims[j] * dims[j]);
    arena.reserve(btlo[j], dims[j] * dims[j]);
    arena.reserve(fthi[j], dims[j] * dims[j]);
    arena.reserve(ftlo[j], dims[j] * dims[j]);
  }
  arena.reserve(inv_lambda, (size_t)W * H);
  arena.commit();
  auto basis = [](size_t n) {
    std::vector<double> b(n * n);
    for (size_t k = 0; k < n; ++k) {
      const double s = k == 0 ? std::sqrt(1.0 / (double)n) : std::sqrt(2.0 / (double)n);
      for (size_t x = 0; x < n; ++x) b[k * n + x] = s * std::cos(kPi * (double)k * ((double)x + 0.5) / (double)n);
    }
    return b;
  };
  const std::vector<double> cw = basis(W), ch = basis(H);
  auto bf16 = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  };
  auto bf16f = [](uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  // operand j: 0 = C_W^T, 1 = C_H^T, 2 = C_H, 3 = C_W  (row-major K x N)
  for (int j = 0; j < 4; ++j) {
    const size_t n = dims[j];
    const std::vector<double> &c = (j == 0 || j == 3) ? cw : ch;
    const bool transpose = j < 2;
    std::vector<float> f32(n * n);
    std::vector<uint16_t> thi(n * n), tlo(n * n), fhi(n * n), flo(n * n);
    // FP16 splits of basis * 2^s with max |basis| * 2^s in [2^14, 2^15)
    double bmax = 0.0;
    for (size_t i = 0; i < n * n; ++i) bmax = std::max(bmax, std::fabs(c[i]));
    int e = 0;
    (void)std::frexp(bmax, &e);
    const float scale = std::ldexp(1.0f, 15 - e);
    f16_scale[j] = scale;
    auto f16_bits = [](float f) {
      const _Float16 h = static_cast<_Float16>(f);
      uint16_t u;
      std::memcpy(&u, &h, 2);
      return u;
    };
    for (size_t r = 0; r < n; ++r)
      for (size_t k = 0; k < n; ++k) {
        const double v = transpose ? c[k * n + r] : c[r * n + k];  // element (K index r, N index k)
        const float fv = static_cast<float>(v);
        f32[r * n + k] = fv;
        const uint16_t hi = bf16(fv);
        const uint16_t lo = bf16(static_cast<float>(v - (double)bf16f(hi)));
        thi[k * n + r] = hi;  // transposed: (N index k, K index r)
        tlo[k * n + r] = lo;
        const double sv = v * static_cast<double>(scale);
        const float h = static_cast<float>(static_cast<_Float16>(static_cast<float>(sv)));
        fhi[k * n + r] = f16_bits(h);
        flo[k * n + r] = f16_bits(static_cast<float>(sv - static_cast<double>(h)));
      }
    HIP(hipMemcpy(b32[j], f32.data(), n * n * sizeof(float), hipMemcpyHostToDevice));
    HIP(hipMemcpy(bthi[j], thi.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP(hipMemcpy(btlo[j], tlo.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP(hipMemcpy(fthi[j], fhi.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP(hipMemcpy(ftlo[j], flo.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  std::vector<float> il((size_t)W * H);
  for (int ky = 0; ky < H; ++ky) {
    const double sy = std::sin(kPi * (double)ky / (2.0 * (double)H));
    for (int kx = 0; kx < W; ++kx) {
      const double sx = std::sin(kPi * (double)kx / (2.0 * (double)W));
      const double lam = 4.0 * sx * sx + 4.0 * sy * sy;
      il[(size_t)ky * W + kx] = (kx == 0 && ky == 0) ? 0.0f : static_cast<float>(1.0 / lam);
    }
  }
  HIP(hipMemcpy(inv_lambda, il.data(), il.size() * sizeof(float), hipMemcpyHostToDevice));
}

void Job::load(const texo_gp_problem *p, const texo_gp_options *o, const std::shared_ptr<BlockPool> &pool) {
  if (!p || p->abi_version != TEXO_GP_ABI_VERSION) fail(TEXO_GP_ERR_INVALID_ARGUMENT);
  if (o) {
    opt.precision = o->dct_precision;
    opt.fixed_iterations = o->fixed_iterations;
    if (o->max_iterations) opt.max_iterations = o->max_iterations;
    if (o->gamma_base > 0.0) opt.gamma_base = o->gamma_base;
    opt.trace = o->trace != 0;
    opt.device = o->device;
    opt.inject = o->inject_failure_at_evaluation;
  }
  if (opt.precision > TEXO_GP_DCT_FP16_SCALED) fail(TEXO_GP_ERR_INVALID_ARGUMENT);
  W = p->width;
  H = p->height;
  if (W == 0 || H == 0) fail(TEXO_GP_ERR_EMPTY_GRID);
  if ((uint64_t)W * H > (1ull << 26)) fail(TEXO_GP_ERR_GRID_TOO_LARGE);
  G = (size_t)W * H;
  U = p->n_units;
  n_cells = p->n_cells;
  best_effort = p->best_effort != 0;
  unit_fixed.assign(p->unit_fixed, p->unit_fixed + U);
  initial_xy.assign(p->unit_initial_xy, p->unit_initial_xy + 2 * U);
  for (size_t c = 0; c < n_cells; ++c)
    if (p->cell_unit[c] >= U) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)c);

  // Density units: movable units with exchangeable members (eplace.rs:261-301)
  std::vector<uint8_t> kind_seen(1u << 16, 0);
  unit_du.assign(U, -1);
  {
    size_t members = 0;
    for (size_t u = 0; u < U; ++u)
      if (!unit_fixed[u]) members += p->unit_member_start[u + 1] - p->unit_member_start[u];
    m_unit.reserve(members);
    m_cell.reserve(members);
    m_off.reserve(2 * members);
  }
  for (size_t u = 0; u < U; ++u) {
    const uint32_t b = p->unit_member_start[u], e = p->unit_member_start[u + 1];
    if (unit_fixed[u] || b == e) continue;
    unit_du[u] = static_cast<int32_t>(du_unit.size());
    du_unit.push_back(static_cast<uint32_t>(u));
    du_mstart.push_back(static_cast<uint32_t>(m_unit.size()));
    for (uint32_t m = b; m < e; ++m) {
      if (!exchangeable(p->member_kind[m]) || p->member_cell[m] >= n_cells) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)m);
      kind_seen[p->member_kind[m]] = 1;
      m_unit.push_back(static_cast<uint32_t>(u));
      m_cell.push_back(p->member_cell[m]);
      m_off.push_back(p->member_offset_xy[2 * m]);
      m_off.push_back(p->member_offset_xy[2 * m + 1]);
    }
  }
  du_mstart.push_back(static_cast<uint32_t>(m_unit.size()));
  DU = du_unit.size();
  // members by unit (an empty range for a unit that is no density unit), so
  // the gradient finds a unit's members with one indirection
  u_mstart.assign(U + 1, 0);
  for (size_t u = 0, m = 0; u < U; ++u) {
    u_mstart[u] = static_cast<uint32_t>(m);
    if (unit_du[u] >= 0) m = du_mstart[unit_du[u] + 1];
  }
  u_mstart[U] = static_cast<uint32_t>(m_unit.size());
  M = m_unit.size();
  m_charge.assign(M, 1.0);
  // fields in ascending kind order (the reference's BTreeMap over ResourceKind)
  for (uint32_t k = 0; k < kind_seen.size(); ++k)
    if (kind_seen[k]) {
      kind_seen[k] = static_cast<uint8_t>(field_kind.size());
      field_kind.push_back(k);
    }
  F = static_cast<int>(field_kind.size());
  if (F > MAXF) fail(TEXO_GP_ERR_UNSUPPORTED);
  m_field.resize(M);
  {
    size_t k = 0;
    for (size_t u = 0; u < U; ++u) {
      const uint32_t b = p->unit_member_start[u], e = p->unit_member_start[u + 1];
      if (unit_fixed[u] || b == e) continue;
      for (uint32_t m = b; m < e; ++m) m_field[k++] = kind_seen[p->member_kind[m]];
    }
  }
  for (size_t c = 0; c < n_cells; ++c)
    if (!unit_fixed[p->cell_unit[c]]) movable_instance_count += 1;

  // Hypergraph (analytical_placement.rs:116-168): nets then timed arcs.
  n_nets = p->n_nets;
  n_arcs = p->n_arcs;
  n_net_total = n_nets + n_arcs;
  net_start.assign(n_net_total + 1, 0);
  const size_t P = n_nets ? p->net_pin_start[n_nets] : 0;
  for (size_t n = 0; n < n_nets; ++n) {
    if (p->net_pin_start[n + 1] < p->net_pin_start[n] + 2) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)n);
    net_start[n + 1] = p->net_pin_start[n + 1];
  }
  first_arc_pin = P;
  n_pins_total = P + 2 * n_arcs;
  for (size_t a = 0; a < n_arcs; ++a) net_start[n_nets + a + 1] = static_cast<uint32_t>(P + 2 * (a + 1));
  std::vector<uint32_t> pin_cell(n_pins_total);
  for (size_t i = 0; i < P; ++i) pin_cell[i] = p->pin_cell[i];
  net_weight.assign(n_net_total, 1.0);
  for (size_t a = 0; a < n_arcs; ++a) {
    if (p->arc_extra_weight[a] == 0) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)a);
    pin_cell[P + 2 * a] = p->arc_cells[2 * a];
    pin_cell[P + 2 * a + 1] = p->arc_cells[2 * a + 1];
    const uint64_t w = p->arc_extra_weight[a];
    net_weight[n_nets + a] = (double)(uint32_t)(w >> 32) * 4294967296.0 + (double)(uint32_t)(w & 0xffffffffu);
  }
  for (size_t i = 0; i < n_pins_total; ++i)
    if (pin_cell[i] >= n_cells) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)i);
  // wirelength_preconditioner (analytical_placement.rs:177-211) and
  // baseline_cell_incidence_weights (:219-231)
  wl_diag.assign(U, 0.0);
  std::vector<double> cell_weight(n_cells, 0.0);
  {
    // distinct units / cells of a net by stamping them with the net's number:
    // each distinct one receives the net's weight once, nets in input order
    std::vector<uint64_t> seen_u(U, UINT64_MAX), seen_c(n_cells, UINT64_MAX);
    for (size_t n = 0; n < n_nets; ++n) {
      const uint64_t first = 2 * (uint64_t)n, second = first + 1;
      const double w = 1.0 / (double)(uint32_t)(net_start[n + 1] - net_start[n] - 1);
      uint32_t distinct = 0;
      for (uint32_t q = net_start[n]; q < net_start[n + 1]; ++q) {
        const uint32_t c = pin_cell[q], u = p->cell_unit[c];
        if (seen_u[u] != first) {
          seen_u[u] = first;
          distinct += 1;
        }
        if (seen_c[c] != first) {
          seen_c[c] = first;
          cell_weight[c] += w;
        }
      }
      if (distinct >= 2)
        for (uint32_t q = net_start[n]; q < net_start[n + 1]; ++q) {
          const uint32_t u = p->cell_unit[pin_cell[q]];
          if (seen_u[u] == first) {
            seen_u[u] = second;
            wl_diag[u] += w;
          }
        }
    }
    for (size_t a = 0; a < n_arcs; ++a) {
      const uint32_t du_ = p->cell_unit[p->arc_cells[2 * a]], su = p->cell_unit[p->arc_cells[2 * a + 1]];
      if (du_ == su) continue;
      wl_diag[du_] += net_weight[n_nets + a];
      wl_diag[su] += net_weight[n_nets + a];
    }
  }
  // Device layout of the nets: the nets with a movable pin ("live"), then
  // those whose pins are all fixed; in each class small nets (at most
  // LARGE_NET_PINS pins) grouped by degree, then large nets longest first;
  // then the timed arcs, live ones first.  Lanes of a wavefront take
  // consecutive small nets, so a wave's pin loops have one length and its
  // pins are contiguous.  Pins are renumbered with their nets; each unit
  // still lists its pins in the input net order, which fixes the summation
  // order of its gradient.  A net of fixed pins has no gradient and a WA
  // value that depends on gamma only: its wavefronts form a separate range
  // that runs once per gamma (an incremental request fixes most units).
  {
    auto degree = [&](uint32_t n) { return net_start[n + 1] - net_start[n]; };
    // a pin of a fixed cell is a fixed point (its cell's coordinate); a net
    // with a pin of a movable cell is live.  unit -> number of incident
    // movable pins (nets before arcs, input order)
    std::vector<uint8_t> live(n_net_total, 0);
    u_pstart.assign(U + 1, 0);
    for (size_t n = 0; n < n_net_total; ++n)
      for (uint32_t i = net_start[n]; i < net_start[n + 1]; ++i) {
        const uint32_t c = pin_cell[i];
        if (p->cell_fixed[c]) continue;
        live[n] = 1;
        u_pstart[p->cell_unit[c] + 1] += 1;
      }
    for (size_t u = 0; u < U; ++u) u_pstart[u + 1] += u_pstart[u];
    // per class: stable counting sort of the small nets by degree, then the
    // large nets stably by descending degree
    std::vector<uint32_t> order;
    order.reserve(n_net_total);
    n_live_nets = 0;
    for (int cls = 1; cls >= 0; --cls) {
      std::vector<uint32_t> at(LARGE_NET_PINS + 2, 0), large;
      const size_t first = order.size();
      for (size_t n = 0; n < n_nets; ++n) {
        if (live[n] != cls) continue;
        const uint32_t d = degree((uint32_t)n);
        if (d <= LARGE_NET_PINS)
          at[d + 1] += 1;
        else
          large.push_back((uint32_t)n);
      }
      for (size_t d = 1; d < at.size(); ++d) at[d] += at[d - 1];
      order.resize(first + at.back() + large.size());
      for (size_t n = 0; n < n_nets; ++n) {
        if (live[n] != cls) continue;
        const uint32_t d = degree((uint32_t)n);
        if (d <= LARGE_NET_PINS) order[first + at[d]++] = (uint32_t)n;
      }
      std::stable_sort(large.begin(), large.end(), [&](uint32_t a, uint32_t b) { return degree(a) > degree(b); });
      std::copy(large.begin(), large.end(), order.end() - large.size());
    }
    for (int cls = 1; cls >= 0; --cls)
      for (size_t n = n_nets; n < n_net_total; ++n)
        if (live[n] == cls) order.push_back((uint32_t)n);
    // nets and pins go straight to their device positions in one pass over
    // the input order: each pin's unit (NO_UNIT for a fixed pin) and offset
    // (or fixed coordinate), and each unit's list of its movable pins in
    // input order (nets first, then timed arcs), which fixes the summation
    // order of its gradient; each movable pin's derivative goes to its entry
    // of that list (pin_slot), so the per-unit sums read contiguously
    std::vector<uint32_t> start2(n_net_total + 1, 0), unit2(n_pins_total);
    std::vector<double> xy2(2 * n_pins_total), w2(n_net_total);
    std::vector<uint8_t> live2(n_net_total);
    u_pins.assign(u_pstart[U], 0);
    pin_slot.assign(n_pins_total, NO_UNIT);
    u_nbase.assign(U, 0);
    {
      std::vector<uint32_t> rank(n_net_total);
      for (size_t k = 0; k < n_net_total; ++k) rank[order[k]] = (uint32_t)k;
      for (size_t n = 0; n < n_net_total; ++n) {
        const uint32_t k = rank[n];
        start2[k + 1] = degree((uint32_t)n);
        w2[k] = net_weight[n];
        live2[k] = live[n];
      }
      for (size_t k = 0; k < n_net_total; ++k) start2[k + 1] += start2[k];
      std::vector<uint32_t> fill(u_pstart.begin(), u_pstart.end() - 1);
      for (size_t n = 0; n < n_net_total; ++n) {
        uint32_t q = start2[rank[n]];
        for (uint32_t i = net_start[n]; i < net_start[n + 1]; ++i, ++q) {
          const uint32_t c = pin_cell[i];
          if (p->cell_fixed[c]) {
            unit2[q] = NO_UNIT;
            xy2[2 * q] = p->cell_fixed_xy[2 * c];
            xy2[2 * q + 1] = p->cell_fixed_xy[2 * c + 1];
          } else {
            const uint32_t u = p->cell_unit[c];
            unit2[q] = u;
            xy2[2 * q] = p->cell_offset_xy[2 * c];
            xy2[2 * q + 1] = p->cell_offset_xy[2 * c + 1];
            const uint32_t k = fill[u]++;
            u_pins[k] = q;
            pin_slot[q] = k;
            if (i < first_arc_pin) u_nbase[u] += 1;
          }
        }
      }
    }
    // WA segments: runs of consecutive nets of one class with one lane-group
    // size and, for small nets and timed arcs, one degree (a lane then
    // computes its pins from the segment's first pin); per class the
    // looping wavefronts (longest nets) come first, and the fixed class
    // starts on a block boundary
    std::vector<WaSeg> runs[2];
    for (size_t k = 0; k < n_net_total; ++k) {
      const uint32_t d = start2[k + 1] - start2[k];
      const int g = wa_lanes(d);
      const int deg = g > 0 && (d <= LARGE_NET_PINS || k >= n_nets) ? (int)d : 0;
      std::vector<WaSeg> &rc = runs[live2[k]];
      if (!rc.empty() && rc.back().lanes == g && rc.back().degree == deg &&
          (size_t)(rc.back().first_net + rc.back().nets) == k)
        rc.back().nets += 1;
      else
        rc.push_back(WaSeg{(int)k, 1, g, 0, (int)start2[k], deg, 0, 0});
    }
    wa_seg.clear();
    int waves = 0;
    for (int cls = 1; cls >= 0; --cls) {
      std::stable_partition(runs[cls].begin(), runs[cls].end(), [](const WaSeg &r) { return r.lanes == 0; });
      for (WaSeg &r : runs[cls]) {
        r.first_wave = waves;
        waves += r.lanes == 0 ? r.nets : (int)(((int64_t)r.nets * r.lanes + 63) / 64);
        wa_seg.push_back(r);
      }
      if (cls == 1) {
        waves = (waves + BLOCK / 64 - 1) / (BLOCK / 64) * (BLOCK / 64);
        wa_waves_live = waves;
      }
    }
    if (wa_seg.size() > (size_t)WA_MAX_SEGMENTS) fail(TEXO_GP_ERR_UNSUPPORTED);
    wa_waves = waves;
    net_start.swap(start2);
    pin_unit.swap(unit2);
    pin_xy.swap(xy2);
    net_weight.swap(w2);
  }
  has_routing = p->routing_horizontal && p->routing_vertical;
  if (has_routing) {
    routing_h.assign(p->routing_horizontal, p->routing_horizontal + G);
    routing_v.assign(p->routing_vertical, p->routing_vertical + G);
  }
  for (size_t i = 0; i < p->n_register_controls; ++i)
    controls[p->register_cell[i]] = {p->register_clock_lsr[2 * i], p->register_clock_lsr[2 * i + 1],
                                     p->register_ce[i]};

  if (DU == 0) {  // eplace.rs:303-310: no density unit, targets are the rounded inputs
    empty = true;
    C = 2 * U;
    lower.resize(C);
    upper.resize(C);
    for (size_t u = 0; u < U; ++u) {
      lower[2 * u] = p->unit_bounds[4 * u];
      lower[2 * u + 1] = p->unit_bounds[4 * u + 1];
      upper[2 * u] = p->unit_bounds[4 * u + 2];
      upper[2 * u + 1] = p->unit_bounds[4 * u + 3];
    }
    return;
  }

  // DensityModel::new (electrostatic_placement.rs:139-196)
  for (size_t k = 0; k < p->n_capacity_kinds; ++k)
    capacity[p->capacity_kind[k]] = std::vector<uint32_t>(p->capacity + k * G, p->capacity + (k + 1) * G);
  for (size_t i = 0; i < p->n_fixed_occupancy; ++i) {
    const uint32_t kind = p->fixed_occupancy_kind[i], x = p->fixed_occupancy_xy[2 * i],
                   y = p->fixed_occupancy_xy[2 * i + 1];
    if (x >= W || y >= H) {
      Failure f;
      f.status = TEXO_GP_ERR_POINT_OUTSIDE_GRID;
      f.x = x;
      f.y = y;
      throw f;
    }
    auto it = capacity.find(kind);
    if (it == capacity.end() || it->second[(size_t)y * W + x] == 0) {
      Failure f;
      f.status = TEXO_GP_ERR_FIXED_CAPACITY_EXHAUSTED;
      f.kind = kind;
      f.x = x;
      f.y = y;
      throw f;
    }
    it->second[(size_t)y * W + x] -= 1;
  }

  // initial_fillers (electrostatic_placement.rs:198-240) with unit checks
  for (size_t i = 0; i < DU; ++i) {
    const uint32_t u = du_unit[i];
    for (uint32_t m = du_mstart[i]; m < du_mstart[i + 1]; ++m) {
      if (!std::isfinite(initial_xy[2 * u]) || !std::isfinite(initial_xy[2 * u + 1]) || !std::isfinite(m_off[2 * m]) ||
          !std::isfinite(m_off[2 * m + 1]))
        fail(TEXO_GP_ERR_NON_FINITE_UNIT_COORDINATE, (int64_t)i, (int64_t)(m - du_mstart[i]));
    }
  }
  available.assign(F, 0);
  real_area.assign(F, 0.0);
  std::vector<size_t> counts(F, 0);
  for (size_t m = 0; m < M; ++m) {
    counts[m_field[m]] += 1;
    real_area[m_field[m]] += m_charge[m];
  }
  for (int f = 0; f < F; ++f) {
    auto it = capacity.find(field_kind[f]);
    size_t total = 0;
    if (it != capacity.end())
      for (uint32_t c : it->second) total += c;
    available[f] = total;
    if (counts[f] > total) {
      Failure e;
      e.status = TEXO_GP_ERR_INSUFFICIENT_CAPACITY;
      e.kind = field_kind[f];
      e.v0 = (double)counts[f];
      e.v1 = (double)total;
      throw e;
    }
    const double avail = (double)total;
    if (!std::isfinite(real_area[f]) || real_area[f] < 0.0 || real_area[f] > avail + 1.0e-9) {
      Failure e;
      e.status = TEXO_GP_ERR_INSUFFICIENT_AREA;
      e.kind = field_kind[f];
      e.v0 = real_area[f];
      e.v1 = avail;
      throw e;
    }
    const double filler_area = avail - real_area[f];
    if (filler_area == 0.0) continue;
    std::vector<size_t> occupied;
    for (size_t i = 0; i < G; ++i)
      if (it->second[i] != 0) occupied.push_back(i);
    double assigned = 0.0;
    for (size_t ord = 0; ord < occupied.size(); ++ord) {
      const size_t index = occupied[ord];
      const double charge = ord + 1 == occupied.size() ? filler_area - assigned
                                                       : filler_area * (double)it->second[index] / avail;
      assigned += charge;
      fl_field.push_back(static_cast<uint8_t>(f));
      fl_charge.push_back(charge);
      fl_xy.push_back((double)(index % W));
      fl_xy.push_back((double)(index / W));
    }
  }
  Fl = fl_field.size();
  C = 2 * (U + Fl);
  // density_gamma_weights (eplace.rs:1424-1439)
  gamma_weight.assign(F, 0.0);
  for (size_t m = 0; m < M; ++m) gamma_weight[m_field[m]] += cell_weight[m_cell[m]];
  if (controls.size() != p->n_register_controls) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);

  // initial_coordinates_and_bounds (eplace.rs:1441-1468), then clamp
  lower.resize(C);
  upper.resize(C);
  for (size_t u = 0; u < U; ++u) {
    lower[2 * u] = p->unit_bounds[4 * u];
    lower[2 * u + 1] = p->unit_bounds[4 * u + 1];
    upper[2 * u] = p->unit_bounds[4 * u + 2];
    upper[2 * u + 1] = p->unit_bounds[4 * u + 3];
  }
  for (size_t j = 0; j < Fl; ++j) {
    lower[2 * (U + j)] = 0.0;
    lower[2 * (U + j) + 1] = 0.0;
    upper[2 * (U + j)] = (double)(W - 1);
    upper[2 * (U + j) + 1] = (double)(H - 1);
  }
  // the WA gradient bound of value-only origin evaluations: every pin
  // position the bounds allow (a unit's coordinates stay inside them)
  {
    double w = 0.0, lx = INFINITY, hx = -INFINITY, ly = INFINITY, hy = -INFINITY;
    size_t movable = 0;
    for (size_t n = 0; n < n_net_total; ++n) w += std::fabs(net_weight[n]) * (double)(net_start[n + 1] - net_start[n]);
    for (size_t q = 0; q < n_pins_total; ++q) {
      const uint32_t u = pin_unit[q];
      const double ox = pin_xy[2 * q], oy = pin_xy[2 * q + 1];
      double x0 = ox, x1 = ox, y0 = oy, y1 = oy;
      if (u != NO_UNIT) {
        movable += 1;
        x0 = lower[2 * (size_t)u] + ox;
        x1 = upper[2 * (size_t)u] + ox;
        y0 = lower[2 * (size_t)u + 1] + oy;
        y1 = upper[2 * (size_t)u + 1] + oy;
      }
      lx = std::min(lx, x0);
      hx = std::max(hx, x1);
      ly = std::min(ly, y0);
      hy = std::max(hy, y1);
    }
    wa_bound_w = w;
    wa_bound_ext = n_pins_total ? (hx - lx) + (hy - ly) : 0.0;
    wa_bound_ok = std::isfinite(wa_bound_w) && std::isfinite(wa_bound_ext);
    // WA bytes per evaluation: 16 per unit coordinate pair gathered and 16
    // per movable pin derivative stored; a value-only evaluation stores none
    wa_value_only_share = U + movable > 0 ? (double)U / (double)(U + movable) : 1.0;
  }
  set_charges(std::vector<double>(m_charge), std::vector<double>(fl_charge));
  stage(pool);
}


void Job::upload(const Basis *b, const std::shared_ptr<BlockPool> &pool, hipStream_t s) {
  basis = b;
  arena.commit(pool);
  HIP(hipMemcpyAsync(arena.base(), image, image_bytes, hipMemcpyHostToDevice, s));
  HIP(hipMemsetAsync(arena.base() + zero_begin, 0, zero_end - zero_begin, s));
  build_js();
}

// Points the design block's arrays at a shared copy (same offsets).
void Job::use_design(const std::shared_ptr<SharedDesign> &sd) {
  char *const from = arena.base(), *const to = static_cast<char *>(sd->base);
  auto move = [&](auto *&p) {
    if (!p) return;
    const char *c = reinterpret_cast<const char *>(p);
    if (c < from || c >= from + design_bytes) fail(TEXO_GP_ERR_DEVICE);
    p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(to + (c - from));
  };
  move(d_base);
  move(d_m_unit);
  move(d_m_off);
  move(d_m_off32);
  move(d_m_field);
  move(d_du_unit);
  move(d_du_mstart);
  move(d_fl_field);
  move(d_net_start);
  move(d_pin_unit);
  move(d_pin_xy);
  move(d_pin_xy32);
  move(d_net_weight);
  move(d_wa_seg);
  move(d_hpwl_large);
  move(d_u_pstart);
  move(d_pin_slot);
  move(d_u_nbase);
  move(d_wl_diag);
  move(d_unit_du);
  move(d_u_mstart);
  move(d_lower);
  move(d_upper);
  design = sd;
  build_js();
}

void Job::build_js() {
  const Basis *b = basis;
  js = JobStatic{};
  js.W = (int)W;
  js.H = (int)H;
  js.F = F;
  js.n_members = (int)M;
  js.n_fillers = (int)Fl;
  js.n_units = (int)U;
  js.n_density_units = (int)DU;
  js.m_unit = d_m_unit;
  js.m_off = OffArr{off32 ? nullptr : d_m_off, off32 ? d_m_off32 : nullptr};
  js.m_charge = d_m_charge;
  js.m_field = d_m_field;
  js.du_unit = d_du_unit;
  js.du_mstart = d_du_mstart;
  js.fl_field = d_fl_field;
  js.fl_charge = d_fl_charge;
  js.base_fx = d_base;
  js.unit_du = d_unit_du;
  js.u_mstart = d_u_mstart;
  js.perm = d_perm;
  js.s_ci = d_s_ci;
  js.s_off = OffArr{off32 ? nullptr : d_s_off, off32 ? d_s_off32 : nullptr};
  js.s_q = d_s_q;
  js.s_layer = d_s_layer;
  js.tiles_x = tiles_x;
  js.n_keys = n_keys;
  js.key_end = d_sort_count;
  js.lower = d_lower;
  js.upper = d_upper;
  js.n_wa_seg = (int)wa_seg.size();
  js.wa_seg = d_wa_seg;
  js.net_start = d_net_start;
  js.net_weight = d_net_weight;
  js.pin_unit = d_pin_unit;
  js.pin_xy = OffArr{off32 ? nullptr : d_pin_xy, off32 ? d_pin_xy32 : nullptr};
  js.u_pstart = d_u_pstart;
  js.pin_slot = d_pin_slot;
  js.u_nbase = d_u_nbase;
  for (int j = 0; j < 4; ++j) {
    js.b32[j] = b->b32[j];
    js.bthi[j] = b->bthi[j];
    js.btlo[j] = b->btlo[j];
    js.fthi[j] = b->fthi[j];
    js.ftlo[j] = b->ftlo[j];
    js.f16_scale[j] = b->f16_scale[j];
  }
  js.inv_lambda = b->inv_lambda;
}

// Installs member and filler charges (apply_area_adjustment, eplace.rs:793-819).
void Job::set_charges(const std::vector<double> &member_areas, const std::vector<double> &filler_areas) {
  m_charge = member_areas;
  fl_charge = filler_areas;
  h_du_fcharge.assign((size_t)std::max(F, 1) * DU, 0.0);
  for (size_t i = 0; i < DU; ++i)
    for (uint32_t m = du_mstart[i]; m < du_mstart[i + 1]; ++m) h_du_fcharge[(size_t)m_field[m] * DU + i] += m_charge[m];
  real_area.assign(F, 0.0);
  for (size_t m = 0; m < M; ++m) real_area[m_field[m]] += m_charge[m];
  field_charge.assign(F, 0.0);
  for (size_t m = 0; m < M; ++m) field_charge[m_field[m]] += std::fabs(m_charge[m]);
  for (size_t k = 0; k < fl_charge.size(); ++k) field_charge[fl_field[k]] += std::fabs(fl_charge[k]);
}

void Job::apply_charges(const std::vector<double> &member_areas, const std::vector<double> &filler_areas,
                        hipStream_t s) {
  set_charges(member_areas, filler_areas);
  if (M) HIP(hipMemcpyAsync(d_m_charge, m_charge.data(), M * 8, hipMemcpyHostToDevice, s));
  if (Fl) HIP(hipMemcpyAsync(d_fl_charge, fl_charge.data(), Fl * 8, hipMemcpyHostToDevice, s));
  if (DU) HIP(hipMemcpyAsync(d_du_fcharge, h_du_fcharge.data(), h_du_fcharge.size() * 8, hipMemcpyHostToDevice, s));
  // the sorted charge copies follow the new charges; the deposit order stays
  const size_t n = M + Fl;
  std::vector<uint32_t> perm(n);
  std::vector<double> q(n);
  if (n) {
    HIP(hipMemcpyAsync(perm.data(), d_perm, n * 4, hipMemcpyDeviceToHost, s));
    HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) q[i] = perm[i] < M ? m_charge[perm[i]] : fl_charge[perm[i] - M];
    HIP(hipMemcpyAsync(d_s_q, q.data(), n * 8, hipMemcpyHostToDevice, s));
  }
  HIP(hipStreamSynchronize(s));
}

void Job::release_image() {
  if (image_pool && image) image_pool->release(image, image_cap);
  image_pool.reset();
  image_heap.reset();
  image = nullptr;
}

// Arena plan and upload image of one job.  The arrays that start with data
// lead the arena, in one block that the image fills byte for byte: the
// problem arrays, the fixed-point background (-smoothed capacity per field),
// the clamped start coordinates (twice: current and previous iterate) and the
// charges in input order with their sorted-copy fields.  The deposit
// accumulators follow (cleared once at upload; k_convert leaves them at
// zero), then the working arrays.
void Job::stage(const std::shared_ptr<BlockPool> &pool) {
  const size_t FG = (size_t)std::max(F, 1) * G, n = M + Fl;
  auto blocks = [](size_t count, size_t per) { return std::max<size_t>(1, (count + per - 1) / per); };
  nb_wa = wa_waves == 0 ? 0 : (int)blocks((size_t)wa_waves, BLOCK / 64);
  nb_wa_live = wa_waves_live / (BLOCK / 64);
  nb_energy = F > 0 ? poisson_energy_blocks((int)W, (int)H, F) : 0;
  nb_gf = (int)blocks(Fl, BLOCK);
  nb_ug = (int)blocks(U, BLOCK);
  nb_conv = (int)blocks(FG, BLOCK);
  nb_tile = (int)std::max<size_t>(1, FG / 1024);
  tiles_x = (int)((W + (1u << SORT_TILE_SHIFT) - 1) >> SORT_TILE_SHIFT);
  n_keys = tiles_x * (int)((H + (1u << SORT_TILE_SHIFT) - 1) >> SORT_TILE_SHIFT);

  struct Copy {
    size_t at;
    const void *src;
    size_t bytes;
  };
  std::vector<Copy> copies;
  auto data = [&](auto *&ptr, size_t count, const void *src) {
    const size_t at = arena.reserve(ptr, count);
    if (src) copies.push_back({at, src, count * sizeof(*ptr)});
    return at;
  };
  // the design block: arrays fixed by the problem alone (not by the start
  // coordinates or the charges), shareable between jobs of one design
  const size_t o_base = data(d_base, FG, nullptr);
  data(d_m_unit, M, m_unit.data());
  // offsets as FP32 pairs when that is exact for all of them
  auto fp32_exact = [](const std::vector<double> &v) {
    for (double x : v)
      if (!(static_cast<double>(static_cast<float>(x)) == x)) return false;
    return true;
  };
  off32 = fp32_exact(m_off) && fp32_exact(pin_xy);
  std::vector<float> m_off32, pin_xy32;
  if (off32) {
    m_off32.assign(m_off.begin(), m_off.end());
    pin_xy32.assign(pin_xy.begin(), pin_xy.end());
    data(d_m_off32, M, m_off32.data());
    data(d_m_off, 0, nullptr);
  } else {
    data(d_m_off, M, m_off.data());
    data(d_m_off32, 0, nullptr);
  }
  data(d_m_field, M, m_field.data());
  data(d_du_unit, DU, du_unit.data());
  data(d_du_mstart, DU + 1, du_mstart.data());
  data(d_fl_field, Fl, fl_field.data());
  data(d_net_start, n_net_total + 1, net_start.data());
  data(d_pin_unit, n_pins_total, pin_unit.data());
  if (off32) {
    data(d_pin_xy32, n_pins_total, pin_xy32.data());
    data(d_pin_xy, 0, nullptr);
  } else {
    data(d_pin_xy, n_pins_total, pin_xy.data());
    data(d_pin_xy32, 0, nullptr);
  }
  data(d_net_weight, n_net_total, net_weight.data());
  data(d_wa_seg, wa_seg.size(), wa_seg.data());
  hpwl_large.clear();
  for (size_t k = 0; k < n_nets; ++k)
    if (net_start[k + 1] - net_start[k] > 32) hpwl_large.push_back((uint32_t)k);
  data(d_hpwl_large, hpwl_large.size(), hpwl_large.data());
  data(d_u_pstart, U + 1, u_pstart.data());
  data(d_pin_slot, n_pins_total, pin_slot.data());
  data(d_u_nbase, U, u_nbase.data());
  data(d_wl_diag, U, wl_diag.data());
  data(d_unit_du, U, unit_du.data());
  data(d_u_mstart, U + 1, u_mstart.data());
  data(d_lower, C, lower.data());
  data(d_upper, C, upper.data());
  design_bytes = arena.planned();
  // per job: charges (area rounds change them) and the start
  data(d_m_charge, M, m_charge.data());
  data(d_du_fcharge, (size_t)std::max(F, 1) * DU, h_du_fcharge.data());
  data(d_fl_charge, Fl, fl_charge.data());
  // the previous iterate starts as a copy of x (DynamicNesterovState::new,
  // on the device); the sorted charge copies are first written by the sort
  // that precedes any deposit, which walks the input order `perm`
  const size_t o_x0 = data(d_vec[ix], C, nullptr);
  const size_t o_perm = data(d_perm, n, nullptr);
  image_bytes = arena.planned();
  zero_begin = image_bytes;
  for (auto &sl : slot) {
    arena.reserve(sl.acc_m, FG);
    arena.reserve(sl.acc_f, FG);
  }
  zero_end = arena.planned();
  arena.reserve(d_metric, C);
  arena.reserve(d_part_metric, blocks(U + Fl, BLOCK) * 3);
  arena.reserve(d_part_step, blocks(C, BLOCK) * 4);
  arena.reserve(d_part_step_spec, blocks(C, BLOCK) * 4);
  for (int k = 0; k < 5; ++k)
    if (k != ix) arena.reserve(d_vec[k], C);
  arena.reserve(d_s_ci, n);
  arena.reserve(d_s_off, off32 ? 0 : n);
  arena.reserve(d_s_off32, off32 ? n : 0);
  arena.reserve(d_s_q, n);
  arena.reserve(d_s_layer, n);
  for (auto &g : d_grad) arena.reserve(g, C);
  arena.reserve(d_best, 2 * U);
  for (auto &sl : slot) {
    arena.reserve(sl.rho, FG);
    arena.reserve(sl.phi, FG);
    arena.reserve(sl.t0, FG);
    arena.reserve(sl.t1, FG);
    arena.reserve(sl.field_coef, 2 * MAXF);
    arena.reserve(sl.partial, (size_t)n_keys * 2 * std::max(F, 1) * DEP_WIN2);
    arena.reserve(sl.tile_flag, (size_t)n_keys);
    arena.reserve(sl.amax_part, (size_t)nb_conv * MAXF + 4 * (size_t)nb_tile);
    arena.reserve(sl.part_conv, (size_t)nb_conv * 2 * MAXF);
    arena.reserve(sl.dpin, u_pins.size());
    arena.reserve(sl.part_wa, (size_t)nb_wa);
    arena.reserve(sl.part_energy, (size_t)std::max(nb_energy, 1) * MAXF);
    arena.reserve(sl.part_gf, (size_t)nb_gf * GF_STRIDE);
    arena.reserve(sl.part_ug, (size_t)nb_ug * UG_STRIDE);
  }
  arena.reserve(d_rudy, 2 * G);
  arena.reserve(d_rudy_fx, 2 * G);
  arena.reserve(d_rudy_err, 1);
  arena.reserve(d_raw_units, (size_t)std::max(F, 1) * DU * 2);
  arena.reserve(d_raw_fillers, 2 * Fl);
  arena.reserve(d_density_out, FG);
  arena.reserve(d_wa_grad, 2 * U);
  arena.reserve(d_eval_coords, C);
  arena.reserve(d_perm_next, n);
  arena.reserve(d_sort_count, (size_t)n_keys + 1);
  arena.reserve(d_fixed_wa, 1);
  arena.reserve(d_hpwl, (size_t)hpwl_blocks((int)n_nets, (int)hpwl_large.size()) + 1);

  if (pool) {
    image = static_cast<char *>(pool->acquire(image_bytes, &image_cap));
    image_pool = pool;
  } else {
    image_heap.reset(new char[image_bytes]);
    image = image_heap.get();
  }
  for (const Copy &c : copies)
    if (c.bytes) std::memcpy(image + c.at, c.src, c.bytes);

  // Smoothed capacity (electrostatic_placement.rs:173-190, 900-916): a
  // capacity point sits on a bin centre, so its spline weights are 1/8,
  // 3/4, 1/8 per axis (clamped at the edges): the smoothing is one pass per
  // axis.  Every partial sum is a small multiple of 1/64, exact in FP64, so
  // this equals depositing the points one by one in any order.
  long long *base = reinterpret_cast<long long *>(image + o_base);
  std::fill(base, base + FG, 0ll);
  std::vector<double> row(G), smooth(G);
  for (int f = 0; f < F; ++f) {
    auto it = capacity.find(field_kind[f]);
    if (it == capacity.end()) continue;
    const uint32_t *cap = it->second.data();
    std::fill(row.begin(), row.end(), 0.0);
    std::fill(smooth.begin(), smooth.end(), 0.0);
    for (size_t y = 0; y < H; ++y)
      for (size_t x = 0; x < W; ++x) {
        const double q = (double)cap[y * W + x];
        if (q == 0.0) continue;
        double *r = row.data() + y * W;
        r[x == 0 ? 0 : x - 1] += 0.125 * q;
        r[x] += 0.75 * q;
        r[std::min<size_t>(x + 1, W - 1)] += 0.125 * q;
      }
    for (size_t y = 0; y < H; ++y) {
      const double *r = row.data() + y * W;
      double *lo = smooth.data() + (y == 0 ? 0 : y - 1) * W, *mid = smooth.data() + y * W,
             *hi = smooth.data() + std::min<size_t>(y + 1, H - 1) * W;
      for (size_t x = 0; x < W; ++x) {
        lo[x] += 0.125 * r[x];
        mid[x] += 0.75 * r[x];
        hi[x] += 0.125 * r[x];
      }
    }
    for (size_t i = 0; i < G; ++i) base[(size_t)f * G + i] = std::llrint(-smooth[i] * FX_SCALE);
  }
  double *x0 = reinterpret_cast<double *>(image + o_x0);
  for (size_t u = 0; u < U; ++u) {
    x0[2 * u] = initial_xy[2 * u];
    x0[2 * u + 1] = initial_xy[2 * u + 1];
  }
  for (size_t k = 0; k < Fl; ++k) {
    x0[2 * (U + k)] = fl_xy[2 * k];
    x0[2 * (U + k) + 1] = fl_xy[2 * k + 1];
  }
  for (size_t i = 0; i < C; ++i) x0[i] = clampd(x0[i], lower[i], upper[i]);
  // input order until the engine sorts the charges by tile
  uint32_t *perm = reinterpret_cast<uint32_t *>(image + o_perm);
  for (size_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
  // 128-bit fingerprint of the design block (array contents at their offsets),
  // matched when jobs of one design share its device copy.  Four independent
  // multiply-xorshift lanes over interleaved 8-byte words (a serial chain
  // hashes about 2 GB/s; the lanes run in parallel), folded into two words.
  constexpr uint64_t K1 = 0x9E3779B97F4A7C15ull, K2 = 0xC2B2AE3D27D4EB4Full;
  uint64_t h[4] = {0x243F6A8885A308D3ull ^ design_bytes, 0x13198A2E03707344ull ^ ((uint64_t)W << 32 | H),
                   0xA4093822299F31D0ull, 0x082EFA98EC4E6C89ull};
  auto lane = [](uint64_t v, uint64_t w, uint64_t k) {
    v = (v ^ w) * k;
    return v ^ (v >> 31);
  };
  auto mix = [&](const void *p, size_t bytes, size_t at) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (int l = 0; l < 4; ++l) h[l] = lane(h[l], at + (uint64_t)l, l & 1 ? K2 : K1);
    size_t k = 0;
    for (; k + 32 <= bytes; k += 32) {
      uint64_t w[4];
      std::memcpy(w, c + k, 32);
      h[0] = lane(h[0], w[0], K1);
      h[1] = lane(h[1], w[1], K2);
      h[2] = lane(h[2], w[2], K1);
      h[3] = lane(h[3], w[3], K2);
    }
    uint64_t w[4] = {0, 0, 0, 0};
    std::memcpy(w, c + k, bytes - k);
    for (int l = 0; l < 4; ++l) h[l] = lane(h[l], w[l] ^ bytes, l & 1 ? K2 : K1);
  };
  mix(image + o_base, FG * sizeof(long long), o_base);
  for (const Copy &c : copies)
    if (c.at < design_bytes) mix(c.src, c.bytes, c.at);
  uint64_t h1 = h[0] ^ (h[2] * K2), h2 = h[1] ^ (h[3] * K1);
  h1 ^= h1 >> 33;
  h1 *= 0xFF51AFD7ED558CCDull;
  h1 ^= h1 >> 33;
  h2 ^= h2 >> 31;
  h2 *= 0xC4CEB9FE1A85EC53ull;
  h2 ^= h2 >> 29;
  design_hash[0] = h1;
  design_hash[1] = h2;
}

Task Job::task(int sl, const double *coords, double *grad, const double *metric, double g, bool skip,
               EvalScalars *scalars, bool fixed_nets, bool value_only) const {
  Task t{};
  t.j = js;
  const Slot &w = slot[sl];
  t.acc_m = w.acc_m;
  t.acc_f = w.acc_f;
  t.rho = w.rho;
  t.phi = w.phi;
  t.t0 = w.t0;
  t.t1 = w.t1;
  t.field_coef = w.field_coef;
  t.partial = w.partial;
  t.amax_part = w.amax_part;
  t.part_conv = w.part_conv;
  t.nb_conv = F > 0 ? nb_conv : 0;
  t.nb_tile = nb_tile;
  t.tile_flag = w.tile_flag;
  t.dpin = w.dpin;
  t.part_wa = w.part_wa;
  t.part_energy = w.part_energy;
  t.part_gf = w.part_gf;
  t.part_ug = w.part_ug;
  // the nets of fixed pins run when their value at this gamma is not held
  t.nb_wa = fixed_nets ? nb_wa : nb_wa_live;
  t.nb_wa_live = nb_wa_live;
  t.fixed_wa = d_fixed_wa;
  t.fixed_wa_mode = nb_wa == nb_wa_live ? 0 : fixed_nets ? 1 : 2;
  t.nb_energy = F > 0 ? nb_energy : 0;
  t.nb_gf = F > 0 && Fl > 0 && !value_only ? nb_gf : 0;
  t.nb_ug = value_only ? 0 : nb_ug;
  t.value_only = value_only ? 1 : 0;
  t.coords = coords;
  t.grad = grad;
  t.metric = metric;
  t.gamma = g;
  t.skip_density = skip ? 1 : 0;
  for (int f = 0; f < MAXF; ++f) {
    t.lambda[f] = f < (int)mult.size() ? mult[f] : 0.0;
    t.scale[f] = f < (int)scales.size() ? scales[f] : 1.0;
  }
  t.scalars = scalars;
  return t;
}

SortTask Job::sort_task(const double *coords) {
  SortTask t{};
  t.j = js;
  t.coords = coords ? coords : d_vec[ix];
  t.count = d_sort_count;
  t.order = d_perm;
  t.perm = d_perm_next;
  // work enqueued after the sort reads the new order
  std::swap(d_perm, d_perm_next);
  js.perm = d_perm;
  t.s_ci = d_s_ci;
  t.s_off = off32 ? nullptr : d_s_off;
  t.s_off32 = off32 ? d_s_off32 : nullptr;
  t.s_q = d_s_q;
  t.s_layer = d_s_layer;
  return t;
}

double Job::overflow(const EvalScalars &s, int f) const {
  return static_cast<double>(s.over_fx[f]) * FX_INV / real_area[f];
}

bool Job::gradient_certainly_finite(const EvalScalars &s, double g) const {
  if (!wa_bound_ok || !(g > 0.0)) return false;
  double b = 3.0 * wa_bound_w * (1.0 + 3.0 * wa_bound_ext / g);
  // a field term of an entry is its coefficient times a raw entry, at most
  // the raw l1 norm of the current point (the same point, field and charges)
  for (int f = 0; f < F; ++f) {
    const double coef = mult[f] * (1.0 + AUGMENTED_DENSITY_BETA * (s.energy[f] / scales[f]));
    b += 2.0 * std::fabs(coef) * current.raw_l1[f];
  }
  return std::isfinite(b) && b <= 1e300;
}

double Job::combined_value(const EvalScalars &s) const {
  // combined_objective (eplace.rs:1022-1044) with
  // augmented_density_value_and_coefficient (eplace.rs:1111-1124)
  double v = s.wa_value;
  for (int f = 0; f < F; ++f) {
    const double e = s.energy[f] / scales[f];
    v += mult[f] * scales[f] * (e + 0.5 * AUGMENTED_DENSITY_BETA * e * e);
  }
  return v;
}

double Job::adaptive_gamma(const EvalScalars &s) const {
  // field_wirelength_gamma + weighted_gamma_mean (eplace.rs:1325-1395)
  double ws = 0.0, wsum = 0.0;
  const double base = opt.gamma_base;
  for (int f = 0; f < F; ++f) {
    const double o = overflow(s, f), target = overflow_target(field_kind[f]);
    if (!std::isfinite(o) || o < 0.0 || !(target >= 0.0 && target < 1.0) || !(base > 0.0))
      fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    const double slope = 2.0 / (1.0 - target);
    const double gamma = base * std::pow(10.0, slope * o + (1.0 - slope));
    if (!(std::isfinite(gamma) && gamma > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    const double w = gamma_weight[f];
    ws += gamma * w;
    wsum += w;
  }
  const double mean = wsum == 0.0 ? base : ws / wsum;
  if (!(std::isfinite(mean) && mean > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  return mean;
}

void Job::validate_initial(const EvalScalars &s) const {
  bool ok = std::isfinite(s.wa_value) && s.nonfinite == 0;
  for (int f = 0; f < F; ++f)
    ok &= std::isfinite(s.energy[f]) && s.energy[f] >= 0.0 && std::isfinite(overflow(s, f)) &&
          std::isfinite(s.force_l1[f]) && s.force_l1[f] >= 0.0;
  if (!ok) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
}

void Job::validate_fields(const EvalScalars &s) const {
  for (int f = 0; f < F; ++f)
    if (!std::isfinite(s.energy[f]) || !std::isfinite(scales[f]) || scales[f] <= 0.0)
      fail(TEXO_GP_ERR_DENSITY_FIELD_CHANGED);
}

std::vector<double> Job::subgradient(const EvalScalars &s) const {
  validate_fields(s);
  std::vector<double> v(F);
  for (int f = 0; f < F; ++f) {
    const double e = s.energy[f] / scales[f];
    v[f] = e + 0.5 * AUGMENTED_DENSITY_BETA * e * e;
    if (!(std::isfinite(v[f]) && v[f] >= 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  }
  return v;
}

bool Job::stop_reached(const EvalScalars &s, bool area_ready) const {
  for (int f = 0; f < F; ++f) {
    const uint32_t k = field_kind[f];
    if (area_ready) {
      if ((is_lut(k) || k == TEXO_KIND_REGISTER) && !(overflow(s, f) <= AREA_ADJUSTMENT_OVERFLOW_TARGET)) return false;
    } else if (!(overflow(s, f) <= overflow_target(k))) {
      return false;
    }
  }
  return true;
}

std::pair<double, double> Job::excess(const EvalScalars &s) const {
  double mx = 0.0, total = 0.0;
  for (int f = 0; f < F; ++f) {
    const double e = std::max(overflow(s, f) - overflow_target(field_kind[f]), 0.0);
    mx = std::max(mx, e);
    total += e;
  }
  return {mx, total};
}

double Job::hpwl(const std::vector<double> &x) const {
  double total = 0.0;
  for (size_t n = 0; n < n_nets; ++n) {
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint32_t p = net_start[n]; p < net_start[n + 1]; ++p) {
      double px = pin_xy[2 * p], py = pin_xy[2 * p + 1];
      if (pin_unit[p] != NO_UNIT) {
        px += x[2 * pin_unit[p]];
        py += x[2 * pin_unit[p] + 1];
      }
      xmin = std::min(xmin, px);
      xmax = std::max(xmax, px);
      ymin = std::min(ymin, py);
      ymax = std::max(ymax, py);
    }
    total += (xmax - xmin) + (ymax - ymin);
  }
  return total;
}

void Job::enqueue_hpwl(const double *coords, double *out, hipStream_t s) const {
  const size_t nb = (size_t)hpwl_blocks((int)n_nets, (int)hpwl_large.size());
  launch_hpwl((int)n_nets, js.net_start, js.pin_unit, js.pin_xy, coords, d_hpwl_large, (int)hpwl_large.size(), d_hpwl,
              d_hpwl + nb, s);
  HIP(hipMemcpyAsync(out, d_hpwl + nb, sizeof(double), hipMemcpyDeviceToHost, s));
}

size_t Job::area_round_limit() const {
  double initial = 0.0, avail = 0.0;
  for (double q : m_charge) initial += q;
  for (int f = 0; f < F; ++f) avail += (double)(uint32_t)available[f];
  if (!(std::isfinite(initial) && initial > 0.0 && avail >= initial)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  double t = initial;
  size_t rounds = 0;
  while (t < avail) {
    t *= 1.01;
    if (!std::isfinite(t)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    rounds += 1;
  }
  return rounds + 1;
}

void Job::update_multipliers() {
  const std::vector<double> s = subgradient(current);
  double n = 0.0;
  for (double v : s) n += v * v;
  n = std::sqrt(n);
  if (n != 0.0) {
    if (!std::isfinite(n) || !std::isfinite(mstep) || mstep <= 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    for (int f = 0; f < F; ++f) {
      mult[f] += mstep * s[f] / n;
      if (!std::isfinite(mult[f]) || mult[f] < 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    }
  }
  validate_fields(current);
  double en = 0.0;
  for (int f = 0; f < F; ++f) {
    const double e = current.energy[f] / scales[f];
    en += e * e;
  }
  const double lg = std::log(AUGMENTED_DENSITY_BETA * std::sqrt(en) + 1.0);
  if (!std::isfinite(lg) || lg < 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  const double growth = MULTIPLIER_ALPHA_LOW + (1.0 - 1.0 / (1.0 + lg)) * (MULTIPLIER_ALPHA_HIGH - MULTIPLIER_ALPHA_LOW);
  if (!(std::isfinite(growth) && growth >= MULTIPLIER_ALPHA_LOW && growth <= MULTIPLIER_ALPHA_HIGH))
    fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  const double next = mstep * growth;
  if (!(std::isfinite(next) && next > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  mstep = next;
}

// Area-round helpers (instance_area.rs:157-372, register_clustering.rs:17-96).
double soft_division_ceiling(double value, double divisor) {
  const double q = value / divisor;
  const double fl = std::floor(q);
  if (q - fl < 1.0 / divisor) return value + (1.0 - divisor) * fl;
  return std::ceil(q);
}

// clustering_areas (register_clustering.rs:17-81) for registers given as
// (cell, x, y); false reproduces the reference's None.
bool register_clustering(const std::vector<std::pair<uint32_t, std::pair<double, double>>> &regs,
                         const std::map<uint32_t, Control> &controls, uint64_t movable, std::vector<double> &areas) {
  areas.assign(regs.size(), 0.0);
  if (regs.empty() || movable == 0) return true;
  if (movable > 0xffffffffull) return false;
  const double sigma = std::sqrt(1.0e-5 * (double)(uint32_t)movable);
  if (!std::isfinite(sigma) || sigma <= 0.0) return false;
  const double half = 2.5 * sigma, scale = sigma * kSqrt2;
  std::map<std::pair<uint64_t, uint64_t>, std::vector<size_t>> groups;
  for (size_t i = 0; i < regs.size(); ++i) {
    auto it = controls.find(regs[i].first);
    if (it == controls.end()) return false;
    groups[{it->second.clock, it->second.lsr}].push_back(i);
  }
  auto prob = [&](double lo, double hi, double mean) {
    return clampd(0.5 * (std::erf((hi - mean) / scale) - std::erf((lo - mean) / scale)), 0.0, 1.0);
  };
  for (size_t r = 0; r < regs.size(); ++r) {
    const Control &focal = controls.at(regs[r].first);
    std::map<uint64_t, double> expected;
    const double fx = regs[r].second.first, fy = regs[r].second.second;
    for (size_t o : groups[{focal.clock, focal.lsr}]) {
      const double px = prob(fx - half, fx + half, regs[o].second.first);
      const double py = prob(fy - half, fy + half, regs[o].second.second);
      expected[controls.at(regs[o].first).ce] += px * py;
    }
    const double ef = expected.count(focal.ce) ? expected[focal.ce] : 0.0;
    if (!(std::isfinite(ef) && ef > 0.0)) return false;
    double total = 0.0;
    for (auto &kv : expected) total += soft_division_ceiling(kv.second, 2.0);
    const double area = 8.0 / ef * soft_division_ceiling(ef, 2.0) / total * soft_division_ceiling(total, 4.0);
    if (!(std::isfinite(area) && area >= 0.0)) return false;
    areas[r] = area;
  }
  return true;
}



void add_checked(std::map<uint32_t, double> &t, uint32_t k, double v) {
  double &x = t[k];
  x += v;
  if (!std::isfinite(x)) fail(TEXO_GP_ERR_ARITHMETIC_OVERFLOW);
}

// adjust_instance_areas (instance_area.rs:157-253) with
// adjusted_filler_areas (:293-332) and validate_postconditions (:334-367).
bool adjust_areas(const std::vector<AreaMember> &members, const std::vector<std::pair<uint32_t, double>> &fillers,
                  std::vector<double> &member_out, std::vector<double> &filler_out) {
  for (size_t i = 0; i < members.size(); ++i) {
    const AreaMember &m = members[i];
    if (!std::isfinite(m.current) || m.current < 0.0 || !std::isfinite(m.routability) || m.routability < 0.0 ||
        !std::isfinite(m.clustering) || m.clustering < 0.0)
      fail(TEXO_GP_ERR_INVALID_MEMBER_AREA, (int64_t)i);
  }
  for (size_t i = 0; i < fillers.size(); ++i)
    if (!std::isfinite(fillers[i].second) || fillers[i].second < 0.0) fail(TEXO_GP_ERR_INVALID_FILLER_AREA, (int64_t)i);
  const size_t n = members.size();
  std::vector<double> requested(n), delta(n);
  std::map<uint32_t, double> desired, avail, scales, target, actual;
  for (size_t i = 0; i < n; ++i) {
    // pin_area is supplied as 0.0 by the caller (eplace.rs:708-710)
    requested[i] = std::max(std::max(std::max(members[i].routability, 0.0), members[i].clustering), members[i].current);
    delta[i] = requested[i] - members[i].current;
    add_checked(desired, members[i].kind, delta[i]);
  }
  for (auto &f : fillers) add_checked(avail, f.first, f.second);
  for (auto &kv : desired) {
    const double a = avail.count(kv.first) ? avail[kv.first] : 0.0;
    scales[kv.first] = kv.second == 0.0 ? 1.0 : std::min(a / kv.second, 1.0);
    target[kv.first] = std::min(kv.second, a);
  }
  std::map<uint32_t, size_t> last;
  for (size_t i = 0; i < n; ++i) last[members[i].kind] = i;
  member_out.assign(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t k = members[i].kind;
    const double scale = scales[k];
    double increase;
    if (last[k] == i) {
      const double assigned = actual.count(k) ? actual[k] : 0.0;
      increase = clampd(target[k] - assigned, 0.0, delta[i]);
    } else {
      increase = delta[i] * scale;
    }
    const double adjusted = scale == 1.0 ? requested[i] : members[i].current + increase;
    if (!std::isfinite(adjusted)) fail(TEXO_GP_ERR_ARITHMETIC_OVERFLOW);
    add_checked(actual, k, adjusted - members[i].current);
    member_out[i] = adjusted;
  }
  bool all_zero = true;
  for (auto &kv : actual) all_zero = all_zero && kv.second == 0.0;
  filler_out.assign(fillers.size(), 0.0);
  if (all_zero) {
    for (size_t i = 0; i < fillers.size(); ++i) filler_out[i] = fillers[i].second;
  } else {
    std::map<uint32_t, double> remaining, assigned;
    for (auto &kv : avail) remaining[kv.first] = std::max(kv.second - (actual.count(kv.first) ? actual[kv.first] : 0.0), 0.0);
    std::map<uint32_t, size_t> lastf;
    for (size_t i = 0; i < fillers.size(); ++i) lastf[fillers[i].first] = i;
    for (size_t i = 0; i < fillers.size(); ++i) {
      const uint32_t k = fillers[i].first;
      const double cur = fillers[i].second, total = avail[k], tgt = remaining[k];
      double area;
      if (lastf[k] == i)
        area = std::max(tgt - (assigned.count(k) ? assigned[k] : 0.0), 0.0);
      else if (total == 0.0)
        area = 0.0;
      else
        area = cur * tgt / total;
      const double tol = 64.0 * DBL_EPSILON * std::max(std::fabs(cur), 1.0);
      if (!std::isfinite(area) || area > cur + tol) fail(TEXO_GP_ERR_ARITHMETIC_OVERFLOW);
      area = std::min(area, cur);
      add_checked(assigned, k, area);
      filler_out[i] = area;
    }
  }
  std::map<uint32_t, double> before, after;
  for (size_t i = 0; i < n; ++i) {
    if (member_out[i] < members[i].current || !std::isfinite(member_out[i])) fail(TEXO_GP_ERR_POSTCONDITION_VIOLATION);
    add_checked(before, members[i].kind, members[i].current);
    add_checked(after, members[i].kind, member_out[i]);
  }
  for (size_t i = 0; i < fillers.size(); ++i) {
    if (filler_out[i] < 0.0 || filler_out[i] > fillers[i].second || !std::isfinite(filler_out[i]))
      fail(TEXO_GP_ERR_POSTCONDITION_VIOLATION);
    add_checked(before, fillers[i].first, fillers[i].second);
    add_checked(after, fillers[i].first, filler_out[i]);
  }
  const double terms = std::max((double)(uint32_t)(n + fillers.size()), 1.0);
  for (auto &kv : before) {
    const double a = after.count(kv.first) ? after[kv.first] : 0.0;
    if (std::fabs(kv.second - a) > 64.0 * DBL_EPSILON * terms * std::max(std::fabs(kv.second), 1.0))
      fail(TEXO_GP_ERR_POSTCONDITION_VIOLATION);
  }
  bool reset = false;
  for (auto &kv : actual) reset = reset || kv.second > 0.0;
  return reset;
}


}  // namespace texo
