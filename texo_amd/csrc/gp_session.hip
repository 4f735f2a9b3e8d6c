// gp_session.hip — host side of the MI355X electrostatic global placement:
// problem set-up (eplace.rs:255-336), the device-resident objective
// evaluation (ContinuousProblem::evaluate, eplace.rs:137-205), the dynamic
// Nesterov step (global_placement.rs:121-335) driven from FP64 device
// reductions, the lambda/gamma/stop schedule (eplace.rs:1111-1415), area
// rounds (eplace.rs:353-420, 627-819) and the C ABI of include/texo_gp.h.

#include "gp_common.h"
#include "../../include/texo_gp.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <string>
#include <vector>

namespace texo {

thread_local KernelTimer *g_timer = nullptr;

struct KernelTimer {
  std::string name;
  bool enabled = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t used = 0;
};

void timer_begin(KernelTimer *t, const char *name, hipStream_t s) {
  if (!t || !t->enabled || t->name != name) return;
  if (t->used == t->events.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    t->events.push_back({a, b});
  }
  (void)hipEventRecord(t->events[t->used].first, s);
}

void timer_end(KernelTimer *t, const char *name, hipStream_t s) {
  if (!t || !t->enabled || t->name != name || t->used >= t->events.size()) return;
  (void)hipEventRecord(t->events[t->used].second, s);
  t->used += 1;
}

namespace {

thread_local std::string g_last_error;

// eplace.rs:125-134
constexpr double INITIAL_DENSITY_WEIGHT = 1.0e-4;
constexpr double AREA_RESET_DENSITY_WEIGHT = 0.1;
constexpr double MULTIPLIER_ALPHA_LOW = 1.05;
constexpr double MULTIPLIER_ALPHA_HIGH = 1.06;
constexpr double AREA_ADJUSTMENT_OVERFLOW_TARGET = 0.15;
constexpr double WIRELENGTH_GAMMA_BASE_TILES = 8.0;
constexpr uint64_t MAX_EPLACE_ITERATIONS = 3000;
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr uint32_t LARGE_NET_PINS = 16;  // nets above this use one wavefront

struct Failure {
  int status;
  int64_t index = -1, index2 = -1;
  uint32_t kind = 0, x = 0, y = 0;
  double v0 = 0.0, v1 = 0.0;
};

[[noreturn]] void fail(int status, int64_t index = -1, int64_t index2 = -1) {
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
#define HIP(x) ::texo::hip_check((x), #x)

bool is_lut(uint32_t k) { return k >= 0x100u && k < 0x200u; }
bool exchangeable(uint32_t k) { return is_lut(k) || k == TEXO_KIND_REGISTER || k == TEXO_KIND_MEMORY; }

// eplace.rs:1300-1308
double overflow_target(uint32_t kind) {
  if (kind == TEXO_KIND_MEMORY || kind == TEXO_KIND_DSP) return 0.20;
  if (is_lut(kind) || kind == TEXO_KIND_REGISTER) return 0.10;
  return 0.0;
}

// crates/texo-pnr/src/lib.rs:4566-4568
uint32_t rounded_coordinate(double v, uint32_t extent) {
  double r = std::round(v);
  const double hi = static_cast<double>(extent - 1);
  if (!(r >= 0.0)) r = 0.0;
  if (r > hi) r = hi;
  return static_cast<uint32_t>(r);
}

double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Quadratic B-spline deposit on the host, used once per call for the
// smoothed capacity background (electrostatic_placement.rs:173-190, 900-916).
void host_deposit(std::vector<double> &grid, size_t W, size_t H, double x, double y, double q) {
  auto axis = [](size_t L, double c, size_t idx[3], double w[3]) {
    const double maximum = static_cast<double>(L - 1);
    c = clampd(c, 0.0, maximum);
    const double cc = std::min(std::floor(c + 0.5), maximum);
    const size_t center = static_cast<size_t>(cc);
    const double o = c - cc;
    idx[0] = center == 0 ? 0 : center - 1;
    idx[1] = center;
    idx[2] = std::min(center + 1, L - 1);
    w[0] = 0.5 * ((0.5 - o) * (0.5 - o));
    w[1] = 0.75 - o * o;
    w[2] = 0.5 * ((0.5 + o) * (0.5 + o));
  };
  size_t ix[3], iy[3];
  double wx[3], wy[3];
  axis(W, x, ix, wx);
  axis(H, y, iy, wy);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) grid[iy[r] * W + ix[c]] += q * wx[c] * wy[r];
}

// ---------------------------------------------------------------------------
// Device arena: one allocation per session, carved into aligned buffers.
// ---------------------------------------------------------------------------
class Arena {
 public:
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  template <class T>
  void reserve(T *&p, size_t count) {
    plan_.push_back({reinterpret_cast<void **>(&p), std::max<size_t>(count, 1) * sizeof(T)});
  }
  size_t commit() {
    size_t total = 0;
    for (auto &e : plan_) total += (e.bytes + 255) & ~size_t(255);
    HIP(hipMalloc(&base_, total));
    char *c = static_cast<char *>(base_);
    for (auto &e : plan_) {
      *e.slot = c;
      c += (e.bytes + 255) & ~size_t(255);
    }
    bytes_ = total;
    plan_.clear();
    return total;
  }
  size_t bytes() const { return bytes_; }

 private:
  struct Entry {
    void **slot;
    size_t bytes;
  };
  std::vector<Entry> plan_;
  void *base_ = nullptr;
  size_t bytes_ = 0;
};

struct Options {
  uint32_t precision = TEXO_GP_DCT_FP32;
  uint32_t fixed_iterations = 0;
  uint64_t max_iterations = MAX_EPLACE_ITERATIONS;
  double gamma_base = WIRELENGTH_GAMMA_BASE_TILES;
  bool trace = false;
  int device = 0;
  uint32_t inject = 0;
};

struct HostEval {
  EvalScalars s;
  double value = 0.0;  // combined objective (filled by the driver)
  double kkt() const {
    double v;
    std::memcpy(&v, &s.kkt_bits, sizeof(v));
    return v;
  }
};

struct Control {
  uint64_t clock, lsr, ce;
};

}  // namespace

// ---------------------------------------------------------------------------
// Session
// ---------------------------------------------------------------------------
struct SessionImpl {
  Options opt;
  hipStream_t stream = nullptr;
  hipStream_t wa_stream = nullptr;  // WA runs beside the density chain
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int device = 0;
  KernelTimer timer;

  // problem (host)
  uint32_t W = 0, H = 0;
  size_t G = 0, U = 0, DU = 0, M = 0, Fl = 0, C = 0, n_cells = 0;
  int F = 0;
  std::vector<uint32_t> field_kind;
  std::vector<uint8_t> unit_fixed;
  std::vector<double> initial_xy, lower, upper;
  std::vector<uint32_t> du_unit, du_mstart, m_unit, m_cell;
  std::vector<uint8_t> m_field;
  std::vector<double> m_off, m_charge;
  std::vector<uint8_t> fl_field;
  std::vector<double> fl_charge, fl_xy;
  std::vector<int32_t> unit_du;
  std::map<uint32_t, std::vector<uint32_t>> capacity;  // after fixed occupancy
  std::vector<size_t> available;                       // per field
  std::vector<double> real_area;                       // per field
  std::vector<double> gamma_weight;                    // per field
  std::vector<double> wl_diag;                         // per unit
  uint64_t movable_instance_count = 0;
  bool empty = false;
  bool best_effort = false;
  bool has_routing = false;
  std::vector<uint32_t> routing_h, routing_v;
  std::map<uint32_t, Control> controls;
  // hypergraph, flattened with timed arcs appended as 2-pin nets
  size_t n_nets = 0, n_arcs = 0, n_net_total = 0, n_pins_total = 0, first_arc_pin = 0;
  std::vector<uint32_t> net_start, pin_unit;
  std::vector<double> pin_xy, net_weight;
  std::vector<uint32_t> small_nets, large_nets, u_pstart, u_pins;

  // device
  Arena arena;
  DensityArgs dens{};
  WaArgs wa{};
  DctPlan dct{};
  long long *d_base = nullptr, *d_acc_m = nullptr, *d_acc_f = nullptr;
  uint32_t *d_m_unit = nullptr, *d_du_unit = nullptr, *d_du_mstart = nullptr;
  double2 *d_m_off = nullptr;
  double *d_m_charge = nullptr, *d_fl_charge = nullptr, *d_du_fcharge = nullptr;
  uint8_t *d_m_field = nullptr, *d_fl_field = nullptr;
  uint32_t *d_net_start = nullptr, *d_pin_unit = nullptr, *d_small = nullptr, *d_large = nullptr;
  uint32_t *d_u_pstart = nullptr, *d_u_pins = nullptr;
  double2 *d_pin_xy = nullptr, *d_dpin = nullptr;
  double *d_net_weight = nullptr, *d_wl_diag = nullptr;
  int32_t *d_unit_du = nullptr;
  float *d_rho = nullptr, *d_phi = nullptr, *d_t0 = nullptr, *d_t1 = nullptr, *d_inv_lambda = nullptr;
  float *d_b32[4] = {};
  uint16_t *d_bhi[4] = {}, *d_blo[4] = {}, *d_bthi[4] = {}, *d_btlo[4] = {};
  double *d_lower = nullptr, *d_upper = nullptr, *d_metric = nullptr;
  double *d_vec[4] = {};   // coordinate buffers (rotated)
  double *d_grad[3] = {};  // origin, y, candidate gradients
  double2 *d_gfu = nullptr;
  double *d_part_wa = nullptr, *d_part_energy = nullptr, *d_part_gu = nullptr, *d_part_gf = nullptr,
         *d_part_ug = nullptr, *d_part_step = nullptr, *d_part_metric = nullptr;
  EvalScalars *d_eval = nullptr;  // 4 slots
  StepScalars *d_step = nullptr;
  MetricScalars *d_metric_s = nullptr;
  unsigned int *d_flag = nullptr;
  uint32_t *d_best = nullptr;
  double *d_rudy = nullptr;       // [2G]
  long long *d_rudy_fx = nullptr; // [2G]
  unsigned int *d_rudy_err = nullptr;
  // scratch for parity outputs
  double *d_raw_units = nullptr, *d_raw_fillers = nullptr, *d_density_out = nullptr, *d_wa_grad = nullptr;
  double *d_eval_coords = nullptr;

  // pinned host mirrors
  EvalScalars *h_eval = nullptr;
  StepScalars *h_step = nullptr;
  MetricScalars *h_metric = nullptr;
  unsigned int *h_flag = nullptr;

  // state
  int ix = 0, ixp = 1, iy = 2, ic = 3;  // indices into d_vec
  uint64_t iterations = 0, evaluations = 0, trials = 0, area_rounds = 0;
  bool started = false;
  double momentum = 1.0, lipschitz = 0.0;
  bool has_lipschitz = false;
  std::vector<double> scales, mult;
  double mstep = 0.0;
  HostEval current;
  double last_wa = 0.0;
  bool has_best = false;
  double best_max = 0.0, best_total = 0.0;
  std::vector<uint32_t> prev_stationary;
  bool has_prev_stationary = false;

  ~SessionImpl() {
    for (auto &e : timer.events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    if (h_eval) (void)hipHostFree(h_eval);
    if (h_step) (void)hipHostFree(h_step);
    if (h_metric) (void)hipHostFree(h_metric);
    if (h_flag) (void)hipHostFree(h_flag);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (wa_stream) (void)hipStreamDestroy(wa_stream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // ---------------- set-up ----------------
  void load(const texo_gp_problem *p, const texo_gp_options *o);
  void upload();
  void build_dct();

  // ---------------- evaluation ----------------
  void evaluate(const double *coords, double gamma, const double *metric, double *grad, int slot,
                bool parity_outputs, double *density_out = nullptr);
  void fetch_eval(int slot) {
    HIP(hipMemcpyAsync(&h_eval[slot], &d_eval[slot], sizeof(EvalScalars), hipMemcpyDeviceToHost, stream));
  }
  void sync() { HIP(hipStreamSynchronize(stream)); }
  double overflow(const EvalScalars &s, int f) const {
    return static_cast<double>(s.over_fx[f]) * FX_INV / real_area[f];
  }
  double combined_value(const EvalScalars &s) const {
    double v = s.wa_value;
    for (int f = 0; f < F; ++f) {
      const double e = s.energy[f] / scales[f];
      v += mult[f] * scales[f] * (e + 0.5 * AUGMENTED_DENSITY_BETA * e * e);
    }
    return v;
  }

  // ---------------- schedule ----------------
  double adaptive_gamma(const EvalScalars &s) const;
  void validate_initial(const EvalScalars &s) const;
  void validate_fields(const EvalScalars &s) const;
  std::vector<double> subgradient(const EvalScalars &s) const;
  void initialise_round(bool reset_after_area, const std::vector<double> *fixed_scales);
  bool stop_reached(const EvalScalars &s, bool area_ready) const;
  std::pair<double, double> excess(const EvalScalars &s) const;
  void consider_best(const EvalScalars &s);
  bool nesterov_step(double gamma, HostEval &payload, bool &accepted);
  void iteration(bool check_stationary, bool &stop_stationary);
  void set_preconditioner();

  // ---------------- area rounds ----------------
  void area_round_metrics(double &total_relative_increase, bool &requires_reset, std::vector<double> &member_areas,
                          std::vector<double> &filler_areas);
  void apply_charges(const std::vector<double> &member_areas, const std::vector<double> &filler_areas);
  size_t area_round_limit() const;

  int run(uint32_t *targets, texo_gp_report *report);
  void fill_report(texo_gp_report *r, const EvalScalars *fields);
  double hpwl_of(const std::vector<double> &x) const;
};

// ===========================================================================
// Set-up
// ===========================================================================
void SessionImpl::load(const texo_gp_problem *p, const texo_gp_options *o) {
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
  if (opt.precision == TEXO_GP_DCT_FP16_SCALED) {
    g_last_error = "FP16-scaled transform is not available in this build";
    fail(TEXO_GP_ERR_UNSUPPORTED);
  }
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
  std::set<uint32_t> kinds;
  unit_du.assign(U, -1);
  for (size_t u = 0; u < U; ++u) {
    const uint32_t b = p->unit_member_start[u], e = p->unit_member_start[u + 1];
    if (unit_fixed[u] || b == e) continue;
    unit_du[u] = static_cast<int32_t>(du_unit.size());
    du_unit.push_back(static_cast<uint32_t>(u));
    du_mstart.push_back(static_cast<uint32_t>(m_unit.size()));
    for (uint32_t m = b; m < e; ++m) {
      if (!exchangeable(p->member_kind[m]) || p->member_cell[m] >= n_cells) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)m);
      kinds.insert(p->member_kind[m]);
      m_unit.push_back(static_cast<uint32_t>(u));
      m_cell.push_back(p->member_cell[m]);
      m_field.push_back(0);  // filled below
      m_off.push_back(p->member_offset_xy[2 * m]);
      m_off.push_back(p->member_offset_xy[2 * m + 1]);
      m_charge.push_back(1.0);
    }
  }
  du_mstart.push_back(static_cast<uint32_t>(m_unit.size()));
  DU = du_unit.size();
  M = m_unit.size();
  field_kind.assign(kinds.begin(), kinds.end());
  F = static_cast<int>(field_kind.size());
  if (F > MAXF) fail(TEXO_GP_ERR_UNSUPPORTED);
  {
    size_t k = 0;
    for (size_t u = 0; u < U; ++u) {
      const uint32_t b = p->unit_member_start[u], e = p->unit_member_start[u + 1];
      if (unit_fixed[u] || b == e) continue;
      for (uint32_t m = b; m < e; ++m)
        m_field[k++] = static_cast<uint8_t>(std::find(field_kind.begin(), field_kind.end(), p->member_kind[m]) -
                                            field_kind.begin());
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
  pin_unit.resize(n_pins_total);
  pin_xy.resize(2 * n_pins_total);
  for (size_t i = 0; i < n_pins_total; ++i) {
    const uint32_t c = pin_cell[i];
    if (c >= n_cells) fail(TEXO_GP_ERR_INVALID_ARGUMENT, (int64_t)i);
    if (p->cell_fixed[c]) {
      pin_unit[i] = NO_UNIT;
      pin_xy[2 * i] = p->cell_fixed_xy[2 * c];
      pin_xy[2 * i + 1] = p->cell_fixed_xy[2 * c + 1];
    } else {
      pin_unit[i] = p->cell_unit[c];
      pin_xy[2 * i] = p->cell_offset_xy[2 * c];
      pin_xy[2 * i + 1] = p->cell_offset_xy[2 * c + 1];
    }
  }
  for (size_t n = 0; n < n_net_total; ++n)
    (net_start[n + 1] - net_start[n] > LARGE_NET_PINS ? large_nets : small_nets).push_back(static_cast<uint32_t>(n));
  // unit -> incident movable pins (pin order; nets before arcs)
  u_pstart.assign(U + 1, 0);
  for (size_t i = 0; i < n_pins_total; ++i)
    if (pin_unit[i] != NO_UNIT) u_pstart[pin_unit[i] + 1] += 1;
  for (size_t u = 0; u < U; ++u) u_pstart[u + 1] += u_pstart[u];
  u_pins.assign(u_pstart[U], 0);
  {
    std::vector<uint32_t> fill(u_pstart.begin(), u_pstart.end() - 1);
    for (size_t i = 0; i < n_pins_total; ++i)
      if (pin_unit[i] != NO_UNIT) u_pins[fill[pin_unit[i]]++] = static_cast<uint32_t>(i);
  }
  // wirelength_preconditioner (analytical_placement.rs:177-211) and
  // baseline_cell_incidence_weights (:219-231)
  wl_diag.assign(U, 0.0);
  std::vector<double> cell_weight(n_cells, 0.0);
  {
    std::vector<uint32_t> units, cells;
    for (size_t n = 0; n < n_nets; ++n) {
      units.clear();
      cells.clear();
      for (uint32_t q = net_start[n]; q < net_start[n + 1]; ++q) {
        units.push_back(p->cell_unit[pin_cell[q]]);
        cells.push_back(pin_cell[q]);
      }
      const double w = 1.0 / (double)(uint32_t)(net_start[n + 1] - net_start[n] - 1);
      std::sort(units.begin(), units.end());
      units.erase(std::unique(units.begin(), units.end()), units.end());
      if (units.size() >= 2)
        for (uint32_t u : units) wl_diag[u] += w;
      std::sort(cells.begin(), cells.end());
      cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
      for (uint32_t c : cells) cell_weight[c] += w;
    }
    for (size_t a = 0; a < n_arcs; ++a) {
      const uint32_t du_ = p->cell_unit[p->arc_cells[2 * a]], su = p->cell_unit[p->arc_cells[2 * a + 1]];
      if (du_ == su) continue;
      wl_diag[du_] += net_weight[n_nets + a];
      wl_diag[su] += net_weight[n_nets + a];
    }
  }
  has_routing = p->routing_horizontal && p->routing_vertical;
  if (has_routing) {
    routing_h.assign(p->routing_horizontal, p->routing_horizontal + G);
    routing_v.assign(p->routing_vertical, p->routing_vertical + G);
  }
  for (size_t i = 0; i < p->n_register_controls; ++i)
    controls[p->register_cell[i]] = {p->register_clock_lsr[2 * i], p->register_clock_lsr[2 * i + 1],
                                     p->register_ce[i]};

  if (DU == 0) {  // eplace.rs:303-310
    empty = true;
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
}

void SessionImpl::build_dct() {
  auto basis = [](size_t n) {
    std::vector<double> b(n * n);
    for (size_t k = 0; k < n; ++k) {
      const double s = k == 0 ? std::sqrt(1.0 / (double)n) : std::sqrt(2.0 / (double)n);
      for (size_t x = 0; x < n; ++x) b[k * n + x] = s * std::cos(kPi * (double)k * ((double)x + 0.5) / (double)n);
    }
    return b;
  };
  const std::vector<double> cw = basis(W), ch = basis(H);
  // operand j: 0 = C_W^T, 1 = C_H^T, 2 = C_H, 3 = C_W  (row-major K x N)
  const size_t dims[4] = {W, H, H, W};
  for (int j = 0; j < 4; ++j) {
    const size_t n = dims[j];
    const std::vector<double> &c = (j == 0 || j == 3) ? cw : ch;
    const bool transpose = j < 2;
    std::vector<float> f32(n * n);
    std::vector<uint16_t> hi(n * n), lo(n * n);
    for (size_t r = 0; r < n; ++r)
      for (size_t k = 0; k < n; ++k) {
        const double v = transpose ? c[k * n + r] : c[r * n + k];
        const float fv = static_cast<float>(v);
        f32[r * n + k] = fv;
        uint32_t u;
        std::memcpy(&u, &fv, 4);
        uint32_t h = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
        float hf;
        uint32_t hb = h << 16;
        std::memcpy(&hf, &hb, 4);
        const float rem = static_cast<float>(v - (double)hf);
        uint32_t ru;
        std::memcpy(&ru, &rem, 4);
        hi[r * n + k] = static_cast<uint16_t>(h);
        lo[r * n + k] = static_cast<uint16_t>((ru + 0x7fffu + ((ru >> 16) & 1u)) >> 16);
      }
    HIP(hipMemcpy(d_b32[j], f32.data(), n * n * sizeof(float), hipMemcpyHostToDevice));
    HIP(hipMemcpy(d_bhi[j], hi.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP(hipMemcpy(d_blo[j], lo.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
    std::vector<uint16_t> thi(n * n), tlo(n * n);
    for (size_t r = 0; r < n; ++r)
      for (size_t k = 0; k < n; ++k) {
        thi[k * n + r] = hi[r * n + k];
        tlo[k * n + r] = lo[r * n + k];
      }
    HIP(hipMemcpy(d_bthi[j], thi.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP(hipMemcpy(d_btlo[j], tlo.data(), n * n * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  // Poisson2d::new (electrostatic_placement.rs:675-714)
  std::vector<float> il(G);
  for (size_t ky = 0; ky < H; ++ky) {
    const double sy = std::sin(kPi * (double)ky / (2.0 * (double)H));
    for (size_t kx = 0; kx < W; ++kx) {
      const double sx = std::sin(kPi * (double)kx / (2.0 * (double)W));
      const double lam = 4.0 * sx * sx + 4.0 * sy * sy;
      il[ky * W + kx] = (kx == 0 && ky == 0) ? 0.0f : static_cast<float>(1.0 / lam);
    }
  }
  HIP(hipMemcpy(d_inv_lambda, il.data(), G * sizeof(float), hipMemcpyHostToDevice));
  dct.W = W;
  dct.H = H;
  dct.F = F;
  dct.precision = static_cast<int>(opt.precision);
  for (int j = 0; j < 4; ++j) {
    dct.b32[j] = d_b32[j];
    dct.bhi[j] = d_bhi[j];
    dct.blo[j] = d_blo[j];
    dct.bthi[j] = d_bthi[j];
    dct.btlo[j] = d_btlo[j];
    dct.bh16[j] = nullptr;
  }
  dct.inv_lambda = d_inv_lambda;
  dct.t0 = d_t0;
  dct.t1 = d_t1;
}

void SessionImpl::upload() {
  HIP(hipSetDevice(opt.device));
  device = opt.device;
  HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP(hipStreamCreateWithFlags(&wa_stream, hipStreamNonBlocking));
  HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  HIP(hipHostMalloc(reinterpret_cast<void **>(&h_eval), 4 * sizeof(EvalScalars), hipHostMallocDefault));
  HIP(hipHostMalloc(reinterpret_cast<void **>(&h_step), sizeof(StepScalars), hipHostMallocDefault));
  HIP(hipHostMalloc(reinterpret_cast<void **>(&h_metric), sizeof(MetricScalars), hipHostMallocDefault));
  HIP(hipHostMalloc(reinterpret_cast<void **>(&h_flag), sizeof(unsigned int), hipHostMallocDefault));
  const size_t FG = (size_t)std::max(F, 1) * G;
  const size_t Cc = std::max<size_t>(C, 2 * U);
  arena.reserve(d_base, FG);
  arena.reserve(d_acc_m, FG);
  arena.reserve(d_acc_f, FG);
  arena.reserve(d_m_unit, M);
  arena.reserve(d_m_off, M);
  arena.reserve(d_m_charge, M);
  arena.reserve(d_m_field, M);
  arena.reserve(d_du_unit, DU);
  arena.reserve(d_du_mstart, DU + 1);
  arena.reserve(d_du_fcharge, (size_t)std::max(F, 1) * DU);
  arena.reserve(d_fl_field, Fl);
  arena.reserve(d_fl_charge, Fl);
  arena.reserve(d_net_start, n_net_total + 1);
  arena.reserve(d_pin_unit, n_pins_total);
  arena.reserve(d_pin_xy, n_pins_total);
  arena.reserve(d_dpin, n_pins_total);
  arena.reserve(d_net_weight, n_net_total);
  arena.reserve(d_small, small_nets.size());
  arena.reserve(d_large, large_nets.size());
  arena.reserve(d_u_pstart, U + 1);
  arena.reserve(d_u_pins, u_pins.size());
  arena.reserve(d_wl_diag, U);
  arena.reserve(d_unit_du, U);
  arena.reserve(d_rho, FG);
  arena.reserve(d_phi, FG);
  arena.reserve(d_t0, FG);
  arena.reserve(d_t1, FG);
  arena.reserve(d_inv_lambda, G);
  const size_t dims[4] = {W, H, H, W};
  for (int j = 0; j < 4; ++j) {
    arena.reserve(d_b32[j], dims[j] * dims[j]);
    arena.reserve(d_bhi[j], dims[j] * dims[j]);
    arena.reserve(d_blo[j], dims[j] * dims[j]);
    arena.reserve(d_bthi[j], dims[j] * dims[j]);
    arena.reserve(d_btlo[j], dims[j] * dims[j]);
  }
  arena.reserve(d_lower, Cc);
  arena.reserve(d_upper, Cc);
  arena.reserve(d_metric, Cc);
  for (auto &v : d_vec) arena.reserve(v, Cc);
  for (auto &g : d_grad) arena.reserve(g, Cc);
  arena.reserve(d_gfu, (size_t)std::max(F, 1) * DU);
  // Partial-sum buffers sized from each launch's exact block count.
  auto blocks = [](size_t n, size_t per) { return std::max<size_t>(1, (n + per - 1) / per); };
  const size_t nb_wa = blocks(small_nets.size(), 256) + blocks(large_nets.size(), 4);
  size_t nb_energy = 1;
  {
    DctPlan probe{};
    probe.W = (int)W;
    probe.H = (int)H;
    probe.F = std::max(F, 1);
    for (int prec = 0; prec <= 2; ++prec) {
      probe.precision = prec;
      nb_energy = std::max(nb_energy, (size_t)poisson_energy_blocks(probe));
    }
  }
  arena.reserve(d_part_wa, nb_wa);
  arena.reserve(d_part_energy, nb_energy * MAXF);
  arena.reserve(d_part_gu, blocks(DU, 256) * 2 * MAXF);
  arena.reserve(d_part_gf, blocks(Fl, 256) * GF_STRIDE);
  arena.reserve(d_part_ug, blocks(U, 256) * UG_STRIDE);
  arena.reserve(d_part_step, blocks(Cc, 256) * 4);
  arena.reserve(d_part_metric, blocks(U + Fl, 256) * 3);
  arena.reserve(d_eval, 4);
  arena.reserve(d_step, 1);
  arena.reserve(d_metric_s, 1);
  arena.reserve(d_flag, 1);
  arena.reserve(d_best, 2 * U);
  arena.reserve(d_rudy, 2 * G);
  arena.reserve(d_rudy_fx, 2 * G);
  arena.reserve(d_rudy_err, 1);
  arena.reserve(d_raw_units, (size_t)std::max(F, 1) * DU * 2);
  arena.reserve(d_raw_fillers, 2 * Fl);
  arena.reserve(d_density_out, FG);
  arena.reserve(d_wa_grad, 2 * U);
  arena.reserve(d_eval_coords, Cc);
  arena.commit();

  auto up = [&](void *dst, const void *src, size_t bytes) {
    if (bytes) HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  };
  // fixed-point background: -smoothed capacity per field
  std::vector<long long> base(FG, 0);
  for (int f = 0; f < F; ++f) {
    std::vector<double> smooth(G, 0.0);
    auto it = capacity.find(field_kind[f]);
    if (it != capacity.end())
      for (size_t i = 0; i < G; ++i)
        if (it->second[i]) host_deposit(smooth, W, H, (double)(i % W), (double)(i / W), (double)it->second[i]);
    for (size_t i = 0; i < G; ++i) base[(size_t)f * G + i] = std::llrint(-smooth[i] * FX_SCALE);
  }
  up(d_base, base.data(), FG * sizeof(long long));
  up(d_m_unit, m_unit.data(), M * 4);
  up(d_m_off, m_off.data(), M * 16);
  up(d_m_field, m_field.data(), M);
  up(d_du_unit, du_unit.data(), DU * 4);
  up(d_du_mstart, du_mstart.data(), (DU + 1) * 4);
  up(d_fl_field, fl_field.data(), Fl);
  apply_charges(std::vector<double>(m_charge), std::vector<double>(fl_charge));
  up(d_net_start, net_start.data(), (n_net_total + 1) * 4);
  up(d_pin_unit, pin_unit.data(), n_pins_total * 4);
  up(d_pin_xy, pin_xy.data(), n_pins_total * 16);
  up(d_net_weight, net_weight.data(), n_net_total * 8);
  up(d_small, small_nets.data(), small_nets.size() * 4);
  up(d_large, large_nets.data(), large_nets.size() * 4);
  up(d_u_pstart, u_pstart.data(), (U + 1) * 4);
  up(d_u_pins, u_pins.data(), u_pins.size() * 4);
  up(d_wl_diag, wl_diag.data(), U * 8);
  up(d_unit_du, unit_du.data(), U * 4);
  up(d_lower, lower.data(), C * 8);
  up(d_upper, upper.data(), C * 8);
  std::vector<double> x0(C);
  for (size_t u = 0; u < U; ++u) {
    x0[2 * u] = initial_xy[2 * u];
    x0[2 * u + 1] = initial_xy[2 * u + 1];
  }
  for (size_t j = 0; j < Fl; ++j) {
    x0[2 * (U + j)] = fl_xy[2 * j];
    x0[2 * (U + j) + 1] = fl_xy[2 * j + 1];
  }
  for (size_t i = 0; i < C; ++i) x0[i] = clampd(x0[i], lower[i], upper[i]);
  up(d_vec[ix], x0.data(), C * 8);
  up(d_vec[ixp], x0.data(), C * 8);
  build_dct();

  dens.W = W;
  dens.H = H;
  dens.F = F;
  dens.n_members = (int)M;
  dens.n_fillers = (int)Fl;
  dens.n_units = (int)U;
  dens.n_density_units = (int)DU;
  dens.m_unit = d_m_unit;
  dens.m_off = d_m_off;
  dens.m_charge = d_m_charge;
  dens.m_field = d_m_field;
  dens.du_unit = d_du_unit;
  dens.du_mstart = d_du_mstart;
  dens.fl_field = d_fl_field;
  dens.fl_charge = d_fl_charge;
  dens.base_fx = d_base;
  dens.acc_m = d_acc_m;
  dens.acc_f = d_acc_f;
  wa.n_units = (int)U;
  wa.n_pins_total = (int)n_pins_total;
  wa.n_small = (int)small_nets.size();
  wa.n_large = (int)large_nets.size();
  wa.small_nets = d_small;
  wa.large_nets = d_large;
  wa.net_start = d_net_start;
  wa.net_weight = d_net_weight;
  wa.pin_unit = d_pin_unit;
  wa.pin_xy = d_pin_xy;
  wa.dpin = d_dpin;
  wa.u_pstart = d_u_pstart;
  wa.u_pins = d_u_pins;
  wa.first_arc_pin = (uint32_t)first_arc_pin;
  wa.unit_du = d_unit_du;
}

// Installs member and filler charges (apply_area_adjustment, eplace.rs:793-819).
void SessionImpl::apply_charges(const std::vector<double> &member_areas, const std::vector<double> &filler_areas) {
  m_charge = member_areas;
  fl_charge = filler_areas;
  std::vector<double> fcharge((size_t)std::max(F, 1) * DU, 0.0);
  for (size_t i = 0; i < DU; ++i)
    for (uint32_t m = du_mstart[i]; m < du_mstart[i + 1]; ++m) fcharge[(size_t)m_field[m] * DU + i] += m_charge[m];
  real_area.assign(F, 0.0);
  for (size_t m = 0; m < M; ++m) real_area[m_field[m]] += m_charge[m];
  if (M) HIP(hipMemcpy(d_m_charge, m_charge.data(), M * 8, hipMemcpyHostToDevice));
  if (Fl) HIP(hipMemcpy(d_fl_charge, fl_charge.data(), Fl * 8, hipMemcpyHostToDevice));
  if (DU) HIP(hipMemcpy(d_du_fcharge, fcharge.data(), fcharge.size() * 8, hipMemcpyHostToDevice));
}

// ===========================================================================
// One objective evaluation, fully on the device stream.
// ===========================================================================
void SessionImpl::evaluate(const double *coords, double gamma, const double *metric, double *grad, int slot,
                           bool parity, double *density_out) {
  evaluations += 1;
  if (opt.inject && evaluations == opt.inject) fail(TEXO_GP_ERR_INJECTED);
  EvalScalars *sc = &d_eval[slot];
  int nb_energy = 0, nb_gu = 0, nb_gf = 0, nb_wa = 0, nb_ug = 0;
  // WA depends only on the coordinates: fork it onto the second stream so it
  // overlaps the deposit / transform / gather chain, join before combining.
  HIP(hipEventRecord(ev_fork, stream));
  HIP(hipStreamWaitEvent(wa_stream, ev_fork, 0));
  launch_wa(wa, coords, gamma, d_part_wa, &nb_wa, wa_stream);
  HIP(hipEventRecord(ev_join, wa_stream));
  if (F == 0) {
    launch_zero_eval(sc, stream);
  } else {
    launch_deposit(dens, coords, sc, stream);
    launch_density_convert(dens, d_rho, density_out, sc, stream);
    launch_poisson(dct, d_rho, d_phi, d_part_energy, &nb_energy, stream);
  }
  GatherArgs g{};
  g.phi = d_phi;
  g.energy_partials = d_part_energy;
  g.nb_energy = nb_energy;
  for (int f = 0; f < F; ++f) {
    g.lambda[f] = started ? mult[f] : 0.0;
    g.scale[f] = started ? scales[f] : 1.0;
  }
  if (F > 0)
    launch_gather(dens, g, coords, d_gfu, parity ? d_raw_units : nullptr, parity ? d_raw_fillers : nullptr, grad,
                  metric, d_lower, d_upper, d_part_gu, d_part_gf, sc, &nb_gu, &nb_gf, stream);
  HIP(hipStreamWaitEvent(stream, ev_join, 0));
  launch_unit_gradient(wa, F, (int)DU, d_gfu, coords, metric, d_lower, d_upper, grad, parity ? d_wa_grad : nullptr,
                       d_part_ug, sc, &nb_ug, stream);
  FinalizeArgs fa{};
  fa.wa_partials = d_part_wa;
  fa.nb_wa = nb_wa;
  fa.energy_partials = d_part_energy;
  fa.nb_energy = nb_energy;
  fa.gu_partials = d_part_gu;
  fa.nb_gu = nb_gu;
  fa.gf_partials = d_part_gf;
  fa.nb_gf = nb_gf;
  fa.ug_partials = d_part_ug;
  fa.nb_ug = nb_ug;
  launch_finalize_eval(fa, sc, stream);
  HIP(hipGetLastError());
  fetch_eval(slot);
}

// ===========================================================================
// Schedules (eplace.rs:1111-1415).
// ===========================================================================
double SessionImpl::adaptive_gamma(const EvalScalars &s) const {
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

void SessionImpl::validate_initial(const EvalScalars &s) const {
  bool ok = std::isfinite(s.wa_value) && s.nonfinite == 0;
  for (int f = 0; f < F; ++f)
    ok &= std::isfinite(s.energy[f]) && s.energy[f] >= 0.0 && std::isfinite(overflow(s, f)) &&
          std::isfinite(s.force_l1[f]) && s.force_l1[f] >= 0.0;
  if (!ok) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
}

void SessionImpl::validate_fields(const EvalScalars &s) const {
  for (int f = 0; f < F; ++f)
    if (!std::isfinite(s.energy[f]) || !std::isfinite(scales[f]) || scales[f] <= 0.0)
      fail(TEXO_GP_ERR_DENSITY_FIELD_CHANGED);
}

std::vector<double> SessionImpl::subgradient(const EvalScalars &s) const {
  validate_fields(s);
  std::vector<double> v(F);
  for (int f = 0; f < F; ++f) {
    const double e = s.energy[f] / scales[f];
    v[f] = e + 0.5 * AUGMENTED_DENSITY_BETA * e * e;
    if (!(std::isfinite(v[f]) && v[f] >= 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  }
  return v;
}

bool SessionImpl::stop_reached(const EvalScalars &s, bool area_ready) const {
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

std::pair<double, double> SessionImpl::excess(const EvalScalars &s) const {
  double mx = 0.0, total = 0.0;
  for (int f = 0; f < F; ++f) {
    const double e = std::max(overflow(s, f) - overflow_target(field_kind[f]), 0.0);
    mx = std::max(mx, e);
    total += e;
  }
  return {mx, total};
}

}  // namespace texo

namespace texo {

// ===========================================================================
// Checkpoint, preconditioner, Nesterov step (global_placement.rs:121-335).
// ===========================================================================
void SessionImpl::consider_best(const EvalScalars &s) {
  // PlacementCheckpoint::consider (eplace.rs:107-120): lexicographic total_cmp
  const auto sc = excess(s);
  const bool better = !has_best || sc.first < best_max || (sc.first == best_max && sc.second < best_total);
  if (better) {
    launch_round_targets((int)U, d_vec[ix], (int)W, (int)H, d_best, stream);
    best_max = sc.first;
    best_total = sc.second;
    has_best = true;
  }
}

void SessionImpl::set_preconditioner() {
  double lam[MAXF] = {0.0, 0.0, 0.0, 0.0};
  for (int f = 0; f < F; ++f) {
    if (!std::isfinite(mult[f]) || mult[f] < 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    lam[f] = mult[f];
  }
  launch_preconditioner((int)U, (int)Fl, F, (int)DU, d_wl_diag, d_unit_du, d_du_fcharge, d_fl_field, d_fl_charge, lam,
                        d_metric, d_part_metric, d_metric_s, stream);
  HIP(hipMemcpyAsync(h_metric, d_metric_s, sizeof(MetricScalars), hipMemcpyDeviceToHost, stream));
}

namespace {
double half_nonzero(double v) {
  const double h = v * 0.5;
  return h == 0.0 ? v : h;
}
double rounding_margin(double a, double b) { return DBL_EPSILON * (1.0 + std::fabs(a) + std::fabs(b)); }
double bits_to_double(unsigned long long b) {
  double v;
  std::memcpy(&v, &b, 8);
  return v;
}
void check_evaluation(const EvalScalars &s, double value) {
  if (!std::isfinite(value)) fail(TEXO_GP_ERR_INVALID_OBJECTIVE);
  if (s.nonfinite) fail(TEXO_GP_ERR_NON_FINITE_GRADIENT);
}
}  // namespace

// One dynamic Nesterov step with the objective evaluated on the device.  The
// host mirrors the reference's control flow exactly; each decision reads FP64
// scalars reduced on the device.  y and every candidate are evaluated
// speculatively before the decision that needs them (evaluation is
// deterministic, so an evaluation of y == origin reproduces the origin's bits)
// which leaves one stream synchronisation per line-search trial.
bool SessionImpl::nesterov_step(double gamma, HostEval &payload, bool &accepted) {
  if (!std::isfinite(momentum) || momentum < 1.0) fail(TEXO_GP_ERR_INVALID_MOMENTUM);
  if (has_lipschitz && (!std::isfinite(lipschitz) || lipschitz <= 0.0)) fail(TEXO_GP_ERR_INVALID_LIPSCHITZ);
  double *X = d_vec[ix];
  evaluate(X, gamma, d_metric, d_grad[0], 0, false);
  bool origin_checked = false;
  HostEval origin;
  auto check_origin = [&]() {
    if (origin_checked) return;
    if (h_metric->nonfinite) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    origin.s = h_eval[0];
    origin.value = combined_value(origin.s);
    check_evaluation(origin.s, origin.value);
    origin_checked = true;
  };
  bool restart = false;
  bool has_wl = has_lipschitz;
  double wl = lipschitz;
  for (;;) {
    const double mom = restart ? 1.0 : momentum;
    const double next = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * mom * mom));
    if (!std::isfinite(next)) fail(TEXO_GP_ERR_ITERATION_OVERFLOW);
    const double beta = (mom - 1.0) / next;
    bool y_is_origin = true;
    const double *Yp = X;
    const double *GY = d_grad[0];
    HostEval ye;
    if (!restart) {
      launch_zero_step(d_step, d_flag, stream);
      launch_extrapolate(C, X, d_vec[ixp], d_lower, d_upper, beta, d_vec[iy], d_flag, stream);
      HIP(hipMemcpyAsync(h_flag, d_flag, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
      evaluate(d_vec[iy], gamma, d_metric, d_grad[1], 1, false);
      sync();
      check_origin();
      y_is_origin = *h_flag == 0;
      if (y_is_origin) {
        evaluations -= 1;  // the reference reuses the origin evaluation
      } else {
        Yp = d_vec[iy];
        GY = d_grad[1];
        ye.s = h_eval[1];
        ye.value = combined_value(ye.s);
        check_evaluation(ye.s, ye.value);
      }
    } else {
      sync();
      check_origin();
    }
    if (y_is_origin) ye = origin;
    const double y_stat = ye.kkt();
    if (!std::isfinite(y_stat)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
    auto stationary = [&](bool set_l, double l) {
      // state.coordinates = origin; previous = origin; momentum = 1
      HIP(hipMemcpyAsync(d_vec[ixp], X, C * sizeof(double), hipMemcpyDeviceToDevice, stream));
      momentum = 1.0;
      if (set_l) {
        has_lipschitz = true;
        lipschitz = l;
      }
      payload = origin;
      accepted = false;
      return true;
    };
    if (y_stat == 0.0) {
      if (!y_is_origin) {
        restart = true;
        continue;
      }
      return stationary(false, 0.0);
    }
    const double mmax = bits_to_double(h_metric->max_bits), mmin = bits_to_double(h_metric->min_bits);
    double L = half_nonzero(has_wl ? wl : y_stat);
    bool may_reduce = true;
    bool restart_outer = false;
    for (;;) {
      trials += 1;
      if (!std::isfinite(L * mmax) || !std::isfinite(L * mmin)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
      if (L * mmin == 0.0) {
        L *= 2.0;
        may_reduce = false;
        if (!std::isfinite(L)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
        continue;
      }
      launch_zero_step(d_step, nullptr, stream);
      launch_candidate(C, Yp, GY, d_metric, L, d_lower, d_upper, X, d_vec[ic], d_part_step, d_step, stream);
      HIP(hipMemcpyAsync(h_step, d_step, sizeof(StepScalars), hipMemcpyDeviceToHost, stream));
      evaluate(d_vec[ic], gamma, d_metric, d_grad[2], 2, false);
      sync();
      if (!h_step->changed) {
        evaluations -= 1;  // candidate == y: the reference does not evaluate it
        if (may_reduce) {
          const double smaller = half_nonzero(L);
          if (smaller < L) {
            L = smaller;
            continue;
          }
        }
        if (!y_is_origin) {
          restart = true;
          has_wl = true;
          wl = L;
          restart_outer = true;
          break;
        }
        return stationary(true, L);
      }
      HostEval ce;
      ce.s = h_eval[2];
      ce.value = combined_value(ce.s);
      check_evaluation(ce.s, ce.value);
      const double upper_bound = ye.value + h_step->gdelta + 0.5 * L * h_step->msq;
      if (!std::isfinite(upper_bound)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
      const bool smooth = ce.value <= upper_bound + rounding_margin(ce.value, upper_bound);
      const bool monotone = ce.value <= origin.value + rounding_margin(ce.value, origin.value);
      if (smooth && monotone) {
        const double stat = ce.kkt();
        if (!std::isfinite(stat)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
        // previous <- origin, coordinates <- candidate
        const int old_prev = ixp;
        ixp = ix;
        ix = ic;
        ic = old_prev;
        momentum = next;
        has_lipschitz = true;
        lipschitz = L;
        payload = ce;
        accepted = true;
        if (opt.trace)
          std::fprintf(stderr,
                       "TEXO_PNR_METRICS eplace iteration=%llu objective=%.9e stationarity=%.9e "
                       "coordinate_change=%.9e gamma=%.9e wa=%.9e\n",
                       (unsigned long long)iterations, ce.value, stat, bits_to_double(h_step->change_bits), gamma,
                       ce.s.wa_value);
        return true;
      }
      if (smooth && !restart) {
        restart = true;
        has_wl = true;
        wl = L;
        restart_outer = true;
        break;
      }
      L *= 2.0;
      may_reduce = false;
      if (!std::isfinite(L)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
    }
    if (restart_outer) continue;
  }
}

// One iteration of the hot loop (eplace.rs:512-573) without the stop tests.
void SessionImpl::iteration(bool check_stationary, bool &stop_stationary) {
  stop_stationary = false;
  const double gamma = adaptive_gamma(current.s);
  iterations += 1;
  set_preconditioner();
  HostEval payload;
  bool accepted = false;
  nesterov_step(gamma, payload, accepted);
  current = payload;
  last_wa = current.s.wa_value;
  validate_fields(current.s);
  consider_best(current.s);
  if (check_stationary) {
    // stationary_target_is_fixed (eplace.rs:973-985)
    if (accepted) {
      has_prev_stationary = false;
    } else {
      std::vector<uint32_t> t(2 * U);
      std::vector<double> x(2 * U);
      HIP(hipMemcpyAsync(x.data(), d_vec[ix], 2 * U * sizeof(double), hipMemcpyDeviceToHost, stream));
      sync();
      for (size_t i = 0; i < 2 * U; ++i) t[i] = rounded_coordinate(x[i], (i & 1) ? H : W);
      stop_stationary = has_prev_stationary && prev_stationary == t;
      prev_stationary = t;
      has_prev_stationary = true;
    }
  }
}

// optimize_density_once set-up (eplace.rs:452-499).
void SessionImpl::initialise_round(bool reset_after_area, const std::vector<double> *fixed_scales) {
  started = false;
  mult.assign(F, 0.0);
  scales.assign(F, 1.0);
  evaluate(d_vec[ix], opt.gamma_base, nullptr, d_grad[0], 0, false);
  sync();
  const EvalScalars probe = h_eval[0];
  validate_initial(probe);
  const double g0 = adaptive_gamma(probe);
  evaluate(d_vec[ix], g0, nullptr, d_grad[0], 0, false);
  sync();
  const EvalScalars init = h_eval[0];
  validate_initial(init);
  if (fixed_scales) {
    scales = *fixed_scales;
  } else {
    for (int f = 0; f < F; ++f) scales[f] = std::max(init.energy[f], 1.0);
  }
  validate_fields(init);
  // new DynamicNesterovState: previous = coordinates, momentum 1, no L
  HIP(hipMemcpyAsync(d_vec[ixp], d_vec[ix], C * sizeof(double), hipMemcpyDeviceToDevice, stream));
  momentum = 1.0;
  has_lipschitz = false;
  lipschitz = 0.0;
  has_best = false;
  current.s = init;
  consider_best(init);
  has_prev_stationary = false;
  if (reset_after_area) {
    // area_adjusted_density_multipliers / _step (eplace.rs:1183-1232)
    const std::vector<double> s = subgradient(init);
    const double wf = init.wa_l1;
    double df = 0.0;
    for (int f = 0; f < F; ++f) df += init.force_l1[f] * s[f];
    if (!std::isfinite(wf) || wf <= 0.0 || !std::isfinite(df) || df <= 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    const double scale = AREA_RESET_DENSITY_WEIGHT * wf / df;
    bool any = false;
    double norm = 0.0;
    for (int f = 0; f < F; ++f) {
      mult[f] = scale * s[f];
      if (!(std::isfinite(mult[f]) && mult[f] >= 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
      any |= mult[f] > 0.0;
      norm += mult[f] * mult[f];
    }
    if (!any) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    mstep = (MULTIPLIER_ALPHA_HIGH - 1.0) * std::sqrt(norm);
    if (!(std::isfinite(mstep) && mstep > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  } else {
    // initial_density_multiplier (eplace.rs:1126-1148)
    const double wf = init.wa_l1;
    double df = 0.0;
    for (int f = 0; f < F; ++f) df += init.raw_l1[f];
    const double m0 = (wf == 0.0 || df == 0.0) ? 1.0 : INITIAL_DENSITY_WEIGHT * wf / df;
    if (!(std::isfinite(m0) && m0 > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    mult.assign(F, m0);
    mstep = MULTIPLIER_ALPHA_HIGH - 1.0;
  }
  started = true;
}

// update_density_multipliers + next_multiplier_step (eplace.rs:1234-1298)
static void update_multipliers(SessionImpl &S) {
  const std::vector<double> s = S.subgradient(S.current.s);
  double n = 0.0;
  for (double v : s) n += v * v;
  n = std::sqrt(n);
  if (n != 0.0) {
    if (!std::isfinite(n) || !std::isfinite(S.mstep) || S.mstep <= 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    for (int f = 0; f < S.F; ++f) {
      S.mult[f] += S.mstep * s[f] / n;
      if (!std::isfinite(S.mult[f]) || S.mult[f] < 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    }
  }
  S.validate_fields(S.current.s);
  double en = 0.0;
  for (int f = 0; f < S.F; ++f) {
    const double e = S.current.s.energy[f] / S.scales[f];
    en += e * e;
  }
  const double lg = std::log(AUGMENTED_DENSITY_BETA * std::sqrt(en) + 1.0);
  if (!std::isfinite(lg) || lg < 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  const double growth = MULTIPLIER_ALPHA_LOW + (1.0 - 1.0 / (1.0 + lg)) * (MULTIPLIER_ALPHA_HIGH - MULTIPLIER_ALPHA_LOW);
  if (!(std::isfinite(growth) && growth >= MULTIPLIER_ALPHA_LOW && growth <= MULTIPLIER_ALPHA_HIGH))
    fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  const double next = S.mstep * growth;
  if (!(std::isfinite(next) && next > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  S.mstep = next;
}

// ===========================================================================
// Area rounds (eplace.rs:584-819, instance_area.rs, register_clustering.rs).
// ===========================================================================
size_t SessionImpl::area_round_limit() const {
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

namespace {

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

struct AreaMember {
  uint32_t kind;
  double current, routability, clustering;
};

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

}  // namespace

// propose_area_adjustment (eplace.rs:627-729): RUDY on the device, the
// sequential charge bookkeeping on the host.
void SessionImpl::area_round_metrics(double &total_relative_increase, bool &requires_reset,
                                     std::vector<double> &member_areas, std::vector<double> &filler_areas) {
  HIP(hipMemsetAsync(d_rudy_err, 0, sizeof(unsigned int), stream));
  launch_rudy((int)n_nets, d_net_start, d_pin_unit, d_pin_xy, d_vec[ix], (int)W, (int)H, d_rudy_fx, d_rudy,
              d_rudy_err, stream);
  std::vector<double> demand(2 * G), x(2 * U);
  unsigned int err = 0;
  HIP(hipMemcpyAsync(demand.data(), d_rudy, 2 * G * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIP(hipMemcpyAsync(x.data(), d_vec[ix], 2 * U * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIP(hipMemcpyAsync(&err, d_rudy_err, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
  sync();
  if (err) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  std::map<uint32_t, double> clustering;
  if (!controls.empty()) {
    std::vector<std::pair<uint32_t, std::pair<double, double>>> regs;
    for (size_t i = 0; i < DU; ++i) {
      const uint32_t u = du_unit[i];
      for (uint32_t m = du_mstart[i]; m < du_mstart[i + 1]; ++m)
        if (field_kind[m_field[m]] == TEXO_KIND_REGISTER)
          regs.push_back({m_cell[m], {x[2 * u] + m_off[2 * m], x[2 * u + 1] + m_off[2 * m + 1]}});
    }
    std::vector<double> areas;
    if (!register_clustering(regs, controls, movable_instance_count, areas)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    for (size_t i = 0; i < regs.size(); ++i) clustering[regs[i].first] = areas[i];
  }
  auto utilization = [&](const double *d, const std::vector<uint32_t> &cap, uint32_t px, uint32_t py) {
    const size_t idx = (size_t)py * W + px;
    if (cap[idx] == 0) return d[idx] == 0.0 ? 0.0 : kSqrt2;
    const double u = d[idx] / (double)cap[idx];
    if (!(std::isfinite(u) && u >= 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    return u;
  };
  std::vector<AreaMember> members(M);
  for (size_t i = 0; i < DU; ++i) {
    const uint32_t u = du_unit[i];
    for (uint32_t m = du_mstart[i]; m < du_mstart[i + 1]; ++m) {
      const uint32_t px = rounded_coordinate(x[2 * u] + m_off[2 * m], W);
      const uint32_t py = rounded_coordinate(x[2 * u + 1] + m_off[2 * m + 1], H);
      const double h = std::min(utilization(demand.data(), routing_h, px, py), kSqrt2);
      const double v = std::min(utilization(demand.data() + G, routing_v, px, py), kSqrt2);
      const uint32_t kind = field_kind[m_field[m]];
      // routability_optimized_area (instance_area.rs:64-84) for LUT/FF only
      double area = m_charge[m];
      if (is_lut(kind) || kind == TEXO_KIND_REGISTER) {
        const double uu = std::max(h, v);
        area = m_charge[m] * std::min(uu * uu, 2.0);
      }
      auto it = clustering.find(m_cell[m]);
      members[m] = {kind, m_charge[m], area, it == clustering.end() ? 0.0 : it->second};
    }
  }
  std::vector<std::pair<uint32_t, double>> fillers(Fl);
  for (size_t j = 0; j < Fl; ++j) fillers[j] = {field_kind[fl_field[j]], fl_charge[j]};
  requires_reset = adjust_areas(members, fillers, member_areas, filler_areas);
  std::map<uint32_t, double> before, inc;
  for (size_t m = 0; m < M; ++m) {
    before[members[m].kind] += members[m].current;
    inc[members[m].kind] += member_areas[m] - members[m].current;
  }
  double bt = 0.0, it = 0.0;
  for (auto &kv : before) bt += kv.second;
  for (auto &kv : inc) it += kv.second;
  total_relative_increase = it / bt;
  if (!std::isfinite(total_relative_increase)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
}

double SessionImpl::hpwl_of(const std::vector<double> &x) const {
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

void SessionImpl::fill_report(texo_gp_report *r, const EvalScalars *fields) {
  if (!r) return;
  r->iterations = iterations;
  r->evaluations = evaluations;
  r->line_search_trials = trials;
  r->area_rounds = area_rounds;
  r->dct_precision = opt.precision;
  r->fixed_iterations = opt.fixed_iterations;
  r->gamma_base = opt.gamma_base;
  r->final_wirelength = last_wa;
  std::vector<double> x(2 * U);
  if (U) {
    if (empty) {
      x = initial_xy;
    } else {
      HIP(hipMemcpyAsync(x.data(), d_vec[ix], 2 * U * sizeof(double), hipMemcpyDeviceToHost, stream));
      sync();
    }
  }
  r->final_hpwl = hpwl_of(x);
  if (fields) {
    r->n_fields = (uint32_t)F;
    for (int f = 0; f < F; ++f) {
      r->field_kind[f] = field_kind[f];
      r->field_overflow[f] = overflow(*fields, f);
      r->field_energy[f] = fields->energy[f];
    }
  }
}

// place() (eplace.rs:353-420) around optimize_density_once (eplace.rs:439-582).
int SessionImpl::run(uint32_t *targets, texo_gp_report *report) {
  const auto t0 = std::chrono::steady_clock::now();
  if (empty) {
    for (size_t i = 0; i < 2 * U; ++i) targets[i] = rounded_coordinate(initial_xy[i], (i & 1) ? H : W);
    fill_report(report, nullptr);
    return 0;
  }
  const bool fixed_mode = opt.fixed_iterations > 0;
  const bool adjust_area = has_routing && !fixed_mode;
  const size_t limit = area_round_limit();
  bool has_prev_change = false;
  double prev_change = 0.0;
  std::vector<double> fixed_scales;
  bool has_scales = false, reset = false;
  size_t round = 0;
  for (;;) {
    const bool area_ready = adjust_area && !(has_prev_change && prev_change < 0.01);
    initialise_round(reset, has_scales ? &fixed_scales : nullptr);
    if (!has_scales) {
      fixed_scales = scales;
      has_scales = true;
    }
    bool exhausted = false;
    if (fixed_mode || !stop_reached(current.s, area_ready)) {
      uint64_t local = 0;
      for (;;) {
        if (fixed_mode) {
          if (local >= opt.fixed_iterations) break;
        } else {
          if (stop_reached(current.s, area_ready)) break;
          if (iterations >= opt.max_iterations) {
            if (!best_effort) {
              Failure f;
              f.status = TEXO_GP_ERR_DID_NOT_CONVERGE;
              f.index = (int64_t)iterations;
              f.v0 = F > 0 ? overflow(current.s, 0) : 0.0;
              f.v1 = F > 1 ? overflow(current.s, 1) : 0.0;
              throw f;
            }
            exhausted = true;
            break;
          }
        }
        bool stationary_fixed = false;
        iteration(!fixed_mode, stationary_fixed);
        local += 1;
        if (!fixed_mode) {
          if (stop_reached(current.s, area_ready)) break;
          if (stationary_fixed) {
            if (!best_effort) {
              Failure f;
              f.status = TEXO_GP_ERR_DID_NOT_CONVERGE;
              f.index = (int64_t)iterations;
              throw f;
            }
            exhausted = true;
            break;
          }
        }
        update_multipliers(*this);
      }
    }
    if (!area_ready || exhausted) {
      HIP(hipMemcpyAsync(targets, d_best, 2 * U * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
      sync();
      fill_report(report, &current.s);
      if (report)
        report->elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      return 0;
    }
    double total_inc = 0.0;
    bool requires_reset = false;
    std::vector<double> ma, fa;
    area_round_metrics(total_inc, requires_reset, ma, fa);
    if (!requires_reset) {
      has_prev_change = true;
      prev_change = 0.0;
      reset = true;
      continue;
    }
    if (round >= limit) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    apply_charges(ma, fa);
    has_prev_change = true;
    prev_change = total_inc;
    reset = true;
    round += 1;
    area_rounds += 1;
  }
}

}  // namespace texo

// ===========================================================================
// C ABI
// ===========================================================================
using texo::Failure;
using texo::SessionImpl;

struct texo_gp_session {
  SessionImpl impl;
};

namespace {

void put_failure(texo_gp_report *r, const Failure &f) {
  if (!r) return;
  r->status = f.status;
  r->err_index = f.index;
  r->err_index2 = f.index2;
  r->err_kind = f.kind;
  r->err_x = f.x;
  r->err_y = f.y;
  r->err_value[0] = f.v0;
  r->err_value[1] = f.v1;
}

struct TimerScope {
  explicit TimerScope(texo_gp_session *s) { texo::g_timer = s ? &s->impl.timer : nullptr; }
  ~TimerScope() { texo::g_timer = nullptr; }
};

int set_device(const texo_gp_session *s) {
  return hipSetDevice(s->impl.device) == hipSuccess ? 0 : TEXO_GP_ERR_DEVICE;
}

}  // namespace

extern "C" {

const char *texo_gp_last_error(void) { return texo::g_last_error.c_str(); }

texo_gp_session *texo_gp_session_create(const texo_gp_problem *problem, const texo_gp_options *options,
                                        texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  std::unique_ptr<texo_gp_session> s;
  try {
    s.reset(new texo_gp_session());
    s->impl.load(problem, options);
    if (s->impl.empty) {
      s->impl.C = 2 * s->impl.U;
      s->impl.lower.resize(s->impl.C);
      s->impl.upper.resize(s->impl.C);
      for (size_t u = 0; u < s->impl.U; ++u) {
        s->impl.lower[2 * u] = problem->unit_bounds[4 * u];
        s->impl.lower[2 * u + 1] = problem->unit_bounds[4 * u + 1];
        s->impl.upper[2 * u] = problem->unit_bounds[4 * u + 2];
        s->impl.upper[2 * u + 1] = problem->unit_bounds[4 * u + 3];
      }
    }
    s->impl.upload();
    return s.release();
  } catch (const Failure &f) {
    put_failure(report, f);
    return nullptr;
  } catch (const std::bad_alloc &) {
    if (report) report->status = TEXO_GP_ERR_OUT_OF_MEMORY;
    return nullptr;
  }
}

void texo_gp_session_destroy(texo_gp_session *s) {
  if (!s) return;
  set_device(s);
  delete s;
}

int texo_gp_session_sizes(const texo_gp_session *s, uint64_t *coordinates, uint32_t *fields, uint64_t *fillers,
                          uint64_t *density_units) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  if (coordinates) *coordinates = s->impl.C;
  if (fields) *fields = (uint32_t)s->impl.F;
  if (fillers) *fillers = s->impl.Fl;
  if (density_units) *density_units = s->impl.DU;
  return 0;
}

int texo_gp_session_initial(const texo_gp_session *s, double *coordinates, double *lower, double *upper,
                            uint32_t *field_kind, uint32_t *filler_field, double *filler_charge) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const SessionImpl &S = s->impl;
  if (coordinates) {
    for (size_t u = 0; u < S.U; ++u) {
      coordinates[2 * u] = texo::clampd(S.initial_xy[2 * u], S.lower[2 * u], S.upper[2 * u]);
      coordinates[2 * u + 1] = texo::clampd(S.initial_xy[2 * u + 1], S.lower[2 * u + 1], S.upper[2 * u + 1]);
    }
    for (size_t j = 0; j < S.Fl; ++j) {
      coordinates[2 * (S.U + j)] = S.fl_xy[2 * j];
      coordinates[2 * (S.U + j) + 1] = S.fl_xy[2 * j + 1];
    }
  }
  if (lower) std::copy(S.lower.begin(), S.lower.end(), lower);
  if (upper) std::copy(S.upper.begin(), S.upper.end(), upper);
  if (field_kind) std::copy(S.field_kind.begin(), S.field_kind.end(), field_kind);
  for (size_t j = 0; j < S.Fl; ++j) {
    if (filler_field) filler_field[j] = S.fl_field[j];
    if (filler_charge) filler_charge[j] = S.fl_charge[j];
  }
  return 0;
}

int texo_gp_session_evaluate(texo_gp_session *s, const double *coordinates, double gamma, double *wa_value,
                             double *wa_gradient, texo_gp_field_eval *fields, double *unit_gradient,
                             double *filler_gradient, double *density, double *potential) {
  if (!s || !coordinates) return TEXO_GP_ERR_INVALID_ARGUMENT;
  TimerScope ts(s);
  SessionImpl &S = s->impl;
  try {
    if (set_device(s)) return TEXO_GP_ERR_DEVICE;
    if (!(std::isfinite(gamma) && gamma > 0.0)) return TEXO_GP_ERR_INVALID_ARGUMENT;
    // validate_units_with_positions / validate_fillers_with_positions
    for (size_t i = 0; i < S.DU; ++i) {
      const uint32_t u = S.du_unit[i];
      if (!std::isfinite(coordinates[2 * u]) || !std::isfinite(coordinates[2 * u + 1]))
        texo::fail(TEXO_GP_ERR_NON_FINITE_UNIT_COORDINATE, (int64_t)i, 0);
    }
    for (size_t j = 0; j < S.Fl; ++j)
      if (!std::isfinite(coordinates[2 * (S.U + j)]) || !std::isfinite(coordinates[2 * (S.U + j) + 1]))
        texo::fail(TEXO_GP_ERR_NON_FINITE_FILLER, (int64_t)j);
    for (size_t u = 0; u < S.U; ++u)
      if (!std::isfinite(coordinates[2 * u]) || !std::isfinite(coordinates[2 * u + 1]))
        return TEXO_GP_ERR_INVALID_ARGUMENT;
    const bool saved = S.started;
    S.started = false;
    HIP(hipMemcpyAsync(S.d_eval_coords, coordinates, S.C * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const uint64_t evals = S.evaluations;
    S.evaluate(S.d_eval_coords, gamma, nullptr, S.d_grad[0], 3, true, density ? S.d_density_out : nullptr);
    S.evaluations = evals;
    S.started = saved;
    const size_t FG = (size_t)S.F * S.G;
    std::vector<float> phi(potential ? FG : 0);
    if (wa_gradient && S.U)
      HIP(hipMemcpyAsync(wa_gradient, S.d_wa_grad, 2 * S.U * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    if (unit_gradient && S.F && S.DU)
      HIP(hipMemcpyAsync(unit_gradient, S.d_raw_units, 2 * (size_t)S.F * S.DU * sizeof(double), hipMemcpyDeviceToHost,
                         S.stream));
    if (filler_gradient && S.Fl)
      HIP(hipMemcpyAsync(filler_gradient, S.d_raw_fillers, 2 * S.Fl * sizeof(double), hipMemcpyDeviceToHost,
                         S.stream));
    if (density && FG)
      HIP(hipMemcpyAsync(density, S.d_density_out, FG * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    if (potential && FG) HIP(hipMemcpyAsync(phi.data(), S.d_phi, FG * sizeof(float), hipMemcpyDeviceToHost, S.stream));
    S.sync();
    const texo::EvalScalars &e = S.h_eval[3];
    if (wa_value) *wa_value = e.wa_value;
    for (int f = 0; f < S.F; ++f) {
      if (fields) {
        fields[f].energy = e.energy[f];
        fields[f].normalized_overflow = S.overflow(e, f);
        fields[f].net_charge = static_cast<double>(e.net_fx[f]) * texo::FX_INV;
        fields[f].force_l1 = e.force_l1[f];
        fields[f].real_area = S.real_area[f];
        fields[f].filler_charge = (double)S.available[f] - S.real_area[f];
      }
    }
    for (size_t i = 0; i < FG && potential; ++i) potential[i] = static_cast<double>(phi[i]);
    return 0;
  } catch (const Failure &f) {
    return f.status;
  }
}

int texo_gp_session_poisson(texo_gp_session *s, const double *density, double *potential, double *energy) {
  if (!s || !density) return TEXO_GP_ERR_INVALID_ARGUMENT;
  TimerScope ts(s);
  SessionImpl &S = s->impl;
  try {
    if (set_device(s) || S.F == 0) return TEXO_GP_ERR_INVALID_ARGUMENT;
    const size_t FG = (size_t)S.F * S.G;
    std::vector<float> rho(FG), phi(FG);
    for (size_t i = 0; i < FG; ++i) {
      if (!std::isfinite(density[i])) return TEXO_GP_ERR_NON_FINITE_DENSITY;
      rho[i] = static_cast<float>(density[i]);
    }
    HIP(hipMemcpyAsync(S.d_rho, rho.data(), FG * sizeof(float), hipMemcpyHostToDevice, S.stream));
    int nb = 0;
    texo::launch_poisson(S.dct, S.d_rho, S.d_phi, S.d_part_energy, &nb, S.stream);
    std::vector<double> parts((size_t)nb * texo::MAXF);
    HIP(hipMemcpyAsync(parts.data(), S.d_part_energy, parts.size() * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP(hipMemcpyAsync(phi.data(), S.d_phi, FG * sizeof(float), hipMemcpyDeviceToHost, S.stream));
    S.sync();
    for (size_t i = 0; i < FG; ++i) potential[i] = phi[i];
    for (int f = 0; f < S.F; ++f) {
      double e = 0.0;
      for (int b = 0; b < nb; ++b) e += parts[(size_t)b * texo::MAXF + f];
      energy[f] = e;
    }
    return 0;
  } catch (const Failure &f) {
    return f.status;
  }
}

int texo_gp_session_run(texo_gp_session *s, uint32_t *out_targets_xy, texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!s || !out_targets_xy) return TEXO_GP_ERR_INVALID_ARGUMENT;
  TimerScope ts(s);
  try {
    if (set_device(s)) return TEXO_GP_ERR_DEVICE;
    return s->impl.run(out_targets_xy, report);
  } catch (const Failure &f) {
    put_failure(report, f);
    if (report) {
      report->iterations = s->impl.iterations;
      report->evaluations = s->impl.evaluations;
    }
    return f.status;
  }
}

int texo_gp_session_iterate(texo_gp_session *s, uint32_t iterations, texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  TimerScope ts(s);
  SessionImpl &S = s->impl;
  try {
    if (set_device(s)) return TEXO_GP_ERR_DEVICE;
    if (S.empty) return TEXO_GP_ERR_INVALID_ARGUMENT;
    if (!S.started) S.initialise_round(false, nullptr);
    for (uint32_t i = 0; i < iterations; ++i) {
      bool stationary = false;
      S.iteration(false, stationary);
      texo::update_multipliers(S);
    }
    S.sync();
    if (report) {
      S.fill_report(report, &S.current.s);
    }
    return 0;
  } catch (const Failure &f) {
    put_failure(report, f);
    return f.status;
  }
}

int texo_gp_session_coordinates(const texo_gp_session *s, double *coordinates) {
  if (!s || !coordinates) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const SessionImpl &S = s->impl;
  if (S.empty) {
    std::copy(S.initial_xy.begin(), S.initial_xy.end(), coordinates);
    return 0;
  }
  if (hipSetDevice(S.device) != hipSuccess) return TEXO_GP_ERR_DEVICE;
  if (hipMemcpyAsync(coordinates, S.d_vec[S.ix], S.C * sizeof(double), hipMemcpyDeviceToHost, S.stream) != hipSuccess ||
      hipStreamSynchronize(S.stream) != hipSuccess)
    return TEXO_GP_ERR_DEVICE;
  return 0;
}

int texo_gp_session_rudy(texo_gp_session *s, const double *coordinates, double *horizontal, double *vertical) {
  if (!s || !coordinates || !horizontal || !vertical) return TEXO_GP_ERR_INVALID_ARGUMENT;
  TimerScope ts(s);
  SessionImpl &S = s->impl;
  try {
    if (set_device(s)) return TEXO_GP_ERR_DEVICE;
    HIP(hipMemcpyAsync(S.d_eval_coords, coordinates, S.C * sizeof(double), hipMemcpyHostToDevice, S.stream));
    HIP(hipMemsetAsync(S.d_rudy_err, 0, sizeof(unsigned int), S.stream));
    texo::launch_rudy((int)S.n_nets, S.d_net_start, S.d_pin_unit, S.d_pin_xy, S.d_eval_coords, (int)S.W, (int)S.H,
                      S.d_rudy_fx, S.d_rudy, S.d_rudy_err, S.stream);
    unsigned int err = 0;
    HIP(hipMemcpyAsync(horizontal, S.d_rudy, S.G * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP(hipMemcpyAsync(vertical, S.d_rudy + S.G, S.G * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP(hipMemcpyAsync(&err, S.d_rudy_err, sizeof(unsigned int), hipMemcpyDeviceToHost, S.stream));
    S.sync();
    return err ? TEXO_GP_ERR_INVALID_NORMALIZATION : 0;
  } catch (const Failure &f) {
    return f.status;
  }
}

int texo_gp_session_set_kernel_timing(texo_gp_session *s, const char *kernel, int enable) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  s->impl.timer.name = kernel ? kernel : "";
  s->impl.timer.enabled = enable != 0;
  s->impl.timer.used = 0;
  return 0;
}

int texo_gp_session_kernel_timing(texo_gp_session *s, double *avg_ms, uint64_t *count) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  SessionImpl &S = s->impl;
  if (set_device(s)) return TEXO_GP_ERR_DEVICE;
  if (hipStreamSynchronize(S.stream) != hipSuccess) return TEXO_GP_ERR_DEVICE;
  double total = 0.0;
  for (size_t i = 0; i < S.timer.used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, S.timer.events[i].first, S.timer.events[i].second) != hipSuccess)
      return TEXO_GP_ERR_DEVICE;
    total += ms;
  }
  if (avg_ms) *avg_ms = S.timer.used ? total / (double)S.timer.used : 0.0;
  if (count) *count = S.timer.used;
  S.timer.used = 0;
  return 0;
}

int texo_gp_place(const texo_gp_problem *problem, const texo_gp_options *options, uint32_t *out_targets_xy,
                  texo_gp_report *report) {
  texo_gp_session *s = texo_gp_session_create(problem, options, report);
  if (!s) return report ? report->status : TEXO_GP_ERR_INVALID_ARGUMENT;
  const int rc = texo_gp_session_run(s, out_targets_xy, report);
  texo_gp_session_destroy(s);
  return rc;
}

}  // extern "C"
