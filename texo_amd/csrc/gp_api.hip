// gp_api.hip — the C ABI of include/texo_gp.h: one-call placement
// (eplace::place), evaluate-level sessions used by the parity tests, and the
// continuous-batching server.

#include "gp_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

using texo::Engine;
using texo::Failure;
using texo::Job;
using texo::Phase;

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

// Final report of a job (status, counts, field overflows; HPWL of `x` when
// given).
void fill_report_at(Job &j, const std::vector<double> *x, texo_gp_report *r) {
  if (!r) return;
  r->status = j.phase == Phase::Failed ? j.status : 0;
  if (j.phase == Phase::Failed) put_failure(r, j.failure);
  r->iterations = j.iterations;
  r->evaluations = j.evaluations;
  r->line_search_trials = j.trials;
  r->area_rounds = j.area_rounds;
  r->dct_precision = j.opt.precision;
  r->fixed_iterations = j.opt.fixed_iterations;
  r->gamma_base = j.opt.gamma_base;
  r->final_wirelength = j.last_wa;
  if (x) r->final_hpwl = j.hpwl(*x);
  r->n_fields = (uint32_t)j.F;
  for (int f = 0; f < j.F && f < TEXO_GP_MAX_FIELDS; ++f) {
    r->field_kind[f] = j.field_kind[f];
    r->field_overflow[f] = j.started ? j.overflow(j.current, f) : 0.0;
    r->field_energy[f] = j.current.energy[f];
  }
}

void fill_report(Engine &e, Job &j, texo_gp_report *r) {
  if (!r) return;
  if (j.U && !j.empty && j.arena.base()) {
    // the final HPWL reduced on the device (FP64) from the current coordinates
    const hipStream_t st = e.stream();
    double h = 0.0;
    j.enqueue_hpwl(j.d_vec[j.ix], &h, st);
    HIP(hipStreamSynchronize(st));
    fill_report_at(j, nullptr, r);
    r->final_hpwl = h;
    return;
  }
  fill_report_at(j, &j.initial_xy, r);
}

// Targets of a finished job: the best checkpoint (eplace.rs:553), or the
// rounded inputs when there is no density unit (eplace.rs:303-310).
void copy_targets(Engine &e, Job &j, uint32_t *out) {
  if (!out) return;
  if (j.empty) {
    for (size_t i = 0; i < 2 * j.U; ++i)
      out[i] = texo::rounded_coordinate(j.initial_xy[i], (i & 1) ? j.H : j.W);
    return;
  }
  const hipStream_t st = e.stream();
  HIP(hipMemcpyAsync(out, j.d_best, 2 * j.U * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP(hipStreamSynchronize(st));
}

void drive(Engine &e, Job &j) {
  std::vector<Job *> one{&j};
  while (j.phase != Phase::Done && j.phase != Phase::Failed && j.phase != Phase::Paused) e.round(one);
}

int timing_set(Engine &e, const char *kernel, int enable) {
  e.timer.name = kernel ? kernel : "";
  e.timer.enabled = enable != 0;
  e.timer.used = 0;
  e.timer.tasks = 0.0;
  return 0;
}

// Average duration of the named kernel's launches since the last call, and
// the average number of evaluations one launch covered.
int timing_get(Engine &e, double *avg_ms, uint64_t *count, double *avg_tasks) {
  texo::KernelTimer &t = e.timer;
  if (hipSetDevice(e.device()) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return TEXO_GP_ERR_DEVICE;
  double total = 0.0;
  for (size_t i = 0; i < t.used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.events[i].first, t.events[i].second) != hipSuccess) return TEXO_GP_ERR_DEVICE;
    total += ms;
  }
  if (avg_ms) *avg_ms = t.used ? total / (double)t.used : 0.0;
  if (count) *count = t.used;
  if (avg_tasks) *avg_tasks = t.used ? t.tasks / (double)t.used : 0.0;
  t.used = 0;
  t.tasks = 0.0;
  return 0;
}

}  // namespace

struct texo_gp_session {
  std::unique_ptr<Engine> engine;
  Job job;
  ~texo_gp_session() {
    if (engine) {
      (void)hipSetDevice(engine->device());
      engine->detach(&job);
    }
  }
};

extern "C" {

const char *texo_gp_last_error(void) { return texo::g_last_error.c_str(); }

texo_gp_session *texo_gp_session_create(const texo_gp_problem *problem, const texo_gp_options *options,
                                        texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  try {
    std::unique_ptr<texo_gp_session> s(new texo_gp_session());
    s->job.load(problem, options);
    s->engine.reset(new Engine(s->job.opt.device, 4));
    s->engine->attach(&s->job);
    return s.release();
  } catch (...) {
    put_failure(report, texo::current_failure());
    return nullptr;
  }
}

void texo_gp_session_destroy(texo_gp_session *s) { delete s; }

int texo_gp_session_sizes(const texo_gp_session *s, uint64_t *coordinates, uint32_t *fields, uint64_t *fillers,
                          uint64_t *density_units) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  if (coordinates) *coordinates = s->job.C;
  if (fields) *fields = (uint32_t)s->job.F;
  if (fillers) *fillers = s->job.Fl;
  if (density_units) *density_units = s->job.DU;
  return 0;
}

int texo_gp_session_initial(const texo_gp_session *s, double *coordinates, double *lower, double *upper,
                            uint32_t *field_kind, uint32_t *filler_field, double *filler_charge) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const Job &J = s->job;
  if (coordinates) {
    for (size_t u = 0; u < J.U; ++u) {
      coordinates[2 * u] = texo::clampd(J.initial_xy[2 * u], J.lower[2 * u], J.upper[2 * u]);
      coordinates[2 * u + 1] = texo::clampd(J.initial_xy[2 * u + 1], J.lower[2 * u + 1], J.upper[2 * u + 1]);
    }
    for (size_t k = 0; k < J.Fl; ++k) {
      coordinates[2 * (J.U + k)] = J.fl_xy[2 * k];
      coordinates[2 * (J.U + k) + 1] = J.fl_xy[2 * k + 1];
    }
  }
  if (lower) std::copy(J.lower.begin(), J.lower.end(), lower);
  if (upper) std::copy(J.upper.begin(), J.upper.end(), upper);
  if (field_kind) std::copy(J.field_kind.begin(), J.field_kind.end(), field_kind);
  for (size_t k = 0; k < J.Fl; ++k) {
    if (filler_field) filler_field[k] = J.fl_field[k];
    if (filler_charge) filler_charge[k] = J.fl_charge[k];
  }
  return 0;
}

int texo_gp_session_evaluate(texo_gp_session *s, const double *coordinates, double gamma, double *wa_value,
                             double *wa_gradient, texo_gp_field_eval *fields, double *unit_gradient,
                             double *filler_gradient, double *density, double *potential) {
  if (!s || !coordinates) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  try {
    if (!(std::isfinite(gamma) && gamma > 0.0)) return TEXO_GP_ERR_INVALID_ARGUMENT;
    // validate_units_with_positions / validate_fillers_with_positions
    // (electrostatic_placement.rs:485-582)
    for (size_t i = 0; i < J.DU; ++i) {
      const uint32_t u = J.du_unit[i];
      if (!std::isfinite(coordinates[2 * u]) || !std::isfinite(coordinates[2 * u + 1]))
        texo::fail(TEXO_GP_ERR_NON_FINITE_UNIT_COORDINATE, (int64_t)i, 0);
    }
    for (size_t k = 0; k < J.Fl; ++k)
      if (!std::isfinite(coordinates[2 * (J.U + k)]) || !std::isfinite(coordinates[2 * (J.U + k) + 1]))
        texo::fail(TEXO_GP_ERR_NON_FINITE_FILLER, (int64_t)k);
    for (size_t u = 0; u < J.U; ++u)
      if (!std::isfinite(coordinates[2 * u]) || !std::isfinite(coordinates[2 * u + 1]))
        return TEXO_GP_ERR_INVALID_ARGUMENT;
    Engine &E = *s->engine;
    E.evaluate_once(&J, coordinates, gamma, density != nullptr);
    const size_t FG = (size_t)J.F * J.G;
    std::vector<float> phi(potential ? FG : 0);
    hipStream_t st = E.stream();
    if (wa_gradient && J.U)
      HIP(hipMemcpyAsync(wa_gradient, J.d_wa_grad, 2 * J.U * sizeof(double), hipMemcpyDeviceToHost, st));
    if (unit_gradient && J.F && J.DU)
      HIP(hipMemcpyAsync(unit_gradient, J.d_raw_units, 2 * (size_t)J.F * J.DU * sizeof(double), hipMemcpyDeviceToHost,
                         st));
    if (filler_gradient && J.Fl)
      HIP(hipMemcpyAsync(filler_gradient, J.d_raw_fillers, 2 * J.Fl * sizeof(double), hipMemcpyDeviceToHost, st));
    if (density && FG) HIP(hipMemcpyAsync(density, J.d_density_out, FG * sizeof(double), hipMemcpyDeviceToHost, st));
    if (potential && FG) HIP(hipMemcpyAsync(phi.data(), J.slot[0].phi, FG * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
    const texo::EvalScalars &e = J.h_scalars->eval[2];
    if (wa_value) *wa_value = e.wa_value;
    for (int f = 0; f < J.F && fields; ++f) {
      fields[f].energy = e.energy[f];
      fields[f].normalized_overflow = J.overflow(e, f);
      fields[f].net_charge = static_cast<double>(e.net_fx[f]) * texo::FX_INV;
      fields[f].force_l1 = e.force_l1[f];
      fields[f].real_area = J.real_area[f];
      fields[f].filler_charge = (double)J.available[f] - J.real_area[f];
    }
    for (size_t i = 0; i < FG && potential; ++i) potential[i] = static_cast<double>(phi[i]);
    return 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    return f.status;
  }
}

int texo_gp_session_poisson(texo_gp_session *s, const double *density, double *potential, double *energy) {
  if (!s || !density || !potential || !energy) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  try {
    if (J.F == 0) return TEXO_GP_ERR_INVALID_ARGUMENT;
    Engine &E = *s->engine;
    hipStream_t st = E.stream();
    const size_t FG = (size_t)J.F * J.G;
    std::vector<float> rho(FG), phi(FG);
    for (size_t i = 0; i < FG; ++i) {
      if (!std::isfinite(density[i])) return TEXO_GP_ERR_NON_FINITE_DENSITY;
      rho[i] = static_cast<float>(density[i]);
    }
    if (J.p_slot == 0) J.p_valid = false;
    HIP(hipMemcpyAsync(J.slot[0].rho, rho.data(), FG * sizeof(float), hipMemcpyHostToDevice, st));
    // input range of the FP16-scaled transform, as the k_convert blocks of a
    // full evaluation record it: max |rho| per block of BLOCK cells and field
    std::vector<unsigned int> amax((size_t)J.nb_conv * texo::MAXF, 0u);
    for (size_t i = 0; i < FG; ++i) {
      const float a = std::fabs(rho[i]);
      unsigned int bits;
      std::memcpy(&bits, &a, 4);
      unsigned int &m = amax[(i / texo::BLOCK) * texo::MAXF + i / J.G];
      m = std::max(m, bits);
    }
    HIP(hipMemcpyAsync(J.slot[0].amax_part, amax.data(), amax.size() * 4, hipMemcpyHostToDevice, st));
    HIP(hipStreamSynchronize(st));
    // transform only: no WA runs
    texo::Task t = J.task(0, J.d_eval_coords, J.d_grad[0], nullptr, 1.0, false, &J.d_scalars->eval[2], false);
    // one task: its transform tiles as four one-entry block plans
    const int F = J.F, W = (int)J.W, H = (int)J.H;
    texo::DctLaunch d{};
    d.precision = (int)J.opt.precision;
    d.fast = texo::poisson_fast_shape(W, H);
    d.wide_ok = d.fast && texo::poisson_wide_ok(d.precision, W, H);
    d.kmax = std::max(W, H);
    d.n_tasks = 1;
    const int tiles[4] = {((F * H + 31) / 32) * ((W + 31) / 32), ((F * W + 31) / 32) * ((H + 31) / 32),
                          d.wide_ok ? (F * H / 32) * (W / 128) : 0, d.wide_ok ? (F * W / 32) * (H / 128) : 0};
    const int starts[8] = {0, tiles[0], 0, tiles[1], 0, tiles[2], 0, tiles[3]};
    texo::Task *d_task = nullptr;
    int *d_starts = nullptr;
    HIP(hipMalloc(&d_task, sizeof(texo::Task)));
    HIP(hipMalloc(&d_starts, sizeof(starts)));
    HIP(hipMemcpy(d_task, &t, sizeof(t), hipMemcpyHostToDevice));
    HIP(hipMemcpy(d_starts, starts, sizeof(starts), hipMemcpyHostToDevice));
    texo::BlockPlan *plans[4] = {&d.h, &d.w, &d.h_wide, &d.w_wide};
    for (int k = 0; k < 4; ++k) {
      plans[k]->start = d_starts + 2 * k;
      plans[k]->n_tasks = 1;
      plans[k]->total = tiles[k];
      plans[k]->per = tiles[k] >= 2 * texo::XCDS ? (tiles[k] + texo::XCDS - 1) / texo::XCDS : 0;
    }
    texo::launch_poisson(d_task, d, st);
    std::vector<double> parts((size_t)J.nb_energy * texo::MAXF);
    HIP(hipMemcpyAsync(parts.data(), J.slot[0].part_energy, parts.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipMemcpyAsync(phi.data(), J.slot[0].phi, FG * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
    HIP(hipFree(d_task));
    HIP(hipFree(d_starts));
    for (size_t i = 0; i < FG; ++i) potential[i] = phi[i];
    for (int f = 0; f < J.F; ++f) {
      double e = 0.0;
      for (int b = 0; b < J.nb_energy; ++b) e += parts[(size_t)b * texo::MAXF + f];
      energy[f] = e;
    }
    return 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    return f.status;
  }
}

int texo_gp_session_run(texo_gp_session *s, uint32_t *out_targets_xy, texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!s || !out_targets_xy) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const auto t0 = std::chrono::steady_clock::now();
  try {
    Job &J = s->job;
    J.begin_run(false);
    drive(*s->engine, J);
    fill_report(*s->engine, J, report);
    if (J.phase == Phase::Failed) return J.status;
    copy_targets(*s->engine, J, out_targets_xy);
    if (report) report->elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    put_failure(report, f);
    return f.status;
  }
}

int texo_gp_session_iterate(texo_gp_session *s, uint32_t iterations, texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    Job &J = s->job;
    if (J.empty) return TEXO_GP_ERR_INVALID_ARGUMENT;
    if (!J.started) J.begin_run(true);
    J.pause_at = J.iterations + iterations;
    if (J.phase == Phase::Paused) J.phase = Phase::IterStart;
    drive(*s->engine, J);
    fill_report(*s->engine, J, report);
    return J.phase == Phase::Failed ? J.status : 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    put_failure(report, f);
    return f.status;
  }
}

int texo_gp_session_coordinates(const texo_gp_session *s, double *coordinates) {
  if (!s || !coordinates) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const Job &J = s->job;
  if (J.empty) {
    std::copy(J.initial_xy.begin(), J.initial_xy.end(), coordinates);
    return 0;
  }
  try {
    HIP(hipSetDevice(s->engine->device()));
    const hipStream_t st = s->engine->stream();
    HIP(hipMemcpyAsync(coordinates, J.d_vec[J.ix], J.C * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
  } catch (...) {
    return TEXO_GP_ERR_DEVICE;
  }
  return 0;
}

int texo_gp_session_rudy(texo_gp_session *s, const double *coordinates, double *horizontal, double *vertical) {
  if (!s || !coordinates || !horizontal || !vertical) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  try {
    hipStream_t st = s->engine->stream();
    HIP(hipMemcpyAsync(J.d_eval_coords, coordinates, J.C * sizeof(double), hipMemcpyHostToDevice, st));
    HIP(hipMemsetAsync(J.d_rudy_err, 0, sizeof(unsigned int), st));
    texo::g_timer = &s->engine->timer;
    texo::launch_rudy((int)J.n_nets, J.d_net_start, J.d_pin_unit, J.js.pin_xy, J.d_eval_coords, (int)J.W, (int)J.H,
                      J.d_rudy_fx, J.d_rudy, J.d_rudy_err, st);
    texo::g_timer = nullptr;
    unsigned int err = 0;
    HIP(hipMemcpyAsync(horizontal, J.d_rudy, J.G * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipMemcpyAsync(vertical, J.d_rudy + J.G, J.G * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipMemcpyAsync(&err, J.d_rudy_err, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
    return err ? TEXO_GP_ERR_INVALID_NORMALIZATION : 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    return f.status;
  }
}

int texo_gp_session_set_kernel_timing(texo_gp_session *s, const char *kernel, int enable) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  return timing_set(*s->engine, kernel, enable);
}

int texo_gp_session_kernel_timing(texo_gp_session *s, double *avg_ms, uint64_t *count, double *avg_tasks) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  return timing_get(*s->engine, avg_ms, count, avg_tasks);
}

int texo_gp_place(const texo_gp_problem *problem, const texo_gp_options *options, uint32_t *out_targets_xy,
                  texo_gp_report *report) {
  // a caller without a report still gets the typed status
  texo_gp_report local;
  texo_gp_report *r = report ? report : &local;
  texo_gp_session *s = texo_gp_session_create(problem, options, r);
  if (!s) return r->status ? r->status : TEXO_GP_ERR_INVALID_ARGUMENT;
  const int rc = texo_gp_session_run(s, out_targets_xy, r);
  texo_gp_session_destroy(s);
  return rc;
}

int texo_gp_session_combined(texo_gp_session *s, const double *coordinates, double gamma, const double *multipliers,
                             const double *scales, double *value, double *gradient) {
  if (!s || !coordinates || !multipliers || !scales || !value || !gradient) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  if (J.empty || !(std::isfinite(gamma) && gamma > 0.0)) return TEXO_GP_ERR_INVALID_ARGUMENT;
  for (int f = 0; f < J.F; ++f)
    if (!(std::isfinite(multipliers[f]) && multipliers[f] >= 0.0) || !(std::isfinite(scales[f]) && scales[f] > 0.0))
      return TEXO_GP_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < J.C; ++i)
    if (!std::isfinite(coordinates[i])) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const std::vector<double> mult = J.mult, scl = J.scales;
  int rc = 0;
  try {
    J.mult.assign(multipliers, multipliers + J.F);
    J.scales.assign(scales, scales + J.F);
    Engine &E = *s->engine;
    E.evaluate_once(&J, coordinates, gamma, false);
    *value = J.combined_value(J.h_scalars->eval[2]);
    const hipStream_t st = E.stream();
    HIP(hipMemcpyAsync(gradient, J.d_grad[0], J.C * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
  } catch (...) {
    rc = texo::current_failure().status;
  }
  J.mult = mult;
  J.scales = scl;
  return rc;
}

int texo_gp_session_preconditioner(texo_gp_session *s, const double *multipliers, double *diagonal) {
  if (!s || !multipliers || !diagonal) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  if (J.empty) return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    // the coefficient check of diagonal_preconditioner (eplace.rs:1060-1063)
    std::vector<double> lambda(multipliers, multipliers + J.F);
    for (double l : lambda)
      if (!std::isfinite(l) || l < 0.0) return TEXO_GP_ERR_INVALID_NORMALIZATION;
    Engine &E = *s->engine;
    if (!E.precondition_once(&J, lambda)) return TEXO_GP_ERR_INVALID_NORMALIZATION;
    const hipStream_t st = E.stream();
    HIP(hipMemcpyAsync(diagonal, J.d_metric, J.C * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
    return 0;
  } catch (...) {
    return texo::current_failure().status;
  }
}

int texo_gp_session_round_targets(texo_gp_session *s, const double *coordinates, uint32_t *out_targets_xy) {
  if (!s || !coordinates || !out_targets_xy) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  if (J.U == 0) return 0;
  if (J.empty) {
    for (size_t i = 0; i < 2 * J.U; ++i)
      out_targets_xy[i] = texo::rounded_coordinate(coordinates[i], (i & 1) ? J.H : J.W);
    return 0;
  }
  uint32_t *d_out = nullptr;
  int rc = 0;
  try {
    Engine &E = *s->engine;
    const hipStream_t st = E.stream();
    HIP(hipMalloc(&d_out, 2 * J.U * sizeof(uint32_t)));
    HIP(hipMemcpyAsync(J.d_eval_coords, coordinates, 2 * J.U * sizeof(double), hipMemcpyHostToDevice, st));
    texo::launch_round_targets_to(J.d_eval_coords, d_out, J.U, J.W, J.H, st);
    HIP(hipGetLastError());
    HIP(hipMemcpyAsync(out_targets_xy, d_out, 2 * J.U * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
  } catch (...) {
    rc = texo::current_failure().status;
  }
  if (d_out) (void)hipFree(d_out);
  return rc;
}

int texo_gp_session_area_proposal(texo_gp_session *s, const double *coordinates, double *member_areas,
                                  double *filler_areas, int32_t *requires_reset, double *total_relative_increase) {
  if (!s || !coordinates || !member_areas || !filler_areas) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &J = s->job;
  if (J.empty || !J.has_routing) return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    Engine &E = *s->engine;
    const hipStream_t st = E.stream();
    HIP(hipMemcpyAsync(J.d_eval_coords, coordinates, J.C * sizeof(double), hipMemcpyHostToDevice, st));
    const Engine::AreaProposal p = E.propose_areas(&J, J.d_eval_coords, st);
    std::copy(p.member_areas.begin(), p.member_areas.end(), member_areas);
    std::copy(p.filler_areas.begin(), p.filler_areas.end(), filler_areas);
    if (requires_reset) *requires_reset = p.requires_reset ? 1 : 0;
    if (total_relative_increase) *total_relative_increase = p.total_relative_increase;
    return 0;
  } catch (...) {
    return texo::current_failure().status;
  }
}

int texo_gp_session_schedule(const texo_gp_session *s, texo_gp_schedule_state *out) {
  if (!s || !out) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const Job &J = s->job;
  std::memset(out, 0, sizeof(*out));
  out->n_fields = (uint32_t)std::min(J.F, TEXO_GP_MAX_FIELDS);
  for (uint32_t f = 0; f < out->n_fields; ++f) {
    out->multiplier[f] = f < J.mult.size() ? J.mult[f] : 0.0;
    out->initial_energy[f] = f < J.scales.size() ? J.scales[f] : 0.0;
    out->gamma_weight[f] = f < J.gamma_weight.size() ? J.gamma_weight[f] : 0.0;
  }
  out->multiplier_step = J.mstep;
  out->gamma = J.gamma;
  out->momentum = J.momentum;
  out->lipschitz = J.has_lipschitz ? J.lipschitz : 0.0;
  return 0;
}

}  // extern "C"

// ===========================================================================
// Lockstep batches
// ===========================================================================
struct texo_gp_batch {
  std::unique_ptr<Engine> engine;
  std::vector<std::unique_ptr<Job>> jobs;
  ~texo_gp_batch() {
    if (!engine) return;
    (void)hipSetDevice(engine->device());
    for (auto &j : jobs) engine->detach(j.get());
    jobs.clear();
  }
  // Rounds until every job is paused, done or failed.  When the jobs are large
  // enough they are dealt into ROUND_CONTEXTS groups whose rounds overlap on
  // the GPU while the host plans the next group's; each job's arithmetic is
  // unchanged.
  void drive_all() {
    size_t charges = 0;
    for (auto &j : jobs) charges += j->M + j->Fl;
    const bool split = jobs.size() > 1 && charges >= texo::PIPELINE_MIN_CHARGES;
    constexpr int NC = texo::ROUND_CONTEXTS;
    std::vector<Job *> group[NC];
    for (size_t k = 0; k < jobs.size(); ++k) group[split ? k % NC : 0].push_back(jobs[k].get());
    bool live[NC];
    for (int c = 0; c < NC; ++c) live[c] = engine->launch(c, group[c]);
    for (;;) {
      bool any = false;
      for (int c = 0; c < NC; ++c) any = any || live[c];
      if (!any) break;
      for (int c = 0; c < NC; ++c) {
        if (!live[c]) continue;
        engine->finish(c);
        live[c] = engine->launch(c, group[c]);
      }
    }
  }
};

extern "C" {

texo_gp_batch *texo_gp_batch_create(int32_t device) {
  try {
    std::unique_ptr<texo_gp_batch> b(new texo_gp_batch());
    b->engine.reset(new Engine(device, 1024));
    return b.release();
  } catch (...) {
    return nullptr;
  }
}

int64_t texo_gp_batch_add(texo_gp_batch *b, const texo_gp_problem *problem, const texo_gp_options *options,
                          texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!b) return -(int64_t)TEXO_GP_ERR_INVALID_ARGUMENT;
  std::unique_ptr<Job> j(new Job());
  try {
    j->load(problem, options, b->engine->image_pool());
    if (j->opt.device != b->engine->device()) texo::fail(TEXO_GP_ERR_INVALID_ARGUMENT);
    b->engine->attach(j.get());
  } catch (...) {
    const Failure f = texo::current_failure();
    put_failure(report, f);
    if (j->pool_index >= 0) b->engine->detach(j.get());
    return -(int64_t)f.status;
  }
  j->id = (int64_t)b->jobs.size();
  b->jobs.push_back(std::move(j));
  return (int64_t)b->jobs.size() - 1;
}

int texo_gp_batch_run(texo_gp_batch *b) {
  if (!b) return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    for (auto &j : b->jobs)
      if (j->phase != Phase::Failed) j->begin_run(false);
    b->drive_all();
    return 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    return f.status;
  }
}

int texo_gp_batch_iterate(texo_gp_batch *b, uint32_t iterations) {
  if (!b) return TEXO_GP_ERR_INVALID_ARGUMENT;
  try {
    for (auto &j : b->jobs) {
      if (j->empty || j->phase == Phase::Failed || j->phase == Phase::Done) continue;
      if (!j->started) j->begin_run(true);
      j->pause_at = j->iterations + iterations;
      if (j->phase == Phase::Paused) j->phase = Phase::IterStart;
    }
    b->drive_all();
    return 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    return f.status;
  }
}

int texo_gp_batch_result(texo_gp_batch *b, uint64_t job, uint32_t *out_targets_xy, texo_gp_report *report) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!b || job >= b->jobs.size()) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Job &j = *b->jobs[job];
  try {
    fill_report(*b->engine, j, report);
    if (j.phase == Phase::Failed) return j.status;
    if (out_targets_xy) {
      if (j.phase != Phase::Done) return TEXO_GP_ERR_NOT_READY;
      copy_targets(*b->engine, j, out_targets_xy);
    }
    return 0;
  } catch (...) {
    const Failure f = texo::current_failure();
    put_failure(report, f);
    return f.status;
  }
}

int texo_gp_batch_sizes(const texo_gp_batch *b, uint64_t job, uint64_t *coordinates, uint32_t *fields,
                        uint64_t *fillers, uint64_t *density_units) {
  if (!b || job >= b->jobs.size()) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const Job &J = *b->jobs[job];
  if (coordinates) *coordinates = J.C;
  if (fields) *fields = (uint32_t)J.F;
  if (fillers) *fillers = J.Fl;
  if (density_units) *density_units = J.DU;
  return 0;
}

int texo_gp_batch_coordinates(texo_gp_batch *b, uint64_t job, double *coordinates) {
  if (!b || job >= b->jobs.size() || !coordinates) return TEXO_GP_ERR_INVALID_ARGUMENT;
  const Job &J = *b->jobs[job];
  if (J.empty) {
    std::copy(J.initial_xy.begin(), J.initial_xy.end(), coordinates);
    return 0;
  }
  try {
    HIP(hipSetDevice(b->engine->device()));
    const hipStream_t st = b->engine->stream();
    HIP(hipMemcpyAsync(coordinates, J.d_vec[J.ix], J.C * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP(hipStreamSynchronize(st));
  } catch (...) {
    return TEXO_GP_ERR_DEVICE;
  }
  return 0;
}

int texo_gp_batch_set_kernel_timing(texo_gp_batch *b, const char *kernel, int enable) {
  if (!b) return TEXO_GP_ERR_INVALID_ARGUMENT;
  return timing_set(*b->engine, kernel, enable);
}

int texo_gp_batch_kernel_timing(texo_gp_batch *b, double *avg_ms, uint64_t *count, double *avg_tasks) {
  if (!b) return TEXO_GP_ERR_INVALID_ARGUMENT;
  return timing_get(*b->engine, avg_ms, count, avg_tasks);
}

void texo_gp_batch_destroy(texo_gp_batch *b) { delete b; }

}  // extern "C"

// ===========================================================================
// Continuous-batching server
// ===========================================================================
struct texo_gp_server {
  struct Entry {
    std::unique_ptr<Job> job;  // host part kept until the result is taken
    std::chrono::steady_clock::time_point submitted;
    bool done = false;
    bool cancel = false;  // texo_gp_server_cancel on a running job
    int status = 0;
    std::vector<uint32_t> targets;
    std::vector<double> x;  // final coordinates, for the report's HPWL
    texo_gp_report report{};
    void *readback = nullptr;  // pinned landing block of the result copies
    size_t readback_cap = 0;
    double hpwl = 0.0;  // final HPWL reduced on the device, when has_hpwl
    bool has_hpwl = false;
  };
  // Result copies of the jobs one retire() found finished, enqueued on the
  // compute stream after their last work; `ev` marks their landing.
  struct Readback {
    hipEvent_t ev;
    std::vector<int64_t> ids;
  };
  std::unique_ptr<Engine> engine;
  uint64_t budget = 0;
  uint32_t max_batch = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> pending;
  std::unordered_map<int64_t, std::unique_ptr<Entry>> entries;
  std::vector<int64_t> active;
  int64_t next_id = 1;
  bool stopping = false;
  uint64_t hbm_in_use = 0;
  size_t charges_in_flight = 0;
  texo_gp_server_stats stats{};
  std::thread worker;
  std::deque<Readback> readbacks;  // server thread only
  std::vector<hipEvent_t> spare_events;

  void loop();
  void retire(int c, std::vector<int64_t> &group);
  void collect(bool block);
  ~texo_gp_server() {
    for (auto &r : readbacks)
      if (r.ev) (void)hipEventDestroy(r.ev);
    for (hipEvent_t ev : spare_events) (void)hipEventDestroy(ev);
  }
};

void texo_gp_server::loop() {
  (void)hipSetDevice(engine->device());
  // Active jobs are held in ROUND_CONTEXTS groups whose rounds are pipelined: the
  // host finishes and re-plans one group while the GPU runs the others'.
  constexpr int NC = texo::ROUND_CONTEXTS;
  std::vector<int64_t> group[NC];
  bool live[NC] = {};
  for (;;) {
    collect(false);
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stopping || !pending.empty() || !active.empty(); });
      if (stopping) {
        lk.unlock();
        for (int c = 0; c < NC; ++c)
          if (live[c]) {
            try {
              engine->finish(c);
            } catch (...) {
            }
          }
        collect(true);
        return;
      }
      // admission in submission order, by HBM footprint and batch size
      while (!pending.empty() && active.size() < max_batch) {
        Entry &e = *entries[pending.front()];
        const uint64_t need = e.job->device_bytes();
        if (!active.empty() && hbm_in_use + need > budget) break;
        const int64_t id = pending.front();
        pending.pop_front();
        lk.unlock();
        const auto t0 = std::chrono::steady_clock::now();
        try {
          engine->attach(e.job.get());
          e.job->begin_run(false);
        } catch (...) {
          const Failure f = texo::current_failure();
          e.job->failure = f;
          e.job->status = f.status;
          e.job->phase = Phase::Failed;
        }
        lk.lock();
        stats.admit_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        active.push_back(id);
        charges_in_flight += e.job->M + e.job->Fl;
        // the groups' rounds overlap on the GPU: once there is enough work to
        // split, a new job joins the group with the fewest charges
        size_t in_group[NC] = {};
        int least = 0;
        for (int c = 0; c < NC; ++c) {
          for (int64_t g : group[c]) in_group[c] += entries[g]->job->M + entries[g]->job->Fl;
          if (in_group[c] < in_group[least]) least = c;
        }
        const bool split = charges_in_flight >= texo::PIPELINE_MIN_CHARGES;
        group[split ? least : 0].push_back(id);
        hbm_in_use += need;
        stats.admitted += 1;
      }
    }
    for (int c = 0; c < NC; ++c) {
      if (live[c]) {
        const auto t0 = std::chrono::steady_clock::now();
        try {
          engine->finish(c);
        } catch (...) {
          const Failure f = texo::current_failure();
          for (int64_t id : group[c]) {
            Job &j = *entries[id]->job;
            if (j.phase != Phase::Done && j.phase != Phase::Failed) {
              j.failure = f;
              j.status = f.status;
              j.phase = Phase::Failed;
            }
          }
        }
        live[c] = false;
        std::lock_guard<std::mutex> lk(mu);
        stats.rounds += 1;
        stats.busy_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats.wait_seconds += engine->last_wait_seconds();
      }
      retire(c, group[c]);
      std::vector<Job *> jobs;
      bool cancelled = false;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (int64_t id : group[c]) {
          Entry &e = *entries[id];
          Job *j = e.job.get();
          // a cancelled job stops at this round boundary: its previous
          // round has finished, so its memory can go back to the pools
          if (e.cancel && j->phase != Phase::Done && j->phase != Phase::Failed) {
            j->failure = Failure{TEXO_GP_ERR_CANCELLED};
            j->status = TEXO_GP_ERR_CANCELLED;
            j->phase = Phase::Failed;
            cancelled = true;
            continue;
          }
          jobs.push_back(j);
        }
      }
      if (cancelled) retire(c, group[c]);
      const auto t1 = std::chrono::steady_clock::now();
      try {
        live[c] = engine->launch(c, jobs);
      } catch (...) {
        const Failure f = texo::current_failure();
        for (Job *j : jobs) {
          if (j->phase != Phase::Done && j->phase != Phase::Failed) {
            j->failure = f;
            j->status = f.status;
            j->phase = Phase::Failed;
          }
        }
        live[c] = false;
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        stats.launch_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
      }
      if (!live[c]) retire(c, group[c]);
    }
    // nothing in flight to overlap with: wait for the pending results
    bool any_live = false;
    for (int c = 0; c < NC; ++c) any_live = any_live || live[c];
    if (!any_live) collect(true);
  }
}

// Takes the finished jobs out of `group` and enqueues their result copies
// on the compute stream, after the work the last round's advance left there
// for them; collect() completes them once the copies have landed, so the
// host never waits here for the round in flight.
void texo_gp_server::retire(int c, std::vector<int64_t> &group) {
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int64_t> still;
  Readback rb{nullptr, {}};
  hipStream_t st = nullptr;
  for (int64_t id : group) {
    Entry *ep;
    {
      std::lock_guard<std::mutex> lk(mu);
      ep = entries[id].get();
    }
    Entry &e = *ep;
    Job &j = *e.job;
    if (j.phase != Phase::Done && j.phase != Phase::Failed) {
      still.push_back(id);
      continue;
    }
    if (!j.empty && j.arena.base() && j.U) {
      try {
        if (!st) st = engine->stream(c);
        e.readback = engine->image_pool()->acquire(j.U * 24 + sizeof(double), &e.readback_cap);
        char *dst = static_cast<char *>(e.readback);
        HIP(hipMemcpyAsync(dst, j.d_vec[j.ix], j.U * 16, hipMemcpyDeviceToHost, st));
        if (j.phase == Phase::Done) HIP(hipMemcpyAsync(dst + j.U * 16, j.d_best, j.U * 8, hipMemcpyDeviceToHost, st));
        // the report's HPWL, reduced on the device from the same coordinates
        j.enqueue_hpwl(j.d_vec[j.ix], reinterpret_cast<double *>(dst + j.U * 24), st);
      } catch (...) {
        const Failure f = texo::current_failure();
        j.failure = f;
        j.status = f.status;
        j.phase = Phase::Failed;
      }
    }
    rb.ids.push_back(id);
  }
  group.swap(still);
  if (!rb.ids.empty()) {
    hipEvent_t ev = nullptr;
    if (!spare_events.empty()) {
      ev = spare_events.back();
      spare_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
    }
    if (ev && hipEventRecord(ev, engine->stream(c)) != hipSuccess) {
      (void)hipEventDestroy(ev);
      ev = nullptr;
    }
    if (!ev) (void)hipStreamSynchronize(engine->stream(c));  // no marker: land them now
    rb.ev = ev;
    readbacks.push_back(std::move(rb));
  }
  std::lock_guard<std::mutex> lk(mu);
  stats.retire_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Completes the retired jobs whose result copies have landed (all of them
// when `block`): results into the entries, device memory back to the pools.
// The HPWL of the report and the release of the job's host data happen on
// the thread that takes the result.
void texo_gp_server::collect(bool block) {
  const auto t0 = std::chrono::steady_clock::now();
  bool any = false;
  while (!readbacks.empty()) {
    Readback &rb = readbacks.front();
    const hipError_t q = !rb.ev ? hipSuccess : block ? hipEventSynchronize(rb.ev) : hipEventQuery(rb.ev);
    if (q == hipErrorNotReady) break;
    for (int64_t id : rb.ids) {
      Entry *ep;
      {
        std::lock_guard<std::mutex> lk(mu);
        ep = entries[id].get();
      }
      Entry &e = *ep;
      Job &j = *e.job;
      e.targets.assign(2 * j.U, 0);
      if (q != hipSuccess && j.phase == Phase::Done) {
        j.status = TEXO_GP_ERR_DEVICE;
        j.phase = Phase::Failed;
      }
      if (j.empty || !e.readback) {
        e.x = j.initial_xy;
        if (j.empty && j.phase == Phase::Done)
          for (size_t i = 0; i < 2 * j.U; ++i)
            e.targets[i] = texo::rounded_coordinate(j.initial_xy[i], (i & 1) ? j.H : j.W);
      } else {
        const char *src = static_cast<const char *>(e.readback);
        e.x.resize(2 * j.U);
        std::memcpy(e.x.data(), src, j.U * 16);
        if (j.phase == Phase::Done) std::memcpy(e.targets.data(), src + j.U * 16, j.U * 8);
        std::memcpy(&e.hpwl, src + j.U * 24, sizeof(double));
        e.has_hpwl = true;
        engine->image_pool()->release(e.readback, e.readback_cap);
        e.readback = nullptr;
      }
      fill_report_at(j, nullptr, &e.report);
      if (e.has_hpwl) e.report.final_hpwl = e.hpwl;
      e.status = j.phase == Phase::Failed ? j.status : 0;
      e.report.status = e.status;
      e.report.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - e.submitted).count();
      const uint64_t held = j.device_bytes();
      const size_t charges = j.M + j.Fl;
      engine->detach(&j);
      j.arena.release();
      j.design.reset();
      j.release_image();
      std::lock_guard<std::mutex> lk(mu);
      hbm_in_use -= std::min(hbm_in_use, held);
      charges_in_flight -= std::min(charges_in_flight, charges);
      if (e.status == 0)
        stats.completed += 1;
      else
        stats.failed += 1;
      active.erase(std::find(active.begin(), active.end(), id));
      e.done = true;
    }
    if (rb.ev) spare_events.push_back(rb.ev);
    readbacks.pop_front();
    any = true;
  }
  if (!any) return;
  {
    std::lock_guard<std::mutex> lk(mu);
    stats.retire_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  cv_done.notify_all();
}

// Moves a finished entry out of the table and fills the caller's buffers;
// the HPWL and the release of the job's host data run outside the lock.
static int take_result(texo_gp_server *s, std::unique_lock<std::mutex> &lk, int64_t id, uint32_t *out,
                       texo_gp_report *report) {
  auto it = s->entries.find(id);
  std::unique_ptr<texo_gp_server::Entry> e = std::move(it->second);
  s->entries.erase(it);
  lk.unlock();
  if (e->job) {
    if (!e->has_hpwl) e->report.final_hpwl = e->job->hpwl(e->x);
    e->job.reset();
  }
  if (out) std::copy(e->targets.begin(), e->targets.end(), out);
  if (report) *report = e->report;
  return e->status;
}

extern "C" {

texo_gp_server *texo_gp_server_create(int32_t device, uint64_t hbm_budget_bytes, uint32_t max_batch) {
  try {
    std::unique_ptr<texo_gp_server> s(new texo_gp_server());
    s->engine.reset(new Engine(device, 1024));
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)64 << 30;
    s->max_batch = max_batch ? std::min<uint32_t>(max_batch, 1000) : 64;
    s->worker = std::thread([p = s.get()] { p->loop(); });
    return s.release();
  } catch (...) {
    return nullptr;
  }
}

int64_t texo_gp_server_submit(texo_gp_server *s, const texo_gp_problem *problem, const texo_gp_options *options) {
  if (!s) return -TEXO_GP_ERR_INVALID_ARGUMENT;
  std::unique_ptr<texo_gp_server::Entry> e(new texo_gp_server::Entry());
  e->submitted = std::chrono::steady_clock::now();
  e->job.reset(new Job());
  try {
    // deep copy, validation, upload image and (for a new grid size) the
    // transform basis on the caller's thread
    e->job->load(problem, options, s->engine->image_pool());
    // a job that can never fit the budget is rejected, not admitted alone
    if (e->job->device_bytes() > s->budget) return -(int64_t)TEXO_GP_ERR_OUT_OF_MEMORY;
    if (!e->job->empty) s->engine->basis((int)e->job->W, (int)e->job->H);
  } catch (...) {
    const Failure f = texo::current_failure();
    return -(int64_t)f.status;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t id = s->next_id++;
  e->job->id = id;
  s->entries[id] = std::move(e);
  s->pending.push_back(id);
  s->stats.submitted += 1;
  s->cv_work.notify_all();
  return id;
}

int texo_gp_server_poll(texo_gp_server *s, int64_t id, uint32_t *out, texo_gp_report *report) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->entries.find(id);
  if (it == s->entries.end()) return TEXO_GP_ERR_INVALID_ARGUMENT;
  if (!it->second->done) return TEXO_GP_ERR_NOT_READY;
  return take_result(s, lk, id, out, report);
}

int texo_gp_server_wait(texo_gp_server *s, int64_t id, uint32_t *out, texo_gp_report *report) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->entries.find(id);
  if (it == s->entries.end()) return TEXO_GP_ERR_INVALID_ARGUMENT;
  s->cv_done.wait(lk, [&] {
    auto e = s->entries.find(id);
    return e == s->entries.end() || e->second->done;
  });
  // another waiter may have taken the result first
  if (s->entries.find(id) == s->entries.end()) return TEXO_GP_ERR_INVALID_ARGUMENT;
  return take_result(s, lk, id, out, report);
}

int texo_gp_server_cancel(texo_gp_server *s, int64_t id) {
  if (!s) return TEXO_GP_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->entries.find(id);
  if (it == s->entries.end()) return TEXO_GP_ERR_INVALID_ARGUMENT;
  texo_gp_server::Entry &e = *it->second;
  if (e.done) return e.status;
  auto p = std::find(s->pending.begin(), s->pending.end(), id);
  if (p != s->pending.end()) {
    // never admitted: complete it here
    s->pending.erase(p);
    Job &j = *e.job;
    j.release_image();
    e.targets.assign(2 * j.U, 0);
    e.x = j.initial_xy;
    std::memset(&e.report, 0, sizeof(e.report));
    e.status = TEXO_GP_ERR_CANCELLED;
    e.report.status = e.status;
    e.report.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - e.submitted).count();
    e.done = true;
    s->stats.failed += 1;
    lk.unlock();
    s->cv_done.notify_all();
    return 0;
  }
  e.cancel = true;
  s->cv_work.notify_all();
  return 0;
}

int texo_gp_server_stats_get(texo_gp_server *s, texo_gp_server_stats *stats) {
  if (!s || !stats) return TEXO_GP_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  *stats = s->stats;
  stats->active = s->active.size();
  stats->pending = s->pending.size();
  stats->hbm_in_use = s->hbm_in_use;
  stats->graph_rounds = s->engine->graph_rounds();
  return 0;
}

void texo_gp_server_destroy(texo_gp_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->engine->device());
  for (auto &kv : s->entries)
    if (kv.second->job && kv.second->job->pool_index >= 0) s->engine->detach(kv.second->job.get());
  s->entries.clear();
  delete s;
}

}  // extern "C"
