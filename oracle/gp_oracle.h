/*
 * gp_oracle.h — FP64 CPU restatement of texo-pnr's electrostatic global
 * placement (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity oracle and the timed CPU baseline.  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product library never links, calls or falls back to it.
 *
 * Every function follows a reference function named in gp_oracle.cpp with its
 * file:line.  The session API mirrors include/texo_gp.h one-for-one with an
 * `oracle_` prefix so a test can feed the same problem to both sides.
 */
#ifndef TEXO_GP_ORACLE_H
#define TEXO_GP_ORACLE_H

#include "../include/texo_gp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Threads used by the evaluation (1 = serial).  Multi-threaded runs keep the
 * reference's fixed-order reductions, so results are bit-identical. */
void oracle_set_threads(int threads);

/* ---- placement and sessions (same contracts as texo_gp.h) --------------- */
int oracle_place(const texo_gp_problem *problem, const texo_gp_options *options,
                 uint32_t *out_targets_xy, texo_gp_report *report);

typedef struct oracle_session oracle_session;
oracle_session *oracle_session_create(const texo_gp_problem *problem,
                                      const texo_gp_options *options,
                                      texo_gp_report *report);
void oracle_session_destroy(oracle_session *session);
int oracle_session_sizes(const oracle_session *session, uint64_t *coordinates,
                         uint32_t *fields, uint64_t *fillers, uint64_t *density_units);
int oracle_session_initial(const oracle_session *session, double *coordinates,
                           double *lower, double *upper, uint32_t *field_kind,
                           uint32_t *filler_field, double *filler_charge);
int oracle_session_evaluate(oracle_session *session, const double *coordinates,
                            double gamma, double *wa_value, double *wa_gradient,
                            texo_gp_field_eval *fields, double *unit_gradient,
                            double *filler_gradient, double *density,
                            double *potential);
int oracle_session_run(oracle_session *session, uint32_t *out_targets_xy,
                       texo_gp_report *report);
int oracle_session_iterate(oracle_session *session, uint32_t iterations,
                           texo_gp_report *report);
int oracle_session_coordinates(const oracle_session *session, double *coordinates);
int oracle_session_rudy(oracle_session *session, const double *coordinates,
                        double *horizontal, double *vertical);
/* Continuous bbox HPWL of the retained nets at `coordinates` (FP64). */
double oracle_session_hpwl(oracle_session *session, const double *coordinates);
/* Combined objective (combined_objective, eplace.rs:1022-1044) at
 * `coordinates` with explicit multipliers and scales; gradient[C]. */
int oracle_session_combined(oracle_session *session, const double *coordinates,
                            double gamma, const double *multipliers,
                            const double *scales, double *value, double *gradient);
/* Diagonal preconditioner (eplace.rs:1046-1100) for `multipliers`. */
int oracle_session_preconditioner(oracle_session *session, const double *multipliers,
                                  double *diagonal);

/* ---- Poisson / B-spline primitives -------------------------------------- */
int oracle_poisson_solve(uint32_t width, uint32_t height, const double *density,
                         double *potential, double *energy);
int oracle_poisson_solve_matrix(uint32_t width, uint32_t height,
                                const double *density, double *potential,
                                double *energy);
void oracle_bspline_axis(uint32_t length, double coordinate, uint32_t *index,
                         double *weight, double *derivative);
void oracle_deposit_bspline(double *density, uint32_t width, uint32_t height,
                            double x, double y, double charge);
void oracle_sample_bspline(const double *potential, uint32_t width, uint32_t height,
                           double x, double y, double *value_gx_gy);

/* ---- weighted-average wirelength on an explicit hypergraph -------------- */
/* pins are (unit, cell); positions[2*n_units]; fixed_mask/fixed_xy/offsets per
 * cell.  Writes value, gradient_x[n_units], gradient_y[n_units];
 * `timing_value` etc. may be NULL. */
int oracle_weighted_average(uint32_t n_units, uint32_t n_cells,
                            const double *positions, const uint8_t *fixed_mask,
                            const double *fixed_xy, const double *offsets,
                            uint32_t n_nets, const uint32_t *net_pin_start,
                            const uint32_t *pin_unit, const uint32_t *pin_cell,
                            uint32_t n_arcs, const uint32_t *arc_units,
                            const uint32_t *arc_cells, const uint64_t *arc_extra,
                            double gamma, double *value, double *gradient_x,
                            double *gradient_y);
int oracle_wirelength_preconditioner(uint32_t n_units, uint32_t n_nets,
                                     const uint32_t *net_pin_start,
                                     const uint32_t *pin_unit, uint32_t n_arcs,
                                     const uint32_t *arc_units,
                                     const uint64_t *arc_extra, double *diagonal,
                                     double *baseline);

/* ---- Nesterov step ------------------------------------------------------ */
typedef struct oracle_nesterov_state {
  double *coordinates;          /* [n] */
  double *previous_coordinates; /* [n] */
  double momentum;
  double lipschitz;
  int has_lipschitz;
} oracle_nesterov_state;

/* Evaluator: fills value and gradient[n]; returns 0 or a nonzero failure that
 * the step reports as TEXO_GP_ERR_EVALUATION. */
typedef int (*oracle_evaluate_fn)(void *context, const double *coordinates,
                                  uint64_t n, double *value, double *gradient);

typedef struct oracle_nesterov_result {
  int32_t status;      /* 0 ok, else texo_gp_status */
  int32_t accepted;    /* 1 Accepted, 0 NumericallyStationary */
  double objective;
  double stationarity;
  double coordinate_change;
  uint64_t line_search_trials;
  uint64_t evaluations;
  int64_t payload_evaluation; /* 1-based index of the evaluation returned */
  int64_t error_index;
} oracle_nesterov_result;

int oracle_nesterov_step(oracle_nesterov_state *state, uint64_t n,
                         const double *lower, const double *upper,
                         const double *metric, oracle_evaluate_fn evaluate,
                         void *context, oracle_nesterov_result *result);
double oracle_box_metric_kkt_residual(uint64_t n, const double *gradient,
                                      const double *coordinates,
                                      const double *lower, const double *upper,
                                      const double *metric);

/* ---- schedule helpers (return NaN where the reference returns an error) -- */
double oracle_field_wirelength_gamma(double overflow, double target, double base);
double oracle_weighted_gamma_mean(uint32_t n, const double *gamma,
                                  const double *weight, double fallback);
double oracle_multiplier_growth(double logarithm);
void oracle_augmented_density(double energy, double initial_energy,
                              double multiplier, double *value, double *coefficient);
double oracle_density_overflow_target(uint32_t kind);
uint32_t oracle_rounded_coordinate(double value, uint32_t extent);

/* ---- area rounds -------------------------------------------------------- */
int oracle_continuous_routing_demand(uint32_t width, uint32_t height, uint64_t n_boxes,
                                     const double *boxes /* [4n] minx,maxx,miny,maxy */,
                                     double *horizontal, double *vertical);
double oracle_routability_optimized_area(double area, double horizontal, double vertical);
typedef struct oracle_area_member {
  uint32_t unit;
  uint32_t kind;
  double current_area;
  double routability_area;
  double pin_area;
  double clustering_area;
} oracle_area_member;
int oracle_adjust_instance_areas(uint64_t n_members, const oracle_area_member *members,
                                 uint64_t n_fillers, const uint32_t *filler_kind,
                                 const double *filler_area, double *member_areas,
                                 double *filler_areas, int *requires_reset);
/* registers: cell, x, y; controls as in texo_gp_problem. Writes areas[n]
 * aligned with the register list; returns 0 or -1 (None). */
int oracle_clustering_areas(uint64_t n_registers, const uint32_t *register_cell,
                            const double *register_xy, uint64_t n_controls,
                            const uint32_t *control_cell,
                            const uint64_t *control_clock_lsr,
                            const uint64_t *control_ce,
                            uint64_t movable_instance_count, double *areas);
double oracle_soft_division_ceiling(double value, double divisor);

#ifdef __cplusplus
}
#endif

#endif
