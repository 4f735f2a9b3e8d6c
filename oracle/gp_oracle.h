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
/* propose_area_adjustment (eplace.rs:627-729) at `coordinates`, as
 * texo_gp_session_area_proposal. */
int oracle_session_area_proposal(oracle_session *session, const double *coordinates, double *member_areas,
                                 double *filler_areas, int32_t *requires_reset, double *total_relative_increase);
/* Schedule and optimizer state after oracle_session_iterate (multipliers,
 * their step, E0 scales, last gamma, gamma weights, momentum, Lipschitz). */
int oracle_session_schedule(const oracle_session *session, texo_gp_schedule_state *state);
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

/* ---- density model with explicit charges (test fixtures) --------------- */
typedef struct oracle_density_member {
  uint32_t kind;
  double offset_x, offset_y, charge;
} oracle_density_member;
typedef struct oracle_field_result {
  uint32_t kind;
  uint64_t available, real_charge;
  double real_area, filler_charge, energy, overflow, net_charge, force_l1;
} oracle_field_result;
/* DensityModel::new + initial_fillers (electrostatic_placement.rs:139-240)
 * for units whose members carry their own charges; capacity is
 * [n_kinds][W*H].  Returns the filler count (at most max_fillers written) or
 * -(status). */
int64_t oracle_density_fillers(uint32_t width, uint32_t height, uint32_t n_kinds, const uint32_t *kinds,
                               const uint32_t *capacity, uint32_t n_units, const uint32_t *unit_member_start,
                               const oracle_density_member *members, const double *unit_xy, uint64_t max_fillers,
                               uint32_t *filler_kind, double *filler_xy, double *filler_charge);
/* DensityModel::evaluate_with_positions (electrostatic_placement.rs:265-421)
 * at unit origins unit_xy[2U] and filler positions filler_xy[2n].  Writes up
 * to max_fields records, unit_gradients[F][U][2] and filler_gradients[n][2]
 * (each filler in its own field; either may be NULL).  Returns the field
 * count or -(status). */
int64_t oracle_density_evaluate(uint32_t width, uint32_t height, uint32_t n_kinds, const uint32_t *kinds,
                                const uint32_t *capacity, uint32_t n_units, const uint32_t *unit_member_start,
                                const oracle_density_member *members, const double *unit_xy, uint64_t n_fillers,
                                const uint32_t *filler_kind, const double *filler_xy, const double *filler_charge,
                                uint32_t max_fields, oracle_field_result *fields, double *unit_gradients,
                                double *filler_gradients);

/* ---- schedule and stop rules (fields given as kind + overflow) ----------- */
int oracle_density_kind_is_exchangeable(uint32_t kind);
/* density_is_converged (eplace.rs:1397-1401) */
int oracle_density_is_converged(uint32_t n, const uint32_t *kinds, const double *overflows);
/* area_adjustment_is_ready (eplace.rs:1403-1408) */
int oracle_area_adjustment_is_ready(uint32_t n, const uint32_t *kinds, const double *overflows);
/* routability_adjusted_member_area (eplace.rs:741-758); NaN on error */
double oracle_routability_adjusted_member_area(uint32_t kind, double area, double horizontal, double vertical);
/* area_adjusted_multiplier_step (eplace.rs:1220-1232); NaN on error */
double oracle_area_adjusted_multiplier_step(uint32_t n, const double *multipliers);
/* area_adjusted_density_multipliers (eplace.rs:1183-1217) for fields with
 * (energy, force_l1) against initial energies and a WA gradient of n_units
 * entries per axis.  Returns 0 or a status. */
int oracle_area_adjusted_multipliers(uint32_t n_fields, const double *energy, const double *force_l1,
                                     const double *initial_energies, uint64_t n_units, const double *wa_gradient_x,
                                     const double *wa_gradient_y, double *multipliers);
/* stationary_target_is_fixed (eplace.rs:973-985) on a caller-held tracker
 * (*has_previous, previous_xy[2n]). */
int oracle_stationary_target_is_fixed(int *has_previous, uint32_t *previous_xy, int accepted,
                                      const uint32_t *targets_xy, uint64_t n_units);
/* stop_without_convergence (eplace.rs:989-999): 0, or DID_NOT_CONVERGE with
 * iterations and the first two overflows in `report` (may be NULL). */
int oracle_stop_without_convergence(int best_effort, uint64_t iterations, uint32_t n, const uint32_t *kinds,
                                    const double *overflows, texo_gp_report *report);
/* eplace_iteration_limit_reached (eplace.rs:1011-1013) */
int oracle_iteration_limit_reached(uint64_t completed_iterations);
/* PlacementCheckpoint (eplace.rs:89-120): candidate 0 makes the checkpoint,
 * candidates 1.. are considered in order; returns the retained index. */
uint32_t oracle_checkpoint_choice(uint32_t n_candidates, uint32_t n_fields, const uint32_t *kinds,
                                  const double *overflows);
/* Session-level schedule inputs: density_gamma_weights (eplace.rs:1424-1439)
 * per field, in field order. */
int oracle_session_gamma_weights(oracle_session *session, double *weights);
/* external_baseline_net_bounding_boxes (analytical_placement.rs:238-276):
 * boxes [4*k] (min x, max x, min y, max y) of the nets spanning two or more
 * units, k returned in *n_boxes (at most n_nets written). */
int oracle_net_bounding_boxes(uint32_t n_units, uint32_t n_cells, const double *positions, const uint8_t *fixed_mask,
                              const double *fixed_xy, const double *offsets, uint32_t n_nets,
                              const uint32_t *net_pin_start, const uint32_t *pin_unit, const uint32_t *pin_cell,
                              double *boxes, uint64_t *n_boxes);
/* dynamic_nesterov_step with bounds and metric of their own lengths (the
 * reference validates them before any evaluation). */
int oracle_nesterov_step_lengths(oracle_nesterov_state *state, uint64_t n, uint64_t n_lower, const double *lower,
                                 uint64_t n_upper, const double *upper, uint64_t n_metric, const double *metric,
                                 oracle_evaluate_fn evaluate, void *context, oracle_nesterov_result *result);

/* ---- GP input builders (routing_demand.rs, texo-flow lib.rs) ------------ */
/* channel_orientation (routing_demand.rs:185-198): 1 H, 2 V, 0 neither. */
int oracle_channel_orientation(const char *name);
/* routing_capacity_map (routing_demand.rs:201-239) over wires (orientation
 * code, tile point, capacity) and pips (from, to, blocked). */
int oracle_routing_capacity_map(uint32_t width, uint32_t height, uint32_t n_wires, const uint8_t *wire_orientation,
                                const uint32_t *wire_xy, const uint16_t *wire_capacity, uint32_t n_pips,
                                const uint32_t *pip_from, const uint32_t *pip_to, const uint8_t *pip_blocked,
                                uint32_t *horizontal, uint32_t *vertical);
/* routing_demand_map_with_capacity (routing_demand.rs:262-313): nets of
 * pins at tile points (0xffffffff: unplaced), clock-driven flags; FP64
 * demand summed in net order. */
int oracle_routing_demand_map(uint32_t width, uint32_t height, uint32_t n_nets, const uint32_t *net_pin_start,
                              const uint32_t *pin_xy, const uint8_t *net_clock, double *horizontal, double *vertical,
                              uint64_t *included_nets, uint64_t *excluded_clock_nets);
/* ecp5_placement_criticality_weight (texo-flow lib.rs:3355-3390) with
 * 128-bit delays given as (lo, hi) words. */
uint64_t oracle_criticality_weight(uint64_t path_lo, uint64_t path_hi, uint64_t worst_lo, uint64_t worst_hi,
                                   uint32_t exponent);
/* ecp5_timing_placement_weights (texo-flow lib.rs:3185-3201): the maximum
 * weight per (net, sink) in (net, sink) order; returns the entry count. */
uint64_t oracle_timing_placement_weights(uint64_t n_edges, const uint32_t *net, const uint32_t *sink,
                                         const uint64_t *path_delay_ps, const uint64_t *worst_path_delay_ps,
                                         uint32_t exponent, uint32_t *out_net, uint32_t *out_sink,
                                         uint64_t *out_weight);

/* ---- coarse quadratic seed (lib.rs:4042-4066, 4404-4692) ----------------- */
/* linearize_axis (analytical_placement.rs:283-356) on one axis of an
 * explicit hypergraph (pins as (unit, cell), driver first; arcs as
 * (driver unit, driver cell, sink unit, sink cell, extra)).  Writes up to
 * max_edges edges (left, right unit; weight; timing flag) and returns the
 * edge count. */
uint64_t oracle_linearize_axis(uint32_t n_units, uint32_t n_cells, const double *positions,
                               const uint8_t *fixed_mask, const double *fixed_coordinate, const double *offsets,
                               uint32_t n_nets, const uint32_t *net_pin_start, const uint32_t *pin_unit,
                               const uint32_t *pin_cell, uint32_t n_arcs, const uint32_t *arc_units,
                               const uint32_t *arc_cells, const uint64_t *arc_extra, uint64_t max_edges,
                               uint32_t *edge_left, uint32_t *edge_right, double *edge_weight, uint8_t *edge_timing);
/* solve_analytical_axis (lib.rs:4529-4556) of explicit AxisEquations
 * (adjacency CSR in insertion order); `iterations` receives the PCG
 * iterations.  DID_NOT_CONVERGE where the reference reports a PCG
 * breakdown or non-convergence. */
int oracle_solve_axis(uint32_t n, const double *diagonal, const uint32_t *adjacency_start,
                      const uint32_t *adjacency_other, const double *adjacency_weight, const double *rhs,
                      const uint8_t *anchored, const double *initial, uint32_t center, double *solution,
                      uint64_t *iterations);
/* The coarse seed of analytical_place (lib.rs:4042-4066): both axes
 * linearized at the device centre and solved; out_xy[2*n_units]
 * interleaved.  report->iterations = PCG iterations of x plus y. */
int oracle_quadratic_seed(const texo_gp_problem *problem, double *out_xy, texo_gp_report *report);

/* ---- legalization auction ----------------------------------------------- */
/* assign_implicitly (legalization/auction.rs:123-253) over explicit domains
 * (CSR by bidder: object, base cost), the two-best search of the reference's
 * tests (auction.rs:258-276).  assignment[n_bidders]: object or -1 (the
 * bidder's private dummy); stats[3] = bids, candidates examined, unmatched.
 * Returns 0 or an AuctionError status (TEXO_GP_ERR_AUCTION_*). */
int oracle_auction(uint64_t n_bidders, uint64_t n_objects, uint64_t max_base_cost, const uint32_t *domain_start,
                   const uint32_t *domain_object, const uint64_t *domain_cost, int64_t *assignment,
                   uint64_t *stats);

/* ---- legalization repair (gp_repair.cpp) -------------------------------- */
/* repair_relaxed (legalization/repair.rs:60-155) over explicit candidate
 * footprints (texo_gp_repair_problem); placed[n_units] = candidate j per
 * unit; stats[12] as texo_gp_repair.  Returns 0, TEXO_GP_ERR_REPAIR_PROJECTION
 * or TEXO_GP_ERR_REPAIR_NO_BEL (*error_unit = first pending unit). */
int oracle_repair(const texo_gp_repair_problem *problem, int64_t *placed, uint64_t *stats, int64_t *error_unit);
/* pin_blocker_branches (repair.rs:1169-1241) of (unit, choice) on a state
 * holding the listed installed units: branches as unit bitmasks (units < 64)
 * in out[max_out]; returns the number of branches (-1: malformed input). */
int oracle_repair_pin_branches(const texo_gp_repair_problem *problem, uint64_t n_installed,
                               const uint64_t *installed_unit, const uint64_t *installed_choice, uint64_t unit,
                               uint64_t choice, uint64_t mandatory_mask, uint64_t *out, int max_out);

/* ChoiceCursor order (repair.rs:649-793) of `unit`'s candidates: up to
 * max_out choices into out; returns their number (-1: malformed input). */
int oracle_repair_cursor(const texo_gp_repair_problem *problem, uint64_t unit, uint64_t *out, int max_out);

#ifdef __cplusplus
}
#endif

#endif
