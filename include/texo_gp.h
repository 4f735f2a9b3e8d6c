/*
 * texo_gp.h — C ABI of the MI355X electrostatic global-placement library.
 *
 * This boundary replaces the body of texo-pnr's crate-private entry point
 *
 *   fn place(graph, units, hypergraph, initial_x, initial_y,
 *            fixed_coordinates, cell_offsets, routing_capacity,
 *            register_controls, best_effort) -> Result<Vec<Point>, EplaceError>
 *
 * (crates/texo-pnr/src/eplace.rs:242-253, single caller
 * crates/texo-pnr/src/lib.rs:4129-4143).  The caller flattens its borrowed
 * inputs into `texo_gp_problem` (plain pointers and sizes, all borrowed for the
 * duration of the call), receives integer tile targets in a caller-allocated
 * `uint32_t[2 * n_units]`, and maps a nonzero status to its `EplaceError`
 * variant with the payload carried in `texo_gp_report`.  Nothing unwinds or
 * aborts across this boundary.
 *
 * All coordinates are in device tiles; one Poisson bin is one tile
 * (crates/texo-pnr/src/eplace.rs:131-133).
 */
#ifndef TEXO_GP_H
#define TEXO_GP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TEXO_GP_ABI_VERSION 1u

/* ResourceKind codes.  Numeric order equals the `derive(Ord)` order of
 * texo_model::ResourceKind (crates/texo-model/src/lib.rs:68-87), so sorting
 * codes reproduces the reference's BTreeMap field order:
 * Logic < Lut(n) < Register < Memory < Clock < Io < Constant < Dsp. */
#define TEXO_KIND_LOGIC 0x000u
#define TEXO_KIND_LUT(n) (0x100u + (uint32_t)(n)) /* n <= 255 */
#define TEXO_KIND_REGISTER 0x200u
#define TEXO_KIND_MEMORY 0x300u
#define TEXO_KIND_CLOCK 0x400u
#define TEXO_KIND_IO 0x500u
#define TEXO_KIND_CONSTANT 0x600u
#define TEXO_KIND_DSP 0x700u

/* Maximum number of electrostatic density fields of one problem. Only
 * Lut(n), Register and Memory are exchangeable density kinds
 * (crates/texo-pnr/src/eplace.rs:1417-1422). */
#define TEXO_GP_MAX_FIELDS 4

/* Status codes.  Each names one variant of EplaceError
 * (crates/texo-pnr/src/eplace.rs:27-38), ElectrostaticError
 * (crates/texo-pnr/src/electrostatic_placement.rs:15-69), NesterovError
 * (crates/texo-pnr/src/global_placement.rs:81-109) or InstanceAreaError
 * (crates/texo-pnr/src/instance_area.rs:55-61); codes >= 900 are
 * boundary-level failures with no reference counterpart. */
enum texo_gp_status {
  TEXO_GP_OK = 0,
  /* EplaceError */
  TEXO_GP_ERR_DENSITY_FIELD_CHANGED = 10,
  TEXO_GP_ERR_INVALID_NORMALIZATION = 11,
  TEXO_GP_ERR_DID_NOT_CONVERGE = 12,
  /* ElectrostaticError (EplaceError::Density) */
  TEXO_GP_ERR_EMPTY_GRID = 100,
  TEXO_GP_ERR_GRID_TOO_LARGE = 101,
  TEXO_GP_ERR_DENSITY_LENGTH = 102,
  TEXO_GP_ERR_NON_FINITE_DENSITY = 103,
  TEXO_GP_ERR_EMPTY_UNIT = 104,
  TEXO_GP_ERR_NON_FINITE_UNIT_COORDINATE = 105,
  TEXO_GP_ERR_INVALID_MEMBER_CHARGE = 106,
  TEXO_GP_ERR_POINT_OUTSIDE_GRID = 107,
  TEXO_GP_ERR_FIXED_CAPACITY_EXHAUSTED = 108,
  TEXO_GP_ERR_INSUFFICIENT_CAPACITY = 109,
  TEXO_GP_ERR_INSUFFICIENT_AREA = 110,
  TEXO_GP_ERR_NON_FINITE_FILLER = 111,
  TEXO_GP_ERR_NEGATIVE_FILLER_CHARGE = 112,
  TEXO_GP_ERR_UNEXPECTED_FILLER_KIND = 113,
  TEXO_GP_ERR_FILLER_CHARGE_MISMATCH = 114,
  /* NesterovError (EplaceError::Optimizer) */
  TEXO_GP_ERR_INVALID_BOUNDS = 200,
  TEXO_GP_ERR_INVALID_METRIC_LENGTH = 201,
  TEXO_GP_ERR_NON_POSITIVE_METRIC = 202,
  TEXO_GP_ERR_INVALID_INITIAL_COORDINATE = 203,
  TEXO_GP_ERR_INVALID_MOMENTUM = 204,
  TEXO_GP_ERR_INVALID_LIPSCHITZ = 205,
  TEXO_GP_ERR_INVALID_OBJECTIVE = 206,
  TEXO_GP_ERR_INVALID_GRADIENT_LENGTH = 207,
  TEXO_GP_ERR_NON_FINITE_GRADIENT = 208,
  TEXO_GP_ERR_LINE_SEARCH_OVERFLOW = 209,
  TEXO_GP_ERR_ITERATION_OVERFLOW = 210,
  TEXO_GP_ERR_EVALUATION = 211,
  /* InstanceAreaError (EplaceError::InstanceArea) */
  TEXO_GP_ERR_INVALID_MEMBER_AREA = 300,
  TEXO_GP_ERR_INVALID_FILLER_AREA = 301,
  TEXO_GP_ERR_POSTCONDITION_VIOLATION = 302,
  TEXO_GP_ERR_ARITHMETIC_OVERFLOW = 303,
  /* AuctionError (crates/texo-pnr/src/legalization/auction.rs:44-53) of the
   * legalization auction */
  TEXO_GP_ERR_AUCTION_OVERFLOW = 400,
  TEXO_GP_ERR_AUCTION_OBJECT_OUT_OF_RANGE = 401,
  TEXO_GP_ERR_AUCTION_BASE_COST_OUT_OF_RANGE = 402,
  TEXO_GP_ERR_AUCTION_DUPLICATE_BEST_OBJECT = 403,
  /* the transactional repair (crates/texo-pnr/src/legalization/repair.rs):
   * PnrError::InvalidPlacement of its projection_error (mod.rs:1184-1188), a
   * relaxed projection it cannot start from, and PnrError::NoBel when no
   * complete legal assignment exists */
  TEXO_GP_ERR_REPAIR_PROJECTION = 410,
  TEXO_GP_ERR_REPAIR_NO_BEL = 411,
  /* Boundary-level failures (no reference counterpart). */
  TEXO_GP_ERR_INVALID_ARGUMENT = 900,
  TEXO_GP_ERR_DEVICE = 901,
  TEXO_GP_ERR_OUT_OF_MEMORY = 902,
  TEXO_GP_ERR_UNSUPPORTED = 903,
  TEXO_GP_ERR_CANCELLED = 904,
  TEXO_GP_ERR_NOT_READY = 905,
  TEXO_GP_ERR_INJECTED = 906
};

/* Precision of the Poisson transform (the only contraction on MFMA).
 * Every other quantity keeps the same arithmetic in all modes: int64
 * fixed-point density, FP64 coordinates and global reductions. */
enum texo_gp_dct_precision {
  TEXO_GP_DCT_FP32 = 0,       /* plain FP32 contraction */
  TEXO_GP_DCT_BF16X3 = 1,     /* split hi/lo BF16 on MFMA, FP32 accumulate */
  TEXO_GP_DCT_FP16_SCALED = 2 /* split hi/lo FP16 on MFMA with power-of-two
                                 scaling per 32-row block, FP32 accumulate */
};

/* Flattened placement problem.  Every pointer is borrowed for the call.
 *
 * Units are PlacementUnits (crates/texo-pnr/src/lib.rs:3729-3733).  A unit
 * with one choice is fixed (`unit_fixed`); its exchangeable BELs are listed in
 * the fixed-occupancy arrays (crates/texo-pnr/src/eplace.rs:263-271).  A
 * movable unit's density members are the exchangeable cells of assignment
 * row 0, with offsets relative to that row's origin BEL
 * (crates/texo-pnr/src/eplace.rs:272-293); a movable unit with no such member
 * is not a density unit. */
typedef struct texo_gp_problem {
  uint32_t abi_version; /* TEXO_GP_ABI_VERSION */
  uint32_t width;       /* device grid (Device::width) */
  uint32_t height;      /* device grid (Device::height) */

  /* Cells (CellId order).  Fixed cells carry exact coordinates
   * (`fixed_coordinates`, crates/texo-pnr/src/lib.rs:4003-4033); every cell
   * carries its offset from its unit origin (`cell_offsets`). */
  uint32_t n_cells;
  const uint32_t *cell_unit;     /* [n_cells] owning unit */
  const uint8_t *cell_fixed;     /* [n_cells] 1 = fixed coordinate present */
  const double *cell_fixed_xy;   /* [2*n_cells] read where cell_fixed */
  const double *cell_offset_xy;  /* [2*n_cells] */

  /* Units. */
  uint32_t n_units;
  const uint8_t *unit_fixed;        /* [n_units] choices.len() == 1 */
  const double *unit_initial_xy;    /* [2*n_units] initial_x/initial_y */
  const double *unit_bounds;        /* [4*n_units] lo_x, lo_y, hi_x, hi_y
                                       (choice_bounds, eplace.rs:1470-1481) */

  /* Density members of movable units, CSR by unit (unit order, then the
   * unit's cell order). */
  const uint32_t *unit_member_start; /* [n_units+1] */
  const uint16_t *member_kind;       /* [n_members] TEXO_KIND_* */
  const double *member_offset_xy;    /* [2*n_members] */
  const uint32_t *member_cell;       /* [n_members] */

  /* BEL capacity per exchangeable kind, before fixed occupancy is removed
   * (DensityModel::new, electrostatic_placement.rs:139-154). */
  uint32_t n_capacity_kinds;
  const uint16_t *capacity_kind; /* [n_capacity_kinds] */
  const uint32_t *capacity;      /* [n_capacity_kinds * width * height]
                                    row-major y*width+x */

  /* Fixed occupancy (kind, point) removed from capacity. */
  uint32_t n_fixed_occupancy;
  const uint16_t *fixed_occupancy_kind; /* [n] */
  const uint32_t *fixed_occupancy_xy;   /* [2n] */

  /* AnalyticalHypergraph (analytical_placement.rs:116-168): retained
   * non-clock nets with at least one sink, pins listed driver first. */
  uint32_t n_nets;
  const uint32_t *net_pin_start; /* [n_nets+1] */
  const uint32_t *pin_cell;      /* [n_pins] */
  /* Sink-local timed arcs with extra weight w-1 > 0. */
  uint32_t n_arcs;
  const uint32_t *arc_cells;        /* [2*n_arcs] driver, sink */
  const uint64_t *arc_extra_weight; /* [n_arcs] */

  /* Optional RoutingCapacityMap (routing_demand.rs:22-27) on the same grid;
   * NULL selects the pure density objective. */
  const uint32_t *routing_horizontal; /* [width*height] or NULL */
  const uint32_t *routing_vertical;   /* [width*height] or NULL */

  /* RegisterControlSet records (crates/texo-pnr/src/lib.rs:2156-2164). */
  uint32_t n_register_controls;
  const uint32_t *register_cell;     /* [n] */
  const uint64_t *register_clock_lsr; /* [2n] */
  const uint64_t *register_ce;       /* [n] */

  uint8_t best_effort; /* ElectrostaticBestEffort */
} texo_gp_problem;

/* Options.  Zero-initialised options reproduce the reference exactly:
 * stop rules (eplace.rs:1397-1415), gamma base 8 tiles (eplace.rs:133) and the
 * 3,000-iteration cap (eplace.rs:134). Extensions are recorded in reports. */
typedef struct texo_gp_options {
  uint32_t dct_precision;    /* texo_gp_dct_precision */
  uint32_t fixed_iterations; /* >0: run exactly this many Nesterov iterations
                                of the hot loop, no stop rule, no area rounds */
  uint32_t max_iterations;   /* 0 -> 3000 */
  uint32_t trace;            /* 1: print reference-shaped metrics lines */
  double gamma_base;         /* 0 -> 8.0 tiles */
  int32_t device;            /* HIP device ordinal */
  uint32_t inject_failure_at_evaluation; /* >0: evaluation N fails
                                            (TEXO_GP_ERR_INJECTED) */
} texo_gp_options;

/* Result record.  Payload fields follow the failing variant:
 * err_index/err_index2 carry unit/member/filler/coordinate indices,
 * err_kind a TEXO_KIND_* code, err_x/err_y a grid point, err_value[0..1]
 * demand/available or expected/actual values. */
typedef struct texo_gp_report {
  int32_t status;
  int64_t err_index;
  int64_t err_index2;
  uint32_t err_kind;
  uint32_t err_x, err_y;
  double err_value[2];

  uint64_t iterations;         /* dynamic_nesterov_step calls */
  uint64_t evaluations;        /* objective evaluations */
  uint64_t line_search_trials;
  uint64_t area_rounds;
  uint32_t dct_precision;
  uint32_t fixed_iterations;
  double gamma_base;

  double final_wirelength;     /* WA value at the final coordinates */
  double final_hpwl;           /* continuous bbox HPWL, FP64 */
  uint32_t n_fields;
  uint32_t field_kind[TEXO_GP_MAX_FIELDS];
  double field_overflow[TEXO_GP_MAX_FIELDS]; /* normalized, physical */
  double field_energy[TEXO_GP_MAX_FIELDS];
  double elapsed_seconds;
} texo_gp_report;

/* ---- one-call entry point (replaces eplace::place) ---------------------- */

/* Runs electrostatic global placement on the GPU and writes
 * `out_targets_xy[2*n_units]` (Point{x,y} per unit, rounded like
 * rounded_coordinate, crates/texo-pnr/src/lib.rs:4566-4568).  Returns the
 * status also stored in `report->status`. */
int texo_gp_place(const texo_gp_problem *problem, const texo_gp_options *options,
                  uint32_t *out_targets_xy, texo_gp_report *report);

/* ---- session API: evaluate-level seams used by parity tests -------------- */

typedef struct texo_gp_session texo_gp_session;

/* Builds the density model, fillers and device-resident problem
 * (eplace.rs:255-336).  On failure returns NULL and fills `report`. */
texo_gp_session *texo_gp_session_create(const texo_gp_problem *problem,
                                        const texo_gp_options *options,
                                        texo_gp_report *report);
void texo_gp_session_destroy(texo_gp_session *session);

/* Sizes: coordinate count C = 2*(n_units + n_fillers), field count, filler
 * count, density-unit count. */
int texo_gp_session_sizes(const texo_gp_session *session, uint64_t *coordinates,
                          uint32_t *fields, uint64_t *fillers,
                          uint64_t *density_units);

/* Initial interleaved coordinates [x0,y0,...] units then fillers, clamped to
 * bounds (initial_coordinates_and_bounds, eplace.rs:1441-1468), plus the
 * per-field kind codes and per-filler (kind, charge). */
int texo_gp_session_initial(const texo_gp_session *session, double *coordinates,
                            double *lower, double *upper, uint32_t *field_kind,
                            uint32_t *filler_field, double *filler_charge);

/* Field of one objective evaluation (ContinuousProblem::evaluate,
 * eplace.rs:137-205, with DensityModel::evaluate_with_positions,
 * electrostatic_placement.rs:265-421).  Any output pointer may be NULL. */
typedef struct texo_gp_field_eval {
  double energy;
  double normalized_overflow; /* physical, before fillers */
  double net_charge;
  double force_l1;
  double real_area;
  double filler_charge;
} texo_gp_field_eval;

/* Evaluates WA wirelength and every density field at `coordinates[C]` with
 * global smoothing `gamma`.  Outputs: wa_value; wa_gradient[2*n_units]
 * interleaved; per field (f-major): unit_gradient[F*2*n_density_units],
 * filler_gradient[2*n_fillers] (each filler in its own field), density[F*W*H],
 * potential[F*W*H]. */
int texo_gp_session_evaluate(texo_gp_session *session, const double *coordinates,
                             double gamma, double *wa_value, double *wa_gradient,
                             texo_gp_field_eval *fields, double *unit_gradient,
                             double *filler_gradient, double *density,
                             double *potential);

/* Poisson solve of one FP64 density grid per field on the device transform
 * path (mean removal, C·rho·C^T, 1/lambda, inverse, energy); `density` is
 * [F*W*H] and outputs are potential[F*W*H], energy[F].  Used by parity tests
 * of the transform alone (Poisson2d::solve, electrostatic_placement.rs:725-776). */
int texo_gp_session_poisson(texo_gp_session *session, const double *density,
                            double *potential, double *energy);

/* Runs the placement loop from the session's current state.  Targets and
 * report as texo_gp_place. */
int texo_gp_session_run(texo_gp_session *session, uint32_t *out_targets_xy,
                        texo_gp_report *report);

/* Final interleaved coordinates after a run. */
int texo_gp_session_coordinates(const texo_gp_session *session, double *coordinates);

/* Continuous RUDY demand of the current coordinates
 * (continuous_routing_demand, eplace.rs:856-937, over
 * external_baseline_net_bounding_boxes, analytical_placement.rs:238-276):
 * horizontal[W*H], vertical[W*H]. */
int texo_gp_session_rudy(texo_gp_session *session, const double *coordinates,
                         double *horizontal, double *vertical);

/* Kernel timing: when enabled, each launch of the named kernel ("deposit",
 * "wa", "gather", "dct", "rudy", "optimizer") is bracketed by HIP events on
 * the stream it runs on; `avg_ms` is the mean duration, `count` the number
 * of launches since the last reset and `avg_tasks` (may be null) the mean
 * number of evaluations one launch covered (for "optimizer": the mean number
 * of coordinate-long vectors it read or wrote).  Timed rounds are issued
 * launch by launch, never replayed from a captured graph. */
int texo_gp_session_set_kernel_timing(texo_gp_session *session, const char *kernel,
                                      int enable);
int texo_gp_session_kernel_timing(texo_gp_session *session, double *avg_ms,
                                  uint64_t *count, double *avg_tasks);

/* Runs exactly `iterations` Nesterov iterations of the hot loop
 * (optimize_density_once, eplace.rs:503-574) with the stop rule disabled:
 * the measured unit of work of the benchmark. */
int texo_gp_session_iterate(texo_gp_session *session, uint32_t iterations,
                            texo_gp_report *report);

/* Combined objective (combined_objective, eplace.rs:1022-1044, with
 * augmented_density_value_and_coefficient, eplace.rs:1111-1124) at
 * `coordinates[C]` with explicit multipliers[F] and initial-energy scales[F]:
 * the value and the interleaved gradient[C] the optimizer steps on. */
int texo_gp_session_combined(texo_gp_session *session, const double *coordinates, double gamma,
                             const double *multipliers, const double *scales, double *value,
                             double *gradient);

/* Diagonal preconditioner (diagonal_preconditioner, eplace.rs:1046-1100):
 * wirelength incidence plus multiplier-weighted charges, floored at 1.0, for
 * `multipliers[F]`; diagonal[C]. */
int texo_gp_session_preconditioner(texo_gp_session *session, const double *multipliers,
                                   double *diagonal);

/* Rounded unit targets of interleaved `coordinates` (the first 2*n_units
 * entries) as the device checkpoint rounds them (rounded_targets_from_coordinates,
 * eplace.rs:954-971, rounded_coordinate, lib.rs:4566-4568); out[2*n_units]. */
int texo_gp_session_round_targets(texo_gp_session *session, const double *coordinates,
                                  uint32_t *out_targets_xy);

/* The area proposal of an area round at `coordinates[C]`
 * (propose_area_adjustment, eplace.rs:627-729): RUDY utilisation
 * (ContinuousRoutingDemand::utilization, eplace.rs:822-854), routability
 * areas (eplace.rs:741-758), register-clustering areas
 * (register_clustering.rs:17-81) and adjust_instance_areas
 * (instance_area.rs:157-372).  Writes the proposed member areas[n_members],
 * filler areas[n_fillers], whether the optimizer would reset and the total
 * relative area increase; applies nothing.  The problem needs a routing
 * capacity map. */
int texo_gp_session_area_proposal(texo_gp_session *session, const double *coordinates, double *member_areas,
                                  double *filler_areas, int32_t *requires_reset,
                                  double *total_relative_increase);

/* clustering_areas (crates/texo-pnr/src/register_clustering.rs:17-81) on
 * `device`: registers (cell, x, y), control sets (cell, clock, lsr, ce) and
 * the movable instance count; areas[n_registers] aligned with the registers.
 * TEXO_GP_ERR_INVALID_NORMALIZATION where the reference returns None. */
int texo_gp_clustering_areas(int32_t device, uint64_t n_registers, const uint32_t *register_cell,
                             const double *register_xy, uint64_t n_controls, const uint32_t *control_cell,
                             const uint64_t *control_clock_lsr, const uint64_t *control_ce,
                             uint64_t movable_instance_count, double *areas);

/* Optimizer and schedule state after the last iterate/run call: density
 * multipliers and their step (eplace.rs:1126-1298), initial energies E0, the
 * wirelength gamma of the last iteration and the per-field gamma weights
 * (eplace.rs:1325-1439), Nesterov momentum and Lipschitz estimate
 * (global_placement.rs:30-58). */
typedef struct texo_gp_schedule_state {
  uint32_t n_fields;
  double multiplier[TEXO_GP_MAX_FIELDS];
  double multiplier_step;
  double initial_energy[TEXO_GP_MAX_FIELDS];
  double gamma;
  double gamma_weight[TEXO_GP_MAX_FIELDS];
  double momentum;
  double lipschitz; /* 0 when not yet estimated */
} texo_gp_schedule_state;
int texo_gp_session_schedule(const texo_gp_session *session, texo_gp_schedule_state *state);

/* ---- lockstep batches: multi-start placement ----------------------------- */
/* No counterpart in the reference (it places one design per call).  A batch
 * owns one GPU and runs every job it holds in lockstep rounds, one launch per
 * kernel family over all jobs: multi-start seeds, timing-weight variants or
 * unrelated designs of any grid size.  Each job's trajectory and result are
 * bit-identical to the same job run alone. */
typedef struct texo_gp_batch texo_gp_batch;

texo_gp_batch *texo_gp_batch_create(int32_t device);
/* Returns the job's index (>= 0), or -(status) with `report` filled. */
int64_t texo_gp_batch_add(texo_gp_batch *batch, const texo_gp_problem *problem,
                          const texo_gp_options *options, texo_gp_report *report);
/* Full placement of every job (eplace::place each); returns once every job
 * has finished or failed.  Per-job outcomes come from texo_gp_batch_result. */
int texo_gp_batch_run(texo_gp_batch *batch);
/* `iterations` more GP iterations of every job with the stop rules off, as
 * texo_gp_session_iterate. */
int texo_gp_batch_iterate(texo_gp_batch *batch, uint32_t iterations);
/* Status of job `job`; targets (best checkpoint, after run) and report. */
int texo_gp_batch_result(texo_gp_batch *batch, uint64_t job, uint32_t *out_targets_xy,
                         texo_gp_report *report);
/* Sizes of job `job`, as texo_gp_session_sizes. */
int texo_gp_batch_sizes(const texo_gp_batch *batch, uint64_t job, uint64_t *coordinates, uint32_t *fields,
                        uint64_t *fillers, uint64_t *density_units);
int texo_gp_batch_coordinates(texo_gp_batch *batch, uint64_t job, double *coordinates);
/* Kernel timing over the batch's launches (one launch covers the
 * evaluations of every job of one of the batch's two job groups). */
int texo_gp_batch_set_kernel_timing(texo_gp_batch *batch, const char *kernel, int enable);
int texo_gp_batch_kernel_timing(texo_gp_batch *batch, double *avg_ms, uint64_t *count,
                                double *avg_tasks);
void texo_gp_batch_destroy(texo_gp_batch *batch);

/* ---- serving: continuous batching of many placement jobs --------------- */
/* No counterpart in the reference (it places one design per call).  A server
 * owns one GPU and a worker thread that runs every admitted job in lockstep
 * rounds: each round issues one launch per kernel family over all jobs.
 * Submitted problems are deep-copied and validated on the caller's thread;
 * jobs are admitted in submission order while their HBM footprint fits
 * `hbm_budget_bytes` and fewer than `max_batch` jobs run; jobs that arrive
 * while others run join at the next round.  A job's result is bit-identical
 * to the same job run alone. */
typedef struct texo_gp_server texo_gp_server;

typedef struct texo_gp_server_stats {
  uint64_t submitted, admitted, completed, failed;
  uint64_t rounds;          /* lockstep rounds executed */
  uint64_t active, pending; /* jobs on the GPU / waiting for admission */
  uint64_t hbm_in_use;      /* bytes held by admitted jobs */
  double busy_seconds;      /* wall time spent in rounds */
  double admit_seconds;     /* server-thread time admitting jobs (arena, upload enqueue) */
  double retire_seconds;    /* server-thread time reading back and releasing finished jobs */
  double wait_seconds;      /* part of busy_seconds spent waiting for the GPU to end a round */
  double launch_seconds;    /* server-thread time planning rounds and enqueueing their work */
  uint64_t graph_rounds;    /* rounds replayed from a captured HIP graph */
} texo_gp_server_stats;

texo_gp_server *texo_gp_server_create(int32_t device, uint64_t hbm_budget_bytes, uint32_t max_batch);
/* Returns a job id > 0, or -(status) if the problem is rejected: invalid
 * input (the EplaceError variants of set-up), or TEXO_GP_ERR_OUT_OF_MEMORY
 * when the job's device footprint exceeds the server's HBM budget. */
int64_t texo_gp_server_submit(texo_gp_server *server, const texo_gp_problem *problem,
                              const texo_gp_options *options);
/* TEXO_GP_ERR_NOT_READY while the job runs; otherwise the job's status, with
 * targets and report filled (report->elapsed_seconds = submission to
 * completion).  A finished job's result is released by the first successful
 * poll/wait. */
int texo_gp_server_poll(texo_gp_server *server, int64_t job, uint32_t *out_targets_xy,
                        texo_gp_report *report);
int texo_gp_server_wait(texo_gp_server *server, int64_t job, uint32_t *out_targets_xy,
                        texo_gp_report *report);
/* Cancels a job: a pending job is never admitted, a running one stops at
 * the next round boundary and releases its device memory; either then
 * completes with TEXO_GP_ERR_CANCELLED (its result is taken by poll/wait as
 * usual).  Returns 0 when the cancellation was registered, the job's status
 * when it had already finished, TEXO_GP_ERR_INVALID_ARGUMENT for an unknown
 * or already-taken id.  No counterpart in the reference. */
int texo_gp_server_cancel(texo_gp_server *server, int64_t job);
int texo_gp_server_stats_get(texo_gp_server *server, texo_gp_server_stats *stats);
void texo_gp_server_destroy(texo_gp_server *server);

/* ---- GP input builders ------------------------------------------------- */

/* channel_orientation (crates/texo-pnr/src/routing_demand.rs:185-198): 1 for
 * a general horizontal wire name ("H##..." after the last '/'), 2 for a
 * vertical one ("V##..."), 0 otherwise.  Host only. */
int texo_gp_channel_orientation(const char *name);

/* routing_capacity_map (crates/texo-pnr/src/routing_demand.rs:201-239):
 * usable general-channel capacity per tile.  A wire (orientation code from
 * texo_gp_channel_orientation, tile point wire_xy[2w], capacity) counts when
 * at least one pip incident to it is not blocked (pip_blocked may be NULL);
 * per-tile sums saturate at UINT32_MAX.  Outputs horizontal/vertical
 * [width*height] row-major.  Runs on `device`; bit-exact. */
int texo_gp_routing_capacity_map(int32_t device, uint32_t width, uint32_t height, uint32_t n_wires,
                                 const uint8_t *wire_orientation, const uint32_t *wire_xy,
                                 const uint16_t *wire_capacity, uint32_t n_pips, const uint32_t *pip_from,
                                 const uint32_t *pip_to, const uint8_t *pip_blocked, uint32_t *horizontal,
                                 uint32_t *vertical);

/* routing_demand_map_with_capacity (crates/texo-pnr/src/routing_demand.rs:262-313):
 * bbox-uniform directional demand of a legal placement.  Nets are CSR over
 * pins at tile points pin_xy[2p] (driver first; 0xffffffff marks a cell with
 * no placement, skipped; a net whose driver is unplaced is skipped);
 * net_clock[n] = 1 excludes a clock-driven net.  Every tile of a net's
 * inclusive bbox receives dx/((dx+1)(dy+1)) horizontal and dy/(...) vertical
 * demand.  Sums are 2^-40 fixed point (within nets * 2^-41 of the
 * reference's FP64 sums in net order; run-to-run identical).  A placed pin
 * outside the grid is TEXO_GP_ERR_POINT_OUTSIDE_GRID. */
int texo_gp_routing_demand_map(int32_t device, uint32_t width, uint32_t height, uint32_t n_nets,
                               const uint32_t *net_pin_start, const uint32_t *pin_xy, const uint8_t *net_clock,
                               double *horizontal, double *vertical, uint64_t *included_nets,
                               uint64_t *excluded_clock_nets);

/* texo_gp_routing_demand_map without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code around HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results, bit for bit; the call's
 * staged inputs, sums and result live in a 64 MiB device array of the library
 * that each GPU holds for the life of the process, so a call whose layout
 * exceeds it (about 4 bytes per net, 8 per pin and 32 per tile) returns
 * TEXO_GP_ERR_OUT_OF_MEMORY, with the sizes in texo_gp_last_error().  Calls
 * are serialised per device inside one process; calls from several processes
 * on one GPU each use their own copy.  Synchronous, on the null stream. */
int texo_gp_routing_demand_map_noalloc(int32_t device, uint32_t width, uint32_t height, uint32_t n_nets,
                                       const uint32_t *net_pin_start, const uint32_t *pin_xy,
                                       const uint8_t *net_clock, double *horizontal, double *vertical,
                                       uint64_t *included_nets, uint64_t *excluded_clock_nets);

/* ecp5_timing_placement_weights (crates/texo-flow/src/lib.rs:3185-3201) with
 * ecp5_placement_criticality_weight (:3355-3390): per timing edge (net,
 * sink pin, path delay, domain worst path delay; delays are 128-bit values
 * given as (low, high) 64-bit words) the integer weight 1 + round(10 c^k) on
 * a 1/1024 grid; the maximum per (net, sink) is written in (net, sink) order
 * to out_*[<= n_edges].  Returns the number of entries.  Host only. */
uint64_t texo_gp_timing_placement_weights(uint64_t n_edges, const uint32_t *net, const uint32_t *sink,
                                          const uint64_t *path_delay_ps, const uint64_t *worst_path_delay_ps,
                                          uint32_t exponent, uint32_t *out_net, uint32_t *out_sink,
                                          uint64_t *out_weight);

/* Ecp5InitialPlacementAlgorithm (crates/texo-flow/src/lib.rs:300-324). */
#define TEXO_GP_ALGORITHM_IMPORTED 0u
#define TEXO_GP_ALGORITHM_ROUTABILITY_ELECTROSTATIC 1u
#define TEXO_GP_ALGORITHM_TIMING_ROUTABILITY_ELECTROSTATIC 2u

/* Provenance: the checkpoint's placement-model record
 * (checkpoint_placement_model, crates/texo-cli/src/checkpoint.rs:87-100) as
 * JSON, extended for the electrostatic algorithms with a "global_placement"
 * object naming this engine, its transform precision, density accumulation
 * and options, so that a checkpoint records how its placement was computed.
 * Returns 0, or -(bytes needed) when `capacity` is too small. */
int texo_gp_placement_model(uint32_t algorithm, uint32_t weight_exponent, const texo_gp_options *options,
                            char *out, size_t capacity);

/* ---- coarse quadratic seed ----------------------------------------------- */
/* The initial coordinates of analytical placement (analytical_place,
 * crates/texo-pnr/src/lib.rs:4042-4066): per axis, the bound-to-bound edges
 * at the device centre (AnalyticalHypergraph::linearize_axis,
 * analytical_placement.rs:283-356), the axis equations
 * (analytical_axis_equations, lib.rs:4464-4527) and their Jacobi-PCG solve
 * (solve_analytical_axis, lib.rs:4529-4556), both axes on `device`.
 * out_xy[2*n_units] interleaved x, y.  report->iterations = PCG iterations
 * of x plus y.  TEXO_GP_ERR_DID_NOT_CONVERGE where the reference's
 * solve_quadratic fails (err_index = iterations, err_index2 = axis,
 * err_kind = 1 for a breakdown, 0 for the dimension bound, err_value[0] =
 * relative residual). */
int texo_gp_quadratic_seed(const texo_gp_problem *problem, int32_t device, double *out_xy,
                           texo_gp_report *report);

/* solve_analytical_axis (lib.rs:4529-4556) of explicit axis equations on
 * `device`: diagonal[n], rhs[n], anchored[n], the off-diagonal terms as a
 * CSR in insertion order (adjacency_start[n+1], adjacency_other,
 * adjacency_weight), the starting solution initial[n] and the device centre
 * that floating components are re-centred on.  `iterations` receives the PCG
 * iterations, also on DID_NOT_CONVERGE. */
int texo_gp_solve_axis(int32_t device, uint32_t n, const double *diagonal, const uint32_t *adjacency_start,
                       const uint32_t *adjacency_other, const double *adjacency_weight, const double *rhs,
                       const uint8_t *anchored, const double *initial, uint32_t center, double *solution,
                       uint64_t *iterations);

/* ---- legalization auction ------------------------------------------------ */
/* The relaxed unit-to-object assignment of legalization::project
 * (assign_implicitly, crates/texo-pnr/src/legalization/auction.rs:123-253,
 * called from auction_units, legalization/mod.rs:781-833) over explicit
 * candidate domains, on `device`: bidder b's domain is
 * domain_object/domain_cost[domain_start[b] .. domain_start[b+1]) (objects
 * < n_objects, base costs <= max_base_cost, no object twice).  Every
 * unassigned bidder bids in the same round at the same prices (the
 * reference's bid: the best object's price + second - best + 1; each bidder's
 * private dummy at n * max + n + 1); an object takes its highest bid, ties to
 * the lowest bidder.  Once at most 512 bidders are active (a price war among
 * the last bidders of a congested region) they bid one at a time on the host
 * by the reference's rule, from the device's prices.  The result is
 * deterministic and carries the reference's guarantee (exact maximum
 * cardinality, base cost within n_bidders of the minimum of that
 * cardinality); where several assignments meet it, it need not be the
 * reference's sequential queue's.  assignment[n_bidders]: object
 * or -1 (unmatched); stats[4] = bids that took an object, candidates examined,
 * unmatched bidders, rounds.  Errors: TEXO_GP_ERR_AUCTION_* as the
 * reference's AuctionError. */
int texo_gp_auction(int32_t device, uint64_t n_bidders, uint64_t n_objects, uint64_t max_base_cost,
                    const uint32_t *domain_start, const uint32_t *domain_object, const uint64_t *domain_cost,
                    int64_t *assignment, uint64_t *stats);

/* ---- legalization repair ------------------------------------------------- */
/* A relaxed unit-to-object projection and the legality of every candidate,
 * as explicit footprints.  Candidate j of unit u is number
 * c = cand_start[u] + j; it occupies the objects
 * cand_object[cand_object_start[c] .. cand_object_start[c + 1]) (one per
 * cell of the unit; the first one's point orders the unit's candidates), and
 * it requires the shared-resource values (resource, value) =
 * (cand_shared[2k], cand_shared[2k + 1]) and puts the nets (wire, net) =
 * (cand_wire[2k], cand_wire[2k + 1]) on pin wires, k over
 * [cand_shared_start[c], cand_shared_start[c + 1]) and
 * [cand_wire_start[c], cand_wire_start[c + 1]).  A legal placement has every
 * object at most once, at most one value per shared resource and at most
 * wire_capacity[w] distinct nets per wire (the reference's
 * PlacementResourceUsage, legalization/mod.rs:1080-1128). */
typedef struct texo_gp_repair_problem {
  uint32_t width, height;          /* device grid; object points lie in it */
  uint64_t n_objects;
  const uint32_t *object_xy;       /* [2 * n_objects] */
  uint64_t n_wires;
  const uint32_t *wire_capacity;   /* [n_wires] */
  uint64_t n_units;
  const uint32_t *unit_cells;      /* [n_units] cells of the unit, >= 1 */
  const uint64_t *unit_first_cell; /* [n_units] id of the unit's first cell */
  const uint32_t *unit_target_xy;  /* [2 * n_units] the GP target */
  const uint8_t *unit_fixed;       /* [n_units] */
  const int64_t *preferred;        /* [n_units] candidate j of the unit, or -1 */
  const uint64_t *cand_start;      /* [n_units + 1] */
  const uint64_t *cand_object_start; /* [n_candidates + 1] */
  const uint32_t *cand_object;
  const uint64_t *cand_shared_start; /* [n_candidates + 1] */
  const uint64_t *cand_shared;
  const uint64_t *cand_wire_start;   /* [n_candidates + 1] */
  const uint64_t *cand_wire;
} texo_gp_repair_problem;

/* repair_relaxed (crates/texo-pnr/src/legalization/repair.rs:60-155), the
 * step of legalization::project after the auctions (mod.rs:377-386): the
 * conflicting part of the relaxed projection is set aside
 * (initial_pending_units, :322-425), the rest installed, and the set-aside
 * units placed by the reference's searches in its order: greedily on free
 * candidates, by single-blocker alternating paths, then by the complete
 * transactional search; each unit's candidates are tried preferred first,
 * then by Manhattan rings around its target (ChoiceCursor, :649-793).  On the
 * host (the search is serial).  placed[n_units] receives the candidate j of
 * every unit; stats[12] = initial pending, direct choices, direct moves,
 * direct fallbacks, augmenting roots, augmenting fallbacks, states visited,
 * failed states, max unplaced, max blocker branches, units evicted, choices
 * examined.  Returns 0, TEXO_GP_ERR_REPAIR_PROJECTION, or
 * TEXO_GP_ERR_REPAIR_NO_BEL with *error_unit = the first unit that was
 * pending; TEXO_GP_ERR_INVALID_ARGUMENT for malformed arrays. */
int texo_gp_repair(const texo_gp_repair_problem *problem, int64_t *placed, uint64_t *stats, int64_t *error_unit);

/* Last error message of the calling thread (static storage). */
const char *texo_gp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TEXO_GP_H */
