// gp_common.h — device-side layout shared by the kernels and the host engine
// of the MI355X electrostatic global-placement library.
//
// Every kernel works on a table of tasks.  A task is one objective evaluation
// (one job, one working slot, one coordinate vector); a launch covers every
// task of a round through a block map (block -> task, local block), so one
// launch evaluates one job or a whole batch of jobs with the same code.  No
// block ever mixes two tasks, so each task's reductions are independent of
// what else is in the batch: a job's results are bit-identical alone and in a
// mix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace texo {

constexpr int MAXF = 4;                            // density fields per job
constexpr double FX_SCALE = 1099511627776.0;       // 2^40 fixed-point density quantum
constexpr double FX_INV = 1.0 / 1099511627776.0;
constexpr uint32_t NO_UNIT = 0xffffffffu;          // fixed pin marker
constexpr double AUGMENTED_DENSITY_BETA = 2000.0;  // eplace.rs:125
constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// Lane-group reductions on the DPP path (no LDS traffic, unlike ds_bpermute
// shuffles).  Inside a row of 16 lanes four data-parallel moves pair lanes
// symmetrically: quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1]
// (lane ^ 2), row_half_mirror (lane i <-> 7 - i within 8) and row_mirror
// (lane i <-> 15 - i within 16); after log2(G) of them every lane of an
// aligned group of G <= 16 lanes holds the same bits (each step combines the
// same two values, and the operations are commutative).  Groups of 32 and 64
// combine the rows' lane-0 values, read with v_readlane, in row order.  Every
// lane of the wavefront must be active.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ unsigned int dpp_u32(unsigned int v) {
  return (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL, class T>
__device__ __forceinline__ T dpp_move(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp_u32<CTRL>(__builtin_bit_cast(unsigned int, v)));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long lo = dpp_u32<CTRL>((unsigned int)b), hi = dpp_u32<CTRL>((unsigned int)(b >> 32));
    return __builtin_bit_cast(T, (hi << 32) | lo);
  }
}
template <class T>
__device__ __forceinline__ T read_lane(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, (unsigned int)__builtin_amdgcn_readlane((int)__builtin_bit_cast(unsigned int, v), lane));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)b, lane);
    const unsigned long long hi = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(b >> 32), lane);
    return __builtin_bit_cast(T, (hi << 32) | lo);
  }
}
template <int G, class T, class Op>
__device__ __forceinline__ T lane_reduce(T v, Op op) {
  static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "group of 2^k lanes");
  if constexpr (G >= 2) v = op(v, dpp_move<0xB1>(v));
  if constexpr (G >= 4) v = op(v, dpp_move<0x4E>(v));
  if constexpr (G >= 8) v = op(v, dpp_move<0x141>(v));
  if constexpr (G >= 16) v = op(v, dpp_move<0x140>(v));
  if constexpr (G == 32) {
    const T r01 = op(read_lane(v, 0), read_lane(v, 16)), r23 = op(read_lane(v, 32), read_lane(v, 48));
    v = (threadIdx.x & 32) ? r23 : r01;
  } else if constexpr (G == 64) {
    v = op(op(op(read_lane(v, 0), read_lane(v, 16)), read_lane(v, 32)), read_lane(v, 48));
  }
  return v;
}
struct OpAdd {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMaxF {
  __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
  __device__ __forceinline__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct OpMinF {
  __device__ __forceinline__ double operator()(double a, double b) const { return fmin(a, b); }
};
struct OpMaxI {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpMinI {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? a : b; }
};
constexpr int GF_STRIDE = 2 * MAXF + 2;  // filler partials: force l1, raw l1 per field, kkt max, flag
constexpr int UG_STRIDE = 2 * MAXF + 3;  // unit partials: wa l1, force l1, raw l1 per field, kkt max, flag

// Block plan of one launch: task t owns logical blocks [start[t], start[t+1]).
// Workgroups are dispatched to the eight XCDs round-robin; with per > 0 the
// grid has 8 * per blocks and XCD x runs the logical slice [x per, (x+1) per),
// so a slice holds one job (or one spatial part of one) and each XCD's L2
// caches that job's arrays.  Only speed depends on the dispatch order.
// Per-element (x, y) offsets, stored as FP64 pairs or, when every value of
// the job is exact in FP32, as FP32 pairs (half the bytes, the same values).
struct OffArr {
  const double2 *d;
  const float2 *f;
  __host__ __device__ __forceinline__ double2 operator[](size_t i) const {
    if (f) {
      const float2 v = f[i];
      return make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
    }
    return d[i];
  }
};

struct BlockPlan {
  const int *start;  // [n_tasks + 1]
  int n_tasks, total, per;
};
// WA lane groups: a net of D pins gets the power-of-two number of lanes that
// gives each lane at most WA_PINS_PER_LANE pins (one lane up to that many);
// nets of more than WA_LOOP_PINS pins get a wavefront that loops.
constexpr int WA_PINS_PER_LANE = 4;
constexpr int WA_LOOP_PINS = 64 * WA_PINS_PER_LANE;
constexpr int WA_MAX_SEGMENTS = 96;
// A WA segment: consecutive nets of one class with one lane-group size
// (`lanes`, 0: one looping wavefront per net), starting at wavefront
// `first_wave`.  `degree` > 0: every net of the segment has that many pins,
// the first at `first_pin`, so a lane finds its pins without loading the net
// starts; 0: the degrees differ.
struct WaSeg {
  int first_net, nets, lanes, first_wave, first_pin, degree, pad0, pad1;
};
__host__ __device__ inline int wa_lanes(uint32_t pins) {
  if (pins > (uint32_t)WA_LOOP_PINS) return 0;
  int g = 1;
  while ((uint32_t)g * WA_PINS_PER_LANE < pins) g <<= 1;
  return g;
}

constexpr int XCDS = 8;
__host__ __device__ inline int plan_grid(const BlockPlan &p) { return p.per > 0 ? XCDS * p.per : p.total; }
__device__ __forceinline__ bool plan_block_at(const BlockPlan &p, int b, int2 &bm) {
  const int lb = p.per > 0 ? (b % XCDS) * p.per + b / XCDS : b;
  if (lb >= p.total) return false;
  int lo = 0, hi = p.n_tasks - 1;  // last task whose first block is <= lb
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.start[mid] <= lb)
      lo = mid;
    else
      hi = mid - 1;
  }
  bm = make_int2(lo, lb - p.start[lo]);
  return true;
}
__device__ __forceinline__ bool plan_block(const BlockPlan &p, int2 &bm) { return plan_block_at(p, blockIdx.x, bm); }

// Reductions finished by k_finalize for one objective evaluation.
struct EvalScalars {
  double wa_value;              // WA wirelength incl. timed arcs
  double wa_l1;                 // l1 norm of the WA gradient
  double energy[MAXF];          // 1/2 <rho_bar, phi> per field
  double force_l1[MAXF];        // sum q (|gx| + |gy|) per field (eq. 27)
  double raw_l1[MAXF];          // l1 norm of the raw field gradient
  long long over_fx[MAXF];      // sum max(rho_phys, 0), fixed point
  long long net_fx[MAXF];       // sum rho, fixed point
  unsigned long long kkt_bits;  // max metric KKT component (bits of a double >= 0)
  unsigned int nonfinite;       // any non-finite gradient entry
  unsigned int moved;           // charges deposited outside their tile's window
  unsigned int moved_acc;       // (counted by k_deposit, moved over by k_finalize)
  // a line-search trial whose gradient blocks did not run: rejected by the
  // acceptance test with its gradient bounded (the gradient sums are zero)
  unsigned int gradient_skipped;
};

// Reductions of one line-search candidate and one extrapolation.
struct StepScalars {
  double gdelta;                   // g(y) . (c - y)
  double msq;                      // sum m (c - y)^2
  unsigned long long change_bits;  // max |c - origin|
  unsigned int changed;            // c != y anywhere
  unsigned int differs;            // y != origin anywhere
};

// Metric extremes reproducing the line search's denominator checks
// (global_placement.rs:215-232).
struct MetricScalars {
  unsigned long long max_bits;
  unsigned long long min_bits;
  unsigned int nonfinite;
  unsigned int pad;
};

constexpr int SORT_TILE_SHIFT = 4;  // charges are binned by 16x16-bin tiles
constexpr int DEP_TILE = 1 << SORT_TILE_SHIFT;
// A tile's deposit window reaches DEP_HALO bins past the tile: the stencil
// (one bin) plus one bin of movement since the charges were last sorted
// (charges that moved further take the global-atomic path).
constexpr int DEP_HALO = 2;
constexpr int DEP_WIN = DEP_TILE + 2 * DEP_HALO;
constexpr int DEP_WIN2 = DEP_WIN * DEP_WIN;

// Centre bin of the quadratic B-spline stencil (electrostatic_placement.rs:950-976);
// a NaN coordinate lands on the last bin, as in bspline_axis.
__device__ __forceinline__ int bin_center(int length, double coordinate) {
  const double maximum = static_cast<double>(length - 1);
  const double c = coordinate < 0.0 ? 0.0 : (coordinate > maximum ? maximum : coordinate);
  double cc = floor(c + 0.5);
  cc = cc < maximum ? cc : maximum;
  return static_cast<int>(cc);
}

// Quadratic B-spline axis stencil (electrostatic_placement.rs:950-976).
struct Axis {
  int i0, i1, i2;
  double w0, w1, w2;
  double d0, d1, d2;
};

__device__ __forceinline__ void bspline_axis(int length, double coordinate, Axis &a) {
  const double maximum = static_cast<double>(length - 1);
  double c = coordinate < 0.0 ? 0.0 : (coordinate > maximum ? maximum : coordinate);
  double cc = floor(c + 0.5);
  cc = cc < maximum ? cc : maximum;
  const int center = static_cast<int>(cc);
  const double o = c - cc;
  a.i0 = center > 0 ? center - 1 : 0;
  a.i1 = center;
  a.i2 = center + 1 < length - 1 ? center + 1 : length - 1;
  const double lo = 0.5 - o, hi = 0.5 + o;
  // 0.5 * (0.5 -/+ o).powi(2): the square is rounded before the halving
  a.w0 = 0.5 * (lo * lo);
  a.w1 = 0.75 - o * o;
  a.w2 = 0.5 * (hi * hi);
  a.d0 = o - 0.5;
  a.d1 = -2.0 * o;
  a.d2 = o + 0.5;
}

// Device-resident static problem of one job (pointers into its arena).
struct JobStatic {
  int W, H, F;
  int n_members, n_fillers, n_units, n_density_units;
  const uint32_t *m_unit;     // member -> unit
  OffArr m_off;               // member offset
  const double *m_charge;     // member charge
  const uint8_t *m_field;     // member field
  const uint32_t *du_unit;    // density unit -> unit
  const uint32_t *du_mstart;  // density unit -> first member
  const uint8_t *fl_field;    // filler field
  const double *fl_charge;    // filler charge
  const long long *base_fx;   // [F*G] -smoothed capacity, fixed point
  const int32_t *unit_du;     // unit -> density unit or -1
  const uint32_t *u_mstart;   // unit -> first member ([n_units + 1]; empty for a unit without members)
  const uint32_t *perm;       // deposit order of the charges (members, then fillers)
  int tiles_x, n_keys;        // DEP_TILE x DEP_TILE-bin tiles of the charge sort
  const uint32_t *key_end;    // [n_keys] end of each tile's charges in deposit order
  // charges in deposit order: coordinate index (unit, or n_units + filler),
  // offset from it, charge, layer = 2 field + (1 for a filler)
  const uint32_t *s_ci;
  OffArr s_off;
  const double *s_q;
  const uint8_t *s_layer;
  const double *lower, *upper;
  // hypergraph: nets then timed arcs as weighted 2-pin nets; the nets in
  // WA segments: {first net, nets, lanes per net (0: one wavefront looping
  // over a long net), first wavefront}
  int n_wa_seg;
  const WaSeg *wa_seg;
  const uint32_t *net_start;
  const double *net_weight;
  const uint32_t *pin_unit;   // NO_UNIT for fixed pins
  OffArr pin_xy;              // offset, or fixed coordinate
  const uint32_t *u_pstart;   // unit -> incident movable pins
  const uint32_t *pin_slot;  // pin -> its entry in the unit-ordered derivative array (NO_UNIT: fixed pin)
  const uint32_t *u_nbase;   // unit -> number of its net pins (its timed-arc pins follow them)
  // transform bases (shared by every job with this grid)
  const float *b32[4];        // C_W^T, C_H^T, C_H, C_W (row-major K x N)
  const uint16_t *bthi[4];    // transposed BF16 hi/lo splits (N x K)
  const uint16_t *btlo[4];
  const uint16_t *fthi[4];    // transposed FP16 hi/lo splits of basis * 2^s (N x K)
  const uint16_t *ftlo[4];
  float f16_scale[4];         // 2^s of each FP16 basis
  const float *inv_lambda;    // [H*W], ky*W + kx
};

// One evaluation request: a job, one of its working slots, the coordinates.
struct Task {
  JobStatic j;
  // working slot
  long long *acc_m, *acc_f;    // [F*G]
  float *rho, *phi, *t0, *t1;  // [F*G]
  double *field_coef;          // [2*MAXF] lambda (1 + beta E/E0) per field, then E per field
  long long *partial;          // [n_keys][2F][DEP_WIN2] per-tile deposit windows
  // max |value| (float bits) per producing block, for the FP16-scaled
  // transform's input ranges: k_convert blocks [nb_conv][MAXF], then one
  // per 32x32 output tile of transform stages 0, 1, 2 [3][nb_tile]
  unsigned int *amax_part;
  long long *part_conv;        // [nb_conv][2*MAXF] over / net fixed-point sums per k_convert block
  uint32_t *tile_flag;         // [n_keys] 1 where the tile's window was written
  double2 *dpin;               // [movable pins] weighted WA derivative, grouped by unit
  double *part_wa, *part_energy, *part_gf, *part_ug;
  int nb_wa, nb_energy, nb_gf, nb_ug, nb_conv, nb_tile;
  // WA blocks [0, nb_wa_live) cover the nets with a movable pin; the blocks
  // after them the nets of fixed pins.  fixed_wa_mode: 0 no such nets, 1 this
  // task evaluates them and stores their value in *fixed_wa, 2 it adds the
  // stored value (same gamma)
  int nb_wa_live, fixed_wa_mode;
  double *fixed_wa;
  // request
  const double *coords;
  double *grad;
  const double *metric;  // nullptr: no KKT residual
  double gamma;
  int skip_density;      // the slot already holds this point's density field
  // value only: WA value and field energies, no gradient (no derivative
  // stores, no gradient blocks; the gradient and KKT scalars stay zero)
  int value_only;
  // A line-search trial (global_placement.rs:241-311): k_coefficients runs
  // its acceptance test (value against the majorizer and the origin) and
  // skips the gradient blocks of a candidate that fails it, when every
  // gradient entry is bounded by trial_wa_bound + sum over fields of
  // |coefficient| x 4 trial_charge[f] x max |phi_f| (finite, so the
  // reference's gradient check passes); the host takes the same decision
  // from the same scalars.
  int trial;
  double trial_y_value, trial_origin_value, trial_L, trial_wa_bound;
  double trial_charge[MAXF];  // sum of |charge| per field
  const StepScalars *trial_step;
  // max |phi| bits (NaN above infinity) per 32x32 tile of the last
  // transform stage, [F][phi_tiles]; null when the transform does not write them
  const unsigned int *phi_amax;
  int phi_tiles;
  double lambda[MAXF], scale[MAXF];
  EvalScalars *scalars;
  // parity outputs (nullptr in the hot loop)
  double *raw_units, *raw_fillers, *density_out, *wa_grad_out;
};

// Spatial sort of one job's charges by sort tile of their stencil centre
// (counting sort).  The order only steers which charges share a deposit
// block; the integer density sums do not depend on it.
struct SortTask {
  JobStatic j;
  const double *coords;
  uint32_t *count;        // [n_keys + 1]
  const uint32_t *order;  // the previous deposit order (walked so that a block's charges share tiles)
  uint32_t *perm;         // [members + fillers] the new order
  uint32_t *s_ci;   // the charges in the new order
  double2 *s_off;
  float2 *s_off32;  // instead of s_off for a job with FP32 offsets
  double *s_q;
  uint8_t *s_layer;
};
struct SortMaps {
  BlockPlan charges, keys;
  const int *list;  // one task per block of the scan
  int n_list;
};
void launch_sort(const SortTask *tasks, const SortMaps &m, hipStream_t s);

// One optimizer elementwise request of one job.
struct OptTask {
  size_t n;  // coordinates
  int n_units, n_fillers, F, DU;
  const double *lower, *upper;
  // preconditioner (eplace.rs:1046-1100)
  const double *wl_diag, *du_fcharge, *fl_charge;
  const int32_t *unit_du;
  const uint8_t *fl_field;
  double lambda[MAXF];
  double *metric, *part_metric;
  MetricScalars *metric_scalars;
  // extrapolation: y = clamp(x + beta (x - xp))
  const double *x, *xp;
  double *y;
  double beta;
  // candidate: c = clamp(yp - gy / (L m)), reductions against yp and origin
  const double *yp, *gy, *origin;
  double *cand;
  double lipschitz;
  double *part_step;  // [blocks][4]: candidate sums, change max, changed flag; extrapolation flag in [3]
  StepScalars *step;
  int do_precond, do_extrapolate, do_candidate;  // the work this round asks of the job
};

struct Maps {
  BlockPlan deposit, convert, wa, gradient;
  int n_tasks;
  int n_density;   // tasks whose density is computed (not skipped)
  int n_gradient;  // tasks with a gradient (not value only)
  // WA work in full evaluations: a value-only task counts its share of a
  // full evaluation's algorithmic bytes (no derivative stores)
  double wa_evaluations;
};

void launch_density(const Task *tasks, const Maps &m, hipStream_t s);
void launch_gather(const Task *tasks, const Maps &m, hipStream_t s);  // field coefficients
void launch_wa(const Task *tasks, const Maps &m, hipStream_t s);
void launch_combine(const Task *tasks, const Maps &m, hipStream_t s);

// Optimizer maps; each finalize list holds one task index per block.
struct OptMaps {
  BlockPlan precond, extrapolate, candidate;
  int n_tasks;  // one finalize block per task
};
// `passes`: the C-long vectors the launch reads or writes, summed over its
// tasks (preconditioner 2, extrapolation 5, candidate 7), for kernel timing.
void launch_optimizer(const OptTask *tasks, const OptMaps &m, hipStream_t s, int passes = 0);
// The optimizer's scalars (read by the host only): may run beside the density chain.
void launch_optimizer_finalize(const OptTask *tasks, const OptMaps &m, hipStream_t s);

// RUDY (eplace.rs:856-937): `fx` is [2G] scratch, `out` receives
// horizontal[G] then vertical[G].
void launch_rudy(int n_nets, const uint32_t *net_start, const uint32_t *pin_unit, OffArr pin_xy,
                 const double *coords, int W, int H, long long *fx, double *out, unsigned int *error, hipStream_t s);

// Continuous bounding-box HPWL of nets [0, n_nets) at `coords`, FP64 on the
// device; `large` lists (ascending) the nets of more than 32 pins, which get
// a wavefront each.  `part` is hpwl_blocks(n_nets, n_large) doubles of
// scratch, `out` one double.
int hpwl_blocks(int n_nets, int n_large);
void launch_hpwl(int n_nets, const uint32_t *net_start, const uint32_t *pin_unit, OffArr pin_xy,
                 const double *coords, const uint32_t *large, int n_large, double *part, double *out, hipStream_t s);

// Poisson transform (gp_dct.hip) of every task of one grid size.  map_h has
// one entry per 32-row block of the F*H-row stages, map_w of the F*W-row
// stages: (task, local row block).
// One transform launch group: tasks of one precision and shape class.  The
// plans count, per task, 32-row blocks x 32-column tiles (h: the stages with
// F*H rows, w: those with F*W rows), and the same with 128-column tiles when
// every task of the group fits the wide FP16-scaled kernel.
struct DctLaunch {
  BlockPlan h, w, h_wide, w_wide;
  int precision;
  bool fast;     // every task's W and H are multiples of 32
  bool wide_ok;  // split-MFMA precision, and W and H multiples of 128, at most 512
  int kmax;      // largest W or H of the launch's tasks
  int n_tasks;   // tasks with tiles in this launch
};
void launch_poisson(const Task *tasks, const DctLaunch &d, hipStream_t s);
int poisson_energy_blocks(int W, int H, int F);  // per task
bool poisson_fast_shape(int W, int H);
bool poisson_wide_ok(int precision, int W, int H);

// Kernel timing hook: the engine brackets launches of one named kernel.
struct KernelTimer;
void timer_begin(KernelTimer *t, const char *name, hipStream_t s);
void timer_end(KernelTimer *t, const char *name, hipStream_t s, double tasks);  // tasks: evaluations the launch covers
extern thread_local KernelTimer *g_timer;

}  // namespace texo
