// gp_common.h — device-side layout shared by the kernels and the host session
// of the MI355X electrostatic global-placement library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace texo {

constexpr int MAXF = 4;                          // density fields per problem
constexpr double FX_SCALE = 1099511627776.0;     // 2^40 fixed-point density quantum
constexpr double FX_INV = 1.0 / 1099511627776.0;
constexpr uint32_t NO_UNIT = 0xffffffffu;        // fixed pin marker
constexpr double AUGMENTED_DENSITY_BETA = 2000.0;  // eplace.rs:125
constexpr int GF_STRIDE = 2 * MAXF + 2;  // filler gather partials: sums, kkt max, flag
constexpr int UG_STRIDE = 3;             // unit gradient partials: wa l1, kkt max, flag

// Reductions finished by k_finalize_eval for one objective evaluation.
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
  unsigned int pad;
};

// Reductions of one line-search candidate.
struct StepScalars {
  double gdelta;                   // g(y) . (c - y)
  double msq;                      // sum m (c - y)^2
  unsigned long long change_bits;  // max |c - origin|
  unsigned int changed;            // c != y anywhere
  unsigned int pad;
};

// Metric extremes used to reproduce the denominator checks of the line
// search (global_placement.rs:215-232).
struct MetricScalars {
  unsigned long long max_bits;
  unsigned long long min_bits;
  unsigned int nonfinite;
  unsigned int pad;
};

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
  a.w0 = 0.5 * (lo * lo);
  a.w1 = 0.75 - o * o;
  a.w2 = 0.5 * (hi * hi);
  a.d0 = o - 0.5;
  a.d1 = -2.0 * o;
  a.d2 = o + 0.5;
}

// --------------------------------------------------------------------------
// Kernel launchers (gp_kernels.hip / gp_dct.hip).  All take the session
// stream; none allocates.
// --------------------------------------------------------------------------
struct DensityArgs {
  int W, H, F;
  int n_members, n_fillers, n_units, n_density_units;
  const uint32_t *m_unit;      // member -> unit
  const double2 *m_off;        // member offset
  const double *m_charge;      // member charge
  const uint8_t *m_field;      // member field
  const uint32_t *du_unit;     // density unit -> unit
  const uint32_t *du_mstart;   // density unit -> first member
  const uint8_t *fl_field;     // filler field
  const double *fl_charge;     // filler charge
  const long long *base_fx;    // [F*G] -smoothed capacity, fixed point
  long long *acc_m;            // [F*G] members
  long long *acc_f;            // [F*G] fillers
};

void launch_zero_eval(EvalScalars *scalars, hipStream_t s);
void launch_deposit(const DensityArgs &a, const double *coords, EvalScalars *scalars, hipStream_t s);
void launch_density_convert(const DensityArgs &a, float *rho, double *density_out, EvalScalars *scalars,
                            hipStream_t s);

struct GatherArgs {
  const float *phi;               // [F*G]
  const double *energy_partials;  // [nb_energy * MAXF]
  int nb_energy;
  double lambda[MAXF];
  double scale[MAXF];
};

// Field gradients: per density unit writes coef_f * G_f into gfu[f*DU+du]
// (raw G_f into raw_units when non-null); per filler writes the combined
// gradient into grad[2*(U+j)] (raw q*g into raw_fillers when non-null).
void launch_gather(const DensityArgs &a, const GatherArgs &g, const double *coords, double2 *gfu,
                   double *raw_units, double *raw_fillers, double *grad, const double *metric,
                   const double *lower, const double *upper, double *partials_units, double *partials_fillers,
                   EvalScalars *scalars, int *nb_units, int *nb_fillers, hipStream_t s);

struct WaArgs {
  int n_units, n_pins_total, n_small, n_large;
  const uint32_t *small_nets, *large_nets;  // net ids
  const uint32_t *net_start;                // [nets+1] incl. arcs as 2-pin nets
  const double *net_weight;                 // 1.0 or arc extra weight
  const uint32_t *pin_unit;                 // NO_UNIT for fixed pins
  const double2 *pin_xy;                    // offset, or fixed coordinate
  double2 *dpin;                            // per-pin weighted derivative
  const uint32_t *u_pstart;                 // unit -> incident movable pins
  const uint32_t *u_pins;
  uint32_t first_arc_pin;                   // pins >= this belong to timed arcs
  const int32_t *unit_du;                   // unit -> density unit or -1
};

void launch_wa(const WaArgs &a, const double *coords, double gamma, double *partials, int *nb, hipStream_t s);
void launch_unit_gradient(const WaArgs &a, int F, int DU, const double2 *gfu, const double *coords,
                          const double *metric, const double *lower, const double *upper, double *grad,
                          double *wa_grad_out, double *partials, EvalScalars *scalars, int *nb, hipStream_t s);

struct FinalizeArgs {
  const double *wa_partials;
  int nb_wa;
  const double *energy_partials;
  int nb_energy;
  const double *gu_partials;  // [nb_gu][2*MAXF]
  int nb_gu;
  const double *gf_partials;  // [nb_gf][2*MAXF]
  int nb_gf;
  const double *ug_partials;  // [nb_ug]
  int nb_ug;
};
void launch_finalize_eval(const FinalizeArgs &a, EvalScalars *scalars, hipStream_t s);

void launch_preconditioner(int n_units, int n_fillers, int F, int DU, const double *wl_diag,
                           const int32_t *unit_du, const double *du_fcharge, const uint8_t *fl_field,
                           const double *fl_charge, const double *lambda, double *metric, double *partials,
                           MetricScalars *ms, hipStream_t s);
void launch_extrapolate(size_t n, const double *x, const double *xp, const double *lower, const double *upper,
                        double beta, double *y, unsigned int *differs, hipStream_t s);
void launch_candidate(size_t n, const double *y, const double *gy, const double *metric, double lipschitz,
                      const double *lower, const double *upper, const double *origin, double *cand,
                      double *partials, StepScalars *step, hipStream_t s);
void launch_zero_step(StepScalars *step, unsigned int *flag, hipStream_t s);

// RUDY (eplace.rs:856-937) over net bounding boxes.
// `fx` is [2G] scratch, `out` receives horizontal[G] then vertical[G].
void launch_rudy(int n_nets, const uint32_t *net_start, const uint32_t *pin_unit, const double2 *pin_xy,
                 const double *coords, int W, int H, long long *fx, double *out, unsigned int *error, hipStream_t s);
void launch_round_targets(int n_units, const double *coords, int W, int H, uint32_t *out, hipStream_t s);

// Poisson transform (gp_dct.hip).
struct DctPlan {
  int W, H, F;
  int precision;  // texo_gp_dct_precision
  // FP32 bases: C_W^T, C_H^T, C_H, C_W (row-major, K x N)
  const float *b32[4];
  // BF16 hi/lo and FP16 splits of the same bases
  const uint16_t *bhi[4];
  const uint16_t *blo[4];
  const uint16_t *bh16[4];
  // transposed (N x K) BF16 hi/lo bases for the register-fragment kernels
  const uint16_t *bthi[4];
  const uint16_t *btlo[4];
  const float *inv_lambda;  // [H*W], layout ky*W+kx
  float *t0, *t1;           // [F*G] scratch
};
// rho [F*G] -> phi [F*G]; energy partials written as [nb * MAXF].
void launch_poisson(const DctPlan &p, const float *rho, float *phi, double *energy_partials, int *nb_energy,
                    hipStream_t s);
int poisson_energy_blocks(const DctPlan &p);

// Kernel timing hook: the session brackets launches of one named kernel.
struct KernelTimer;
void timer_begin(KernelTimer *t, const char *name, hipStream_t s);
void timer_end(KernelTimer *t, const char *name, hipStream_t s);
extern thread_local KernelTimer *g_timer;

}  // namespace texo
