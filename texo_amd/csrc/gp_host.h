// gp_host.h — host-side types of the placement engine: jobs (one placement
// problem each, with its device arena and optimizer state), the transform
// bases shared per grid size, and the lockstep round engine.
#pragma once

#include "gp_common.h"
#include "../../include/texo_gp.h"

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace texo {

extern thread_local std::string g_last_error;

// Typed failure carried to the ABI boundary (status + EplaceError payload).
struct Failure {
  int status;
  int64_t index = -1, index2 = -1;
  uint32_t kind = 0, x = 0, y = 0;
  double v0 = 0.0, v1 = 0.0;
};
[[noreturn]] void fail(int status, int64_t index = -1, int64_t index2 = -1);
void hip_check(hipError_t e, const char *what);
#define HIP(x) ::texo::hip_check((x), #x)

bool is_lut(uint32_t kind);
double overflow_target(uint32_t kind);
uint32_t rounded_coordinate(double v, uint32_t extent);
double clampd(double v, double lo, double hi);

struct KernelTimer {
  std::string name;
  bool enabled = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t used = 0;
  double tasks = 0.0;  // evaluations covered by the timed launches
  ~KernelTimer();
};

// Blocks kept for reuse once a job ends: device blocks for the job arenas
// (hipFree waits for the whole device) and pinned host blocks for the upload
// images (pinning is slow).  A serving loop that admits and retires jobs
// continuously then neither allocates nor frees after it has warmed up.  At
// most `keep_bytes` of unused blocks are held.  Thread-safe.
class BlockPool {
 public:
  BlockPool(size_t keep_bytes, bool pinned_host) : keep_(keep_bytes), host_(pinned_host) {}
  ~BlockPool();
  void *acquire(size_t bytes, size_t *capacity);
  void release(void *p, size_t capacity);

 private:
  void free_block(void *p);
  std::mutex mu_;
  std::multimap<size_t, void *> free_;  // capacity -> block
  size_t kept_ = 0, keep_;
  bool host_;
};

// Design-static device arrays shared by the jobs of one design (multi-start
// placements, repeated submissions): one device copy of the design block that
// all their kernels read, returned to its pool with the last job using it.
struct SharedDesign {
  std::shared_ptr<BlockPool> pool;
  void *base = nullptr;
  size_t capacity = 0;
  ~SharedDesign() {
    if (base && pool) pool->release(base, capacity);
  }
};

// One allocation carved into 256-byte aligned buffers.
class Arena {
 public:
  ~Arena() { release(); }
  // Plans a buffer of `count` elements; returns its byte offset in the arena.
  template <class T>
  size_t reserve(T *&p, size_t count) {
    const size_t at = planned_;
    plan_.push_back({reinterpret_cast<void **>(&p), at});
    planned_ += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
    return at;
  }
  size_t planned() const { return planned_; }
  // Allocates the planned bytes (from `pool` when given) and sets every
  // reserved pointer.
  void commit(std::shared_ptr<BlockPool> pool = nullptr);
  void release();  // returns the memory (to its pool); pointers dangle after
  char *base() const { return static_cast<char *>(base_); }

 private:
  struct Entry {
    void **slot;
    size_t offset;
  };
  std::vector<Entry> plan_;
  std::shared_ptr<BlockPool> pool_;
  void *base_ = nullptr;
  size_t planned_ = 0, capacity_ = 0;
};

struct Options {
  uint32_t precision = TEXO_GP_DCT_FP32;
  uint32_t fixed_iterations = 0;
  uint64_t max_iterations = 3000;  // eplace.rs:134
  double gamma_base = 8.0;         // eplace.rs:133
  bool trace = false;
  int device = 0;
  uint32_t inject = 0;
};

// DCT bases and 1/lambda for one (W, H, precision), shared by jobs.
struct Basis {
  int W = 0, H = 0;
  Arena arena;
  float *b32[4] = {};
  uint16_t *bthi[4] = {}, *btlo[4] = {};
  uint16_t *fthi[4] = {}, *ftlo[4] = {};
  float f16_scale[4] = {};
  float *inv_lambda = nullptr;
  void build(int W, int H);
};

struct Control {
  uint64_t clock, lsr, ce;
};

// Scalars of one job read back by the host after every round.
struct JobScalars {
  EvalScalars eval[4];  // working slots 0 and 1, the parity evaluation, working slot 2
  StepScalars step, step_spec;  // the trial candidate, the speculative next one
  MetricScalars metric;
};
// Working slots of a job: its current point's field, a trial's, and the
// speculative next trial's.
constexpr int WORKING_SLOTS = 3;
// The scalar record of working slot s (record 2 is the parity evaluation's).
inline int scalar_of_slot(int s) { return s == 2 ? 3 : s; }
// A round of at most this many jobs, or of at most this many charges
// (members + fillers) over its jobs, evaluates beside each trial candidate
// the candidate of the step length the trial doubles to when it is rejected
// (a single job or a few small ones leave most of the GPU idle; a batch
// fills it).
constexpr size_t SPECULATE_MAX_JOBS = 2;
constexpr size_t SPECULATE_MAX_CHARGES = 1u << 19;

struct AreaMember {
  uint32_t kind;
  double current, routability, clustering;
};
// clustering_areas (register_clustering.rs:17-81); false is the reference's None.
bool register_clustering(const std::vector<std::pair<uint32_t, std::pair<double, double>>> &regs,
                         const std::map<uint32_t, Control> &controls, uint64_t movable, std::vector<double> &areas);
// The same on the GPU (gp_cluster.hip): 1 done, 0 the reference's None, -1
// a group with more clock-enable classes than the kernel holds (use the host
// version).  Device memory comes from `pool` when given.
int register_clustering_device(const std::vector<std::pair<uint32_t, std::pair<double, double>>> &regs,
                               const std::map<uint32_t, Control> &controls, uint64_t movable,
                               std::vector<double> &areas, BlockPool *pool, hipStream_t s);
// adjust_instance_areas (instance_area.rs:157-372); returns requires_optimizer_reset.
bool adjust_areas(const std::vector<AreaMember> &members, const std::vector<std::pair<uint32_t, double>> &fillers,
                  std::vector<double> &member_out, std::vector<double> &filler_out);
struct Job;
// Jobs per launch of the batched best-checkpoint rounding.
constexpr int TARGET_BATCH = 64;
void launch_round_targets(const std::vector<Job *> &jobs, hipStream_t s);
// Rounds 2*units interleaved coordinates into `out` (one job, one launch).
void launch_round_targets_to(const double *coords, uint32_t *out, size_t units, uint32_t W, uint32_t H,
                             hipStream_t s);
// The typed failure of the exception being handled: a Failure as thrown,
// std::bad_alloc as OUT_OF_MEMORY, anything else as DEVICE (with its message
// in the thread's last error).  Call only inside a catch block.
Failure current_failure();
void launch_collect_scalars(const int *index, int n, const JobScalars *pool, JobScalars *out, hipStream_t s);

// OriginFull: the gradient of an origin whose value was evaluated alone,
// before a decision that reads it (y at the origin, a restart).
enum class Phase { Probe, Init, IterStart, OriginFull, Trial, RoundEnd, Paused, Done, Failed };

// From this many charges (members + fillers) in flight the jobs are split
// into two groups whose rounds overlap on the GPU (each context has its own
// streams) while the host plans the other group; below it one group's
// launches are too small to gain from the split.
constexpr size_t PIPELINE_MIN_CHARGES = 1u << 20;
// Round contexts: job groups whose rounds overlap.  Two are fastest on the
// 16-job lockstep batch and the serving stream (three and four groups make
// each launch smaller than the overlap gains).
constexpr int ROUND_CONTEXTS = 2;

// Unused device blocks / pinned image blocks an engine keeps for the jobs it
// admits next.
constexpr size_t BLOCK_POOL_KEEP = size_t(16) << 30;
constexpr size_t IMAGE_POOL_KEEP = size_t(4) << 30;

struct Job {
  // ---- problem (host) ----
  Options opt;
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
  std::vector<uint32_t> u_mstart;  // unit -> first member (JobStatic::u_mstart)
  std::map<uint32_t, std::vector<uint32_t>> capacity;  // after fixed occupancy
  std::vector<size_t> available;                       // per field
  std::vector<double> real_area, gamma_weight, wl_diag;
  std::vector<double> field_charge;  // sum of |charge| of the members and fillers per field
  uint64_t movable_instance_count = 0;
  bool empty = false, best_effort = false, has_routing = false;
  std::vector<uint32_t> routing_h, routing_v;
  std::map<uint32_t, Control> controls;
  size_t n_nets = 0, n_arcs = 0, n_net_total = 0, n_pins_total = 0, first_arc_pin = 0;
  std::vector<uint32_t> net_start, pin_unit, u_pstart, u_pins, pin_slot, u_nbase;
  std::vector<WaSeg> wa_seg;  // WA segments (JobStatic::wa_seg)
  int wa_waves = 0;          // wavefronts of one WA evaluation
  int wa_waves_live = 0;     // ... of its nets with a movable pin (a multiple of BLOCK / 64)
  size_t n_live_nets = 0;
  // WA value of the nets of fixed pins: held on the device for one gamma
  double *d_fixed_wa = nullptr;
  // final HPWL on the device: block sums of the nets, then the total; the
  // nets of more than 32 pins (ascending) take a wavefront each
  double *d_hpwl = nullptr;
  std::vector<uint32_t> hpwl_large;
  uint32_t *d_hpwl_large = nullptr;
  bool fixed_wa_valid = false;
  double fixed_wa_gamma = 0.0;
  std::vector<double> pin_xy, net_weight;
  std::vector<double> h_du_fcharge;  // charge per field and density unit
  // Upload image: the bytes of the arena's leading block of arrays that
  // start with data, built by load() on the submitting thread (in a pinned
  // block of `image_pool` when given) and copied to the device in one piece.
  char *image = nullptr;
  size_t image_bytes = 0, image_cap = 0, zero_begin = 0, zero_end = 0;
  // the design block: the leading design_bytes of the arena (arrays fixed by
  // the problem alone) and its 128-bit fingerprint
  size_t design_bytes = 0;
  uint64_t design_hash[2] = {0, 0};
  std::shared_ptr<SharedDesign> design;  // set when the job reads a shared copy
  std::shared_ptr<BlockPool> image_pool;
  std::unique_ptr<char[]> image_heap;

  ~Job() { release_image(); }
  void load(const texo_gp_problem *p, const texo_gp_options *o,
            const std::shared_ptr<BlockPool> &image_pool = nullptr);
  size_t device_bytes() const { return arena.planned(); }  // for admission
  void release_image();  // once the upload has completed

  // ---- device ----
  Arena arena;
  const Basis *basis = nullptr;
  JobStatic js{};
  long long *d_base = nullptr;
  uint32_t *d_m_unit = nullptr, *d_du_unit = nullptr, *d_du_mstart = nullptr;
  double2 *d_m_off = nullptr, *d_pin_xy = nullptr;
  float2 *d_m_off32 = nullptr, *d_pin_xy32 = nullptr;  // instead, when off32
  bool off32 = false;  // every member offset and pin offset / fixed coordinate is exact in FP32
  double *d_m_charge = nullptr, *d_fl_charge = nullptr, *d_du_fcharge = nullptr;
  uint8_t *d_m_field = nullptr, *d_fl_field = nullptr;
  uint32_t *d_net_start = nullptr, *d_pin_unit = nullptr;
  WaSeg *d_wa_seg = nullptr;
  uint32_t *d_u_pstart = nullptr, *d_pin_slot = nullptr, *d_u_nbase = nullptr;
  double *d_net_weight = nullptr, *d_wl_diag = nullptr;
  int32_t *d_unit_du = nullptr;
  uint32_t *d_u_mstart = nullptr;
  double *d_lower = nullptr, *d_upper = nullptr, *d_metric = nullptr, *d_part_metric = nullptr, *d_part_step = nullptr,
         *d_part_step_spec = nullptr;
  double *d_vec[5] = {};   // coordinate buffers (rotated by index)
  double *d_grad[4] = {};  // origin, y, candidate, speculative candidate gradients
  uint32_t *d_best = nullptr;
  struct Slot {
    long long *acc_m = nullptr, *acc_f = nullptr;
    float *rho = nullptr, *phi = nullptr, *t0 = nullptr, *t1 = nullptr;
    double2 *dpin = nullptr;
    double *field_coef = nullptr;
    long long *partial = nullptr;
    unsigned int *amax_part = nullptr;
    long long *part_conv = nullptr;
    uint32_t *tile_flag = nullptr;
    double *part_wa = nullptr, *part_energy = nullptr, *part_gf = nullptr, *part_ug = nullptr;
  } slot[WORKING_SLOTS];
  double *d_rudy = nullptr;
  long long *d_rudy_fx = nullptr;
  unsigned int *d_rudy_err = nullptr;
  double *d_raw_units = nullptr, *d_raw_fillers = nullptr, *d_density_out = nullptr, *d_wa_grad = nullptr;
  double *d_eval_coords = nullptr;
  uint32_t *d_perm = nullptr, *d_perm_next = nullptr, *d_sort_count = nullptr, *d_s_ci = nullptr;
  double2 *d_s_off = nullptr;
  float2 *d_s_off32 = nullptr;
  double *d_s_q = nullptr;
  uint8_t *d_s_layer = nullptr;
  int tiles_x = 0, n_keys = 0;
  int nb_wa = 0, nb_wa_live = 0, nb_energy = 0, nb_gf = 0, nb_ug = 0, nb_conv = 0, nb_tile = 0;

  // Allocates the arena (from `pool` when given) and enqueues the image copy
  // and the accumulator clear on `s`; the image stays alive until
  // release_image(), so nothing waits here.
  void upload(const Basis *b, const std::shared_ptr<BlockPool> &pool, hipStream_t s);
  void use_design(const std::shared_ptr<SharedDesign> &sd);  // read the design block from sd
  void build_js();
  // New member / filler charges: host state, then the device copies on `s`
  // (waited for before returning).
  void apply_charges(const std::vector<double> &member_areas, const std::vector<double> &filler_areas, hipStream_t s);
  bool sorted = false;  // charges ordered by tile at least once
  // fixed_nets: the task evaluates the nets of fixed pins too (their value
  // at `gamma` is then held for later tasks at the same gamma)
  // value_only: the WA value and field energies only (skip must be set)
  Task task(int slot, const double *coords, double *grad, const double *metric, double gamma, bool skip,
            EvalScalars *scalars, bool fixed_nets, bool value_only = false) const;
  // next sort, at `coords` (null: the current iterate); the new order
  // becomes the job's order
  SortTask sort_task(const double *coords = nullptr);

  // ---- engine bookkeeping ----
  int64_t id = 0;
  int pool_index = -1;
  JobScalars *d_scalars = nullptr;  // in the engine pool
  const JobScalars *h_scalars = nullptr;

  // ---- optimizer / schedule state (eplace.rs:439-582, global_placement.rs:121-335) ----
  Phase phase = Phase::Probe;
  int status = 0;
  Failure failure{};
  uint64_t iterations = 0, evaluations = 0, trials = 0, area_rounds = 0;
  uint64_t pause_at = UINT64_MAX;  // iterate(): stop when iterations reach this
  bool no_stop = false;            // fixed-iteration or iterate mode: stop rules off
  uint64_t fixed_limit = UINT64_MAX;
  bool fixed_mode = false, area_ready = false, exhausted = false, started = false;
  bool has_prev_change = false, reset_after_area = false, has_fixed_scales = false;
  double prev_change = 0.0;
  size_t round_limit = 0, area_round = 0;
  uint64_t local_iterations = 0;
  std::vector<double> scales, mult, fixed_scales;
  double mstep = 0.0;
  EvalScalars current{};
  double last_wa = 0.0, gamma = 8.0, g0 = 8.0;
  bool has_best = false;
  const double *best_src = nullptr;  // coordinates of a new best, not yet rounded
  double best_max = 0.0, best_total = 0.0;
  bool want_targets = false;
  std::vector<uint32_t> prev_stationary;
  bool has_prev_stationary = false;
  int ix = 0, ixp = 1, iy = 2, ic = 3, ic2 = 4;
  int p_slot = 0;  // working slot of the current point's field; trials use the next two
  bool p_valid = false;
  double momentum = 1.0, lipschitz = 0.0;
  bool has_lipschitz = false;
  // one step's working state
  bool restart = false, has_wl = false, y_is_origin = true, may_reduce = true, y_evaluated = false;
  double wl = 0.0, L = 0.0, next_momentum = 1.0, beta = 0.0;
  EvalScalars origin_s{}, y_s{};
  double origin_value = 0.0, y_value = 0.0;
  // The origin evaluation of an iteration with momentum evaluates only the
  // value (the line search reads the origin's gradient only when y is the
  // origin or on a restart); origin_full says whether origin_s holds the
  // gradient scalars too.  held_y: y's evaluation, taken after an
  // OriginFull round that checks the origin's gradient first.
  bool origin_value_only = false, origin_full = true, held_y = false, held_y_has = false;
  EvalScalars held_y_s{};
  // A bound on every WA gradient entry: each pin's derivative is at most
  // 3 (1 + 3 extent / gamma) times its net weight in magnitude, so a unit's
  // entry is at most 3 wa_bound_w (1 + 3 wa_bound_ext / gamma), with
  // wa_bound_w the sum over nets of weight x pins and wa_bound_ext the x
  // plus y extent of every pin position the bounds allow.
  double wa_bound_w = 0.0, wa_bound_ext = 0.0;
  bool wa_bound_ok = false;
  // WA bytes of a value-only evaluation per full evaluation's
  double wa_value_only_share = 1.0;
  // requests planned for the next round
  struct EvalReq {
    int slot;
    const double *coords;
    double *grad;
    bool metric;
    double gamma;
    bool skip;
    bool parity;
    int scalar_slot;
    bool value_only;
    int trial;  // 1 the line-search trial, 2 the speculative one (device acceptance test)
  };
  std::vector<EvalReq> evals;
  bool want_precond = false, want_extrapolate = false, want_candidate = false, want_sort = false;
  // the speculative trial of this round: its step length (twice the trial's)
  bool want_spec = false;
  double spec_L = 0.0;
  uint32_t moved = 0;  // most charges any density evaluation of the last round deposited outside their window

  // schedule helpers
  double overflow(const EvalScalars &s, int f) const;
  double combined_value(const EvalScalars &s) const;
  double adaptive_gamma(const EvalScalars &s) const;
  void validate_initial(const EvalScalars &s) const;
  void validate_fields(const EvalScalars &s) const;
  std::vector<double> subgradient(const EvalScalars &s) const;
  bool stop_reached(const EvalScalars &s, bool area_ready) const;
  std::pair<double, double> excess(const EvalScalars &s) const;
  size_t area_round_limit() const;
  void update_multipliers();
  // true when the gradient of an evaluation at the current point with the
  // value-only scalars `s` (energies at gamma) has only finite entries
  bool gradient_certainly_finite(const EvalScalars &s, double gamma) const;
  double hpwl(const std::vector<double> &x) const;
  // HPWL of the device coordinates `coords` into `out` (device or pinned
  // host memory), enqueued on `s` (a job with an arena)
  void enqueue_hpwl(const double *coords, double *out, hipStream_t s) const;
  void stage(const std::shared_ptr<BlockPool> &pool);
  void set_charges(const std::vector<double> &member_areas, const std::vector<double> &filler_areas);
  void begin_run(bool iterate_mode);
  void begin_round();
  void count_evaluation();
};

// Lockstep round engine: one launch per kernel family covers every active job.
class Engine {
 public:
  explicit Engine(int device, size_t pool_capacity = 256);
  ~Engine();
  int device() const { return device_; }
  // Compute stream, ordered after every upload enqueued so far.
  // The compute stream of context 0, ordered after the uploads and after
  // all work issued so far on the other contexts: host reads and set-up work
  // enqueued here see every job's latest state.
  hipStream_t stream();
  // The compute stream of context c (a server reads a group's results there).
  hipStream_t stream(int c);
  // Stream of uploads and result read-backs: work there waits for no round.
  hipStream_t io_stream() const { return io_stream_; }
  // Pinned host blocks for upload images (usable from any thread).
  const std::shared_ptr<BlockPool> &image_pool() const { return images_; }
  const Basis *basis(int W, int H);  // built on first use; thread-safe
  void attach(Job *job);  // enqueue the upload and reserve a scalar slot
  void detach(Job *job);  // release the slot (job memory is freed by the job)
  void sort_now(Job *job);  // sorts the job's charges by tile at its current coordinates
  // Runs one round for every job in `jobs` whose phase needs the device.
  void round(const std::vector<Job *> &jobs);
  // Pipelined rounds: launch() plans one round for `jobs` in context c
  // (< ROUND_CONTEXTS), issues it and returns (false when no job needs the
  // device); finish(c) waits for it and advances its jobs.  Each context has
  // its own streams, so with one job group per context the host plans one
  // group while the GPU runs the others' rounds, and the groups' rounds
  // overlap on the GPU.  A job stays in one context while it runs.
  bool launch(int c, const std::vector<Job *> &jobs);
  void finish(int c);
  bool pending(int c) const { return ctx_[c].pending; }
  // Seconds the last finish() spent waiting for the GPU to end the round.
  double last_wait_seconds() const { return last_wait_; }
  // Rounds replayed from a captured graph since the engine started.
  uint64_t graph_rounds() const { return graph_rounds_; }
  // Parity evaluation of one job at host coordinates.
  void evaluate_once(Job *job, const double *coords, double gamma, bool density_out);
  // propose_area_adjustment (eplace.rs:627-729) at device coordinates
  // `coords`: RUDY and register clustering on the GPU, utilisation,
  // routability areas and adjust_instance_areas on the host.
  struct AreaProposal {
    std::vector<double> member_areas, filler_areas;
    bool requires_reset = false;
    double total_relative_increase = 0.0;
  };
  AreaProposal propose_areas(Job *job, const double *coords, hipStream_t s);
  // Preconditioner of one job for `lambda` into job->d_metric (parity seam);
  // returns false when an entry is not finite or negative before the floor.
  bool precondition_once(Job *job, const std::vector<double> &lambda);
  KernelTimer timer;

 private:
  void plan(Job *job);
  void advance(Job *job);
  void area_round(Job *job);
  void await_uploads(int c);  // orders context c's compute stream after the enqueued uploads
  void share_design(Job *job);  // at attach: read the design block from a copy shared with its design's jobs
  // Designs seen by attached jobs: the first job of a design owns the
  // design block in its arena; the second promotes it to a shared copy.
  struct DesignEntry {
    uint64_t hash[2];
    size_t bytes;
    Job *owner;  // first job, until promoted or detached
    std::weak_ptr<SharedDesign> shared;
  };
  std::vector<DesignEntry> designs_;
  hipStream_t cur() const { return ctx_[cur_].stream; }  // stream of the context being planned or advanced
  int device_;
  hipStream_t io_stream_ = nullptr;
  hipEvent_t ev_upload_ = nullptr;
  int cur_ = 0;
  double last_wait_ = 0.0;
  std::shared_ptr<BlockPool> blocks_, images_;
  std::map<std::pair<int, int>, std::unique_ptr<Basis>> bases_;  // never erased while the engine lives
  std::mutex bases_mu_;
  JobScalars *d_pool_ = nullptr, *h_pool_ = nullptr;
  std::vector<bool> pool_used_;
  // Staging of one round: the task tables go up in one copy, the active
  // jobs' scalars come back in one copy (gathered into `d_compact`).
  struct RoundCtx {
    hipStream_t stream = nullptr, wa_stream = nullptr;  // compute stream; WA beside the density chain
    hipEvent_t fork = nullptr, join = nullptr, mark = nullptr;  // mark: joins this stream into context 0's
    bool upload_wait = false;  // uploads enqueued since the stream last waited for them
    char *h_stage = nullptr, *d_stage = nullptr;
    size_t stage_bytes = 0;
    JobScalars *d_compact = nullptr, *h_compact = nullptr;
    size_t compact_cap = 0;
    hipEvent_t done = nullptr;
    std::vector<Job *> active;
    bool pending = false;
    // captured launch sequences by their argument bytes, and how often each
    // sequence was issued before capture
    std::unordered_map<std::string, hipGraphExec_t> graphs;
    std::unordered_map<std::string, int> seen;
  };
  // A sequence is captured on its GRAPH_AFTER_ROUNDS-th occurrence; at most
  // GRAPH_CACHE graphs are kept per context.
  static constexpr int GRAPH_AFTER_ROUNDS = 3;
  static constexpr size_t GRAPH_CACHE = 64;
  bool use_graphs_ = false;  // TEXO_GP_GRAPHS=1 replays recurring rounds from captured graphs
  bool speculate_ = false;   // this round evaluates speculative trials (few jobs)
  // origin evaluations of iterations with momentum: 1 value only (default),
  // 0 full, 2 value only with every gradient check taken by an OriginFull
  // round (TEXO_GP_ORIGIN_VALUE_ONLY; results are the same in all three)
  int origin_mode_ = 1;
  // the gradient blocks of a rejected line-search trial with a bounded
  // gradient are skipped (TEXO_GP_SKIP_REJECTED=0 runs them; same results)
  bool skip_rejected_ = true;
  uint64_t graph_rounds_ = 0;
  void drop_graphs(RoundCtx &c);
  RoundCtx ctx_[ROUND_CONTEXTS], aux_;  // aux_: set-up sorts and parity evaluations (on context 0's stream)
  void ensure_stage(RoundCtx &c, size_t bytes);
  void ensure_compact(RoundCtx &c, size_t n);
  void free_ctx(RoundCtx &c);
};

}  // namespace texo
