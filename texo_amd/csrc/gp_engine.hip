// gp_engine.hip — lockstep round engine.  Every job is a state machine that
// restates optimize_density_once (eplace.rs:439-582) around
// dynamic_nesterov_step (global_placement.rs:121-335).  A round gathers, for
// every job, the device work its next decision needs (preconditioner,
// extrapolation, candidate step, one or two objective evaluations), issues it
// as one launch per kernel family over all jobs, copies every job's FP64
// scalars back with one transfer, synchronises once, and advances each job.
//
// Exactness of the reuse rules:
//  * the density field depends only on coordinates and charges, so the origin
//    re-evaluation of an iteration (new gamma and lambda) reuses the accepted
//    candidate's field, and the initial evaluation reuses the probe's;
//  * when the extrapolation factor is exactly 0, y == origin bit for bit and
//    the reference reuses the origin evaluation (global_placement.rs:176-181).

#include <chrono>
#include "gp_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unordered_map>
#include <array>
#include <tuple>

namespace texo {

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
double kkt_of(const EvalScalars &s) { return bits_to_double(s.kkt_bits); }

// checked_evaluation (global_placement.rs:562-586)
void check_evaluation(const EvalScalars &s, double value) {
  if (!std::isfinite(value)) fail(TEXO_GP_ERR_INVALID_OBJECTIVE);
  if (s.nonfinite) fail(TEXO_GP_ERR_NON_FINITE_GRADIENT);
}

constexpr double INITIAL_DENSITY_WEIGHT = 1.0e-4;  // eplace.rs:126
constexpr double AREA_RESET_DENSITY_WEIGHT = 0.1;  // eplace.rs:127
constexpr double MULTIPLIER_ALPHA_HIGH = 1.06;     // eplace.rs:129
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
// Iterations between re-sorts of a job's charges for the deposit; the order
// only affects speed, never the density sums.  (Every 32 iterations or at
// 1/32 drift measured 0.7-1% faster than every 8 or at 1/64 on the 84k-cell
// batch and the serving stream: fewer sorts outweigh the added atomics; 1/8
// drift then measured another 0.4-0.8% on the batch, 1/16 and 1/64 less.)
constexpr uint64_t SORT_PERIOD = 32;
// ... or at the next iteration once more than 1/MOVED_RESORT of the charges
// fell outside their tile's window (those take global atomics)
constexpr size_t MOVED_RESORT = 8;

// Rounded targets (rounded_coordinate, lib.rs:4566-4568) of up to
// TARGET_BATCH jobs in one launch; the table travels as the kernel argument.
struct TargetBatch {
  const double *coords[TARGET_BATCH];
  uint32_t *out[TARGET_BATCH];
  int n[TARGET_BATCH];  // 2 x units
  int W[TARGET_BATCH], H[TARGET_BATCH];
  int block_end[TARGET_BATCH];  // prefix sums of the jobs' blocks
  int count;
};

__global__ void k_round_targets(TargetBatch b) {
  int k = 0;
  while (k + 1 < b.count && (int)blockIdx.x >= b.block_end[k]) ++k;
  const int i = ((int)blockIdx.x - (k ? b.block_end[k - 1] : 0)) * BLOCK + threadIdx.x;
  if (i >= b.n[k]) return;
  const double extent = static_cast<double>(((i & 1) ? b.H[k] : b.W[k]) - 1);
  double r = round(b.coords[k][i]);
  if (!(r >= 0.0)) r = 0.0;
  if (r > extent) r = extent;
  b.out[k][i] = static_cast<uint32_t>(r);
}

}  // namespace

namespace {

template <class T>
size_t append(std::vector<char> &buf, const std::vector<T> &v) {
  size_t off = (buf.size() + 255) & ~size_t(255);
  buf.resize(off + v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

// Blocks per task of one launch family, as prefix sums (see BlockPlan).
struct Plan {
  std::vector<int> start{0};
  void add(size_t n, size_t per_block) { start.push_back(start.back() + (int)((n + per_block - 1) / per_block)); }
  void none() { start.push_back(start.back()); }
  int total() const { return start.back(); }
  BlockPlan device(const char *stage, size_t offset) const {
    BlockPlan b{};
    b.start = reinterpret_cast<const int *>(stage + offset);
    b.n_tasks = (int)start.size() - 1;
    b.total = total();
    b.per = b.total >= 2 * XCDS ? (b.total + XCDS - 1) / XCDS : 0;
    return b;
  }
};

// Transform tiles of the density tasks of one launch group (precision and
// shape class), one plan entry per task of the round.
struct DctGroup {
  Plan h, w, h_wide, w_wide;
  int precision = 0, n_tasks = 0, kmax = 0;
  bool fast = false, wide_ok = false;
  void pad(size_t n_tasks) {
    for (Plan *p : {&h, &w, &h_wide, &w_wide})
      while (p->start.size() < n_tasks + 1) p->none();
  }
  void add(size_t task, const Job &j) {
    pad(task);
    n_tasks += 1;
    const size_t F = (size_t)j.F, W = j.W, H = j.H;
    kmax = std::max(kmax, (int)std::max(W, H));
    h.add(((F * H + 31) / 32) * ((W + 31) / 32), 1);
    w.add(((F * W + 31) / 32) * ((H + 31) / 32), 1);
    if (wide_ok) {
      h_wide.add((F * H / 32) * (W / 128), 1);
      w_wide.add((F * W / 32) * (H / 128), 1);
    } else {
      h_wide.none();
      w_wide.none();
    }
  }
  DctLaunch device(const char *stage, const size_t (&offset)[4]) const {
    DctLaunch d{};
    d.h = h.device(stage, offset[0]);
    d.w = w.device(stage, offset[1]);
    d.h_wide = h_wide.device(stage, offset[2]);
    d.w_wide = w_wide.device(stage, offset[3]);
    d.precision = precision;
    d.fast = fast;
    d.wide_ok = wide_ok;
    d.kmax = kmax;
    d.n_tasks = n_tasks;
    return d;
  }
};

DctGroup &dct_group(std::map<std::tuple<int, bool, bool>, DctGroup> &groups, const Job &j) {
  const bool fast = poisson_fast_shape((int)j.W, (int)j.H);
  const bool wide = fast && poisson_wide_ok((int)j.opt.precision, (int)j.W, (int)j.H);
  DctGroup &g = groups[std::make_tuple((int)j.opt.precision, fast, wide)];
  g.precision = (int)j.opt.precision;
  g.fast = fast;
  g.wide_ok = wide;
  return g;
}

}  // namespace

// Gathers the scalar blocks of a round's jobs into one contiguous buffer for
// a single device-to-host copy.
__global__ void k_collect_scalars(const int *__restrict__ index, const JobScalars *__restrict__ pool,
                                  JobScalars *__restrict__ out) {
  constexpr int WORDS = (int)(sizeof(JobScalars) / sizeof(unsigned long long));
  static_assert(sizeof(JobScalars) % sizeof(unsigned long long) == 0, "JobScalars is a whole number of words");
  const unsigned long long *src = reinterpret_cast<const unsigned long long *>(pool + index[blockIdx.x]);
  unsigned long long *dst = reinterpret_cast<unsigned long long *>(out + blockIdx.x);
  for (int w = threadIdx.x; w < WORDS; w += blockDim.x) dst[w] = src[w];
}

void launch_collect_scalars(const int *index, int n, const JobScalars *pool, JobScalars *out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_collect_scalars, dim3(n), dim3(64), 0, s, index, pool, out);
}

void launch_round_targets_to(const double *coords, uint32_t *out, size_t units, uint32_t W, uint32_t H,
                             hipStream_t s) {
  if (units == 0) return;
  TargetBatch b{};
  b.coords[0] = coords;
  b.out[0] = out;
  b.n[0] = (int)(2 * units);
  b.W[0] = (int)W;
  b.H[0] = (int)H;
  b.block_end[0] = (int)((2 * units + BLOCK - 1) / BLOCK);
  b.count = 1;
  hipLaunchKernelGGL(k_round_targets, dim3(b.block_end[0]), dim3(BLOCK), 0, s, b);
}

Failure current_failure() {
  Failure f{};
  try {
    throw;
  } catch (const Failure &e) {
    f = e;
  } catch (const std::bad_alloc &) {
    f.status = TEXO_GP_ERR_OUT_OF_MEMORY;
    g_last_error = "host memory exhausted";
  } catch (const std::exception &e) {
    f.status = TEXO_GP_ERR_DEVICE;
    g_last_error = e.what();
  } catch (...) {
    f.status = TEXO_GP_ERR_DEVICE;
    g_last_error = "unknown exception";
  }
  return f;
}

// Rounds the pending best checkpoints of `jobs` (PlacementCheckpoint::consider
// stores the rounded targets, eplace.rs:107-120) in launches of TARGET_BATCH.
void launch_round_targets(const std::vector<Job *> &jobs, hipStream_t s) {
  TargetBatch b{};
  int blocks = 0;
  auto flush = [&] {
    if (b.count > 0) hipLaunchKernelGGL(k_round_targets, dim3(blocks), dim3(BLOCK), 0, s, b);
    b = TargetBatch{};
    blocks = 0;
  };
  for (Job *j : jobs) {
    if (!j->best_src) continue;
    if (j->U > 0) {
      const int k = b.count++;
      b.coords[k] = j->best_src;
      b.out[k] = j->d_best;
      b.n[k] = (int)(2 * j->U);
      b.W[k] = (int)j->W;
      b.H[k] = (int)j->H;
      blocks += (int)((2 * j->U + BLOCK - 1) / BLOCK);
      b.block_end[k] = blocks;
      if (b.count == TARGET_BATCH) flush();
    }
    j->best_src = nullptr;
  }
  flush();
}

// ---------------------------------------------------------------------------
// Job run set-up
// ---------------------------------------------------------------------------
void Job::count_evaluation() {
  evaluations += 1;
  if (opt.inject && evaluations == opt.inject) fail(TEXO_GP_ERR_INJECTED);
}

void Job::begin_run(bool iterate_mode) {
  started = true;
  fixed_mode = iterate_mode || opt.fixed_iterations > 0;
  no_stop = fixed_mode;
  fixed_limit = (!iterate_mode && opt.fixed_iterations > 0) ? opt.fixed_iterations : UINT64_MAX;
  has_prev_change = false;
  prev_change = 0.0;
  has_fixed_scales = false;
  reset_after_area = false;
  area_round = 0;
  round_limit = empty ? 0 : area_round_limit();
  begin_round();
}

void Job::begin_round() {
  // place() (eplace.rs:357-368): AreaAdjustmentReady until the installed
  // area change falls below one percent, then Final.
  const bool adjust_area = has_routing && !fixed_mode;
  area_ready = adjust_area && !(has_prev_change && prev_change < 0.01);
  exhausted = false;
  phase = empty ? Phase::Done : Phase::Probe;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
Engine::Engine(int device, size_t pool_capacity) : device_(device) {
  HIP(hipSetDevice(device));
  // Replaying recurring rounds from captured graphs is opt-in: measured on
  // MI355X with ROCm 7.2 it gave single jobs no gain and slowed the
  // two-context lockstep batch by 8% (hipGraphLaunch holds the host about as
  // long as the launches it replaces).
  const char *graphs = std::getenv("TEXO_GP_GRAPHS");
  use_graphs_ = graphs && graphs[0] == '1';
  const char *origin = std::getenv("TEXO_GP_ORIGIN_VALUE_ONLY");
  if (origin && (origin[0] == '0' || origin[0] == '2')) origin_mode_ = origin[0] - '0';
  const char *skip = std::getenv("TEXO_GP_SKIP_REJECTED");
  skip_rejected_ = !(skip && skip[0] == '0');
  // every stream keeps the default priority
  for (RoundCtx &c : ctx_) {
    HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP(hipStreamCreateWithFlags(&c.wa_stream, hipStreamNonBlocking));
    HIP(hipEventCreateWithFlags(&c.fork, hipEventDisableTiming));
    HIP(hipEventCreateWithFlags(&c.join, hipEventDisableTiming));
    HIP(hipEventCreateWithFlags(&c.mark, hipEventDisableTiming));
  }
  HIP(hipStreamCreateWithFlags(&io_stream_, hipStreamNonBlocking));
  HIP(hipEventCreateWithFlags(&ev_upload_, hipEventDisableTiming));
  blocks_ = std::make_shared<BlockPool>(BLOCK_POOL_KEEP, false);
  images_ = std::make_shared<BlockPool>(IMAGE_POOL_KEEP, true);
  HIP(hipMalloc(&d_pool_, pool_capacity * sizeof(JobScalars)));
  HIP(hipHostMalloc(reinterpret_cast<void **>(&h_pool_), pool_capacity * sizeof(JobScalars), hipHostMallocDefault));
  std::memset(h_pool_, 0, pool_capacity * sizeof(JobScalars));
  HIP(hipMemset(d_pool_, 0, pool_capacity * sizeof(JobScalars)));
  pool_used_.assign(pool_capacity, false);
  for (RoundCtx &c : ctx_) HIP(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  for (RoundCtx &c : ctx_) {
    (void)hipStreamSynchronize(c.stream);
    (void)hipStreamSynchronize(c.wa_stream);
  }
  (void)hipStreamSynchronize(io_stream_);
  bases_.clear();
  if (d_pool_) (void)hipFree(d_pool_);
  if (h_pool_) (void)hipHostFree(h_pool_);
  for (RoundCtx &c : ctx_) {
    free_ctx(c);
    if (c.done) (void)hipEventDestroy(c.done);
  }
  free_ctx(aux_);
  for (RoundCtx &c : ctx_) {
    (void)hipEventDestroy(c.fork);
    (void)hipEventDestroy(c.join);
    (void)hipEventDestroy(c.mark);
    (void)hipStreamDestroy(c.wa_stream);
    (void)hipStreamDestroy(c.stream);
  }
  (void)hipEventDestroy(ev_upload_);
  (void)hipStreamDestroy(io_stream_);
}

hipStream_t Engine::stream() {
  HIP(hipSetDevice(device_));
  await_uploads(0);
  for (int c = 1; c < ROUND_CONTEXTS; ++c) {
    HIP(hipEventRecord(ctx_[c].mark, ctx_[c].stream));
    HIP(hipStreamWaitEvent(ctx_[0].stream, ctx_[c].mark, 0));
  }
  return ctx_[0].stream;
}

hipStream_t Engine::stream(int c) {
  await_uploads(c);
  return ctx_[c].stream;
}

const Basis *Engine::basis(int W, int H) {
  std::lock_guard<std::mutex> lk(bases_mu_);
  auto key = std::make_pair(W, H);
  auto it = bases_.find(key);
  if (it != bases_.end()) return it->second.get();
  HIP(hipSetDevice(device_));
  std::unique_ptr<Basis> b(new Basis());
  b->build(W, H);
  const Basis *r = b.get();
  bases_[key] = std::move(b);
  return r;
}

void Engine::attach(Job *job) {
  size_t idx = 0;
  while (idx < pool_used_.size() && pool_used_[idx]) ++idx;
  if (idx == pool_used_.size()) fail(TEXO_GP_ERR_OUT_OF_MEMORY);
  pool_used_[idx] = true;
  job->pool_index = (int)idx;
  job->d_scalars = d_pool_ + idx;
  job->h_scalars = h_pool_ + idx;
  if (!job->empty) {
    // the first round sorts the charges by tile (the probe evaluation)
    job->upload(basis((int)job->W, (int)job->H), blocks_, io_stream_);
    share_design(job);
    HIP(hipEventRecord(ev_upload_, io_stream_));
    for (RoundCtx &c : ctx_) c.upload_wait = true;
  }
}

// Jobs of one design (same design-block bytes, matched by their 128-bit
// fingerprint and size) read one device copy of the design block: the first
// job's arena holds it until a second job arrives, which copies it once into
// a block of its own that both then use.  The copy runs on the upload stream
// after the first job's upload; every context waits for the upload stream
// before its next round, and the block is never written after its upload.
void Engine::share_design(Job *job) {
  if (job->design_bytes == 0) return;
  designs_.erase(std::remove_if(designs_.begin(), designs_.end(),
                                [](const DesignEntry &e) { return !e.owner && e.shared.expired(); }),
                 designs_.end());
  DesignEntry *e = nullptr;
  for (DesignEntry &d : designs_)
    if (d.bytes == job->design_bytes && d.hash[0] == job->design_hash[0] && d.hash[1] == job->design_hash[1])
      e = &d;
  if (!e) {
    designs_.push_back({{job->design_hash[0], job->design_hash[1]}, job->design_bytes, job, {}});
    return;
  }
  std::shared_ptr<SharedDesign> sd = e->shared.lock();
  if (!sd) {
    sd = std::make_shared<SharedDesign>();
    sd->pool = blocks_;
    sd->base = blocks_->acquire(job->design_bytes, &sd->capacity);
    HIP(hipMemcpyAsync(sd->base, e->owner->arena.base(), job->design_bytes, hipMemcpyDeviceToDevice, io_stream_));
    e->owner->use_design(sd);
    e->owner = nullptr;
    e->shared = sd;
  }
  job->use_design(sd);
}

void Engine::await_uploads(int c) {
  if (!ctx_[c].upload_wait) return;
  HIP(hipStreamWaitEvent(ctx_[c].stream, ev_upload_, 0));
  ctx_[c].upload_wait = false;
}

void Engine::detach(Job *job) {
  if (job->pool_index >= 0) pool_used_[job->pool_index] = false;
  job->pool_index = -1;
  for (DesignEntry &d : designs_)
    if (d.owner == job) d.owner = nullptr;
}

void Engine::ensure_stage(RoundCtx &c, size_t bytes) {
  if (bytes <= c.stage_bytes) return;
  size_t cap = std::max<size_t>(bytes, 2 * c.stage_bytes);
  cap = (cap + 4095) & ~size_t(4095);
  drop_graphs(c);  // their kernel arguments point into the old buffers
  if (c.d_stage) HIP(hipFree(c.d_stage));
  if (c.h_stage) HIP(hipHostFree(c.h_stage));
  c.d_stage = nullptr;
  c.h_stage = nullptr;
  HIP(hipMalloc(reinterpret_cast<void **>(&c.d_stage), cap));
  HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_stage), cap, hipHostMallocDefault));
  c.stage_bytes = cap;
}

void Engine::ensure_compact(RoundCtx &c, size_t n) {
  if (n <= c.compact_cap) return;
  const size_t cap = std::max<size_t>(n, 2 * c.compact_cap);
  drop_graphs(c);
  if (c.h_compact) HIP(hipHostFree(c.h_compact));
  c.d_compact = nullptr;
  c.h_compact = nullptr;
  // the collecting kernel writes straight into coherent pinned host memory
  HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_compact), cap * sizeof(JobScalars),
                    hipHostMallocMapped | hipHostMallocCoherent));
  HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&c.d_compact), c.h_compact, 0));
  c.compact_cap = cap;
}

void Engine::drop_graphs(RoundCtx &c) {
  for (auto &kv : c.graphs) (void)hipGraphExecDestroy(kv.second);
  c.graphs.clear();
  c.seen.clear();
}

void Engine::free_ctx(RoundCtx &c) {
  drop_graphs(c);
  if (c.d_stage) (void)hipFree(c.d_stage);
  if (c.h_stage) (void)hipHostFree(c.h_stage);
  if (c.h_compact) (void)hipHostFree(c.h_compact);
  c.d_stage = c.h_stage = nullptr;
  c.d_compact = c.h_compact = nullptr;
  c.stage_bytes = c.compact_cap = 0;
}

void Engine::sort_now(Job *j) {
  if (j->F == 0 || j->M + j->Fl == 0) return;
  HIP(hipSetDevice(device_));
  const hipStream_t st = stream();
  j->sorted = true;
  std::vector<SortTask> tasks{j->sort_task()};
  Plan pc, pk;
  pc.add(j->M + j->Fl, BLOCK);
  pk.add((size_t)j->n_keys + 1, BLOCK);
  std::vector<int> list{0};
  std::vector<char> buf;
  const size_t o_t = append(buf, tasks), o_c = append(buf, pc.start), o_k = append(buf, pk.start),
               o_l = append(buf, list);
  // wait for any round in flight before reusing the set-up staging buffer
  HIP(hipStreamSynchronize(st));
  ensure_stage(aux_, buf.size());
  std::memcpy(aux_.h_stage, buf.data(), buf.size());
  HIP(hipMemcpyAsync(aux_.d_stage, aux_.h_stage, buf.size(), hipMemcpyHostToDevice, st));
  SortMaps sm{};
  sm.charges = pc.device(aux_.d_stage, o_c);
  sm.keys = pk.device(aux_.d_stage, o_k);
  sm.list = reinterpret_cast<const int *>(aux_.d_stage + o_l);
  sm.n_list = 1;
  launch_sort(reinterpret_cast<const SortTask *>(aux_.d_stage + o_t), sm, st);
  HIP(hipGetLastError());
  HIP(hipStreamSynchronize(st));
}

// Schedule decisions that need no device work run here, before a job's
// requests for the next round are recorded.
void Engine::plan(Job *j) {
  j->evals.clear();
  j->want_precond = j->want_extrapolate = j->want_candidate = j->want_sort = j->want_spec = false;
  for (;;) {
    switch (j->phase) {
      case Phase::Probe: {
        // density probe at gamma = base (eplace.rs:455)
        j->want_sort = true;
        j->evals.push_back({0, j->d_vec[j->ix], j->d_grad[0], false, j->opt.gamma_base, false, false, 0, false, 0});
        return;
      }
      case Phase::Init: {
        // initial evaluation at the adaptive gamma (eplace.rs:457-458): same
        // coordinates as the probe, whose density field slot 0 still holds
        j->evals.push_back({0, j->d_vec[j->ix], j->d_grad[0], false, j->g0, true, false, 0, false, 0});
        return;
      }
      case Phase::IterStart: {
        // loop head of optimize_density_once (eplace.rs:503-517)
        if (j->no_stop) {
          if (j->local_iterations >= j->fixed_limit) {
            j->phase = Phase::RoundEnd;
            continue;
          }
        } else {
          if (j->stop_reached(j->current, j->area_ready)) {
            j->phase = Phase::RoundEnd;
            continue;
          }
          if (j->iterations >= j->opt.max_iterations) {
            // stop_without_convergence (eplace.rs:989-999)
            if (!j->best_effort) {
              Failure f;
              f.status = TEXO_GP_ERR_DID_NOT_CONVERGE;
              f.index = (int64_t)j->iterations;
              f.v0 = j->F > 0 ? j->overflow(j->current, 0) : 0.0;
              f.v1 = j->F > 1 ? j->overflow(j->current, 1) : 0.0;
              throw f;
            }
            j->exhausted = true;
            j->phase = Phase::RoundEnd;
            continue;
          }
        }
        if (j->iterations >= j->pause_at) {
          j->phase = Phase::Paused;
          return;
        }
        j->gamma = j->adaptive_gamma(j->current);
        j->iterations += 1;
        j->local_iterations += 1;
        // periodic re-sort, or sooner when many charges left their windows
        j->want_sort = j->iterations % SORT_PERIOD == 0 || (size_t)j->moved * MOVED_RESORT > j->M + j->Fl;
        if (!std::isfinite(j->momentum) || j->momentum < 1.0) fail(TEXO_GP_ERR_INVALID_MOMENTUM);
        if (j->has_lipschitz && (!std::isfinite(j->lipschitz) || j->lipschitz <= 0.0))
          fail(TEXO_GP_ERR_INVALID_LIPSCHITZ);
        for (int f = 0; f < j->F; ++f)
          if (!std::isfinite(j->mult[f]) || j->mult[f] < 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
        j->restart = false;
        j->has_wl = j->has_lipschitz;
        j->wl = j->lipschitz;
        const double m = j->momentum;
        j->next_momentum = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * m * m));
        if (!std::isfinite(j->next_momentum)) fail(TEXO_GP_ERR_ITERATION_OVERFLOW);
        j->beta = (m - 1.0) / j->next_momentum;
        j->want_precond = true;
        const int p = j->p_slot, q = (p + 1) % WORKING_SLOTS;
        // with momentum, y is evaluated beside the origin, whose gradient is
        // then read only when y turns out to be the origin or on a restart
        // (global_placement.rs:176-181, 291-301): the origin's value alone,
        // its field being the slot's
        j->origin_value_only = origin_mode_ != 0 && j->p_valid && j->beta != 0.0 && j->wa_bound_ok;
        j->evals.push_back({p, j->d_vec[j->ix], j->d_grad[0], true, j->gamma, j->p_valid, false, scalar_of_slot(p),
                            j->origin_value_only, 0});
        j->y_evaluated = j->beta != 0.0;
        if (j->y_evaluated) {
          j->want_extrapolate = true;
          j->evals.push_back({q, j->d_vec[j->iy], j->d_grad[1], true, j->gamma, false, false, scalar_of_slot(q), false, 0});
        }
        return;
      }
      case Phase::OriginFull: {
        // the origin again, with its gradient: same point, field, gamma and
        // multipliers, so the same value bit for bit
        if (!j->p_valid) fail(TEXO_GP_ERR_DEVICE);
        const int p = j->p_slot;
        j->evals.push_back({p, j->d_vec[j->ix], j->d_grad[0], true, j->gamma, true, false, scalar_of_slot(p), false, 0});
        return;
      }
      case Phase::Trial: {
        // global_placement.rs:213-240
        const double mmax = bits_to_double(j->h_scalars->metric.max_bits);
        const double mmin = bits_to_double(j->h_scalars->metric.min_bits);
        for (;;) {
          j->trials += 1;
          if (!std::isfinite(j->L * mmax) || !std::isfinite(j->L * mmin)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
          if (j->L * mmin == 0.0) {
            j->L *= 2.0;
            j->may_reduce = false;
            if (!std::isfinite(j->L)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
            continue;
          }
          break;
        }
        j->want_candidate = true;
        // many charges outside their windows at the last evaluation: sort
        // again, at this candidate
        j->want_sort = (size_t)j->moved * MOVED_RESORT > j->M + j->Fl;
        const int q = (j->p_slot + 1) % WORKING_SLOTS, r = (j->p_slot + 2) % WORKING_SLOTS;
        j->evals.push_back({q, j->d_vec[j->ic], j->d_grad[2], true, j->gamma, false, false, scalar_of_slot(q), false, 1});
        // the candidate of 2L beside it: what the next trial evaluates when
        // this one only grows L (global_placement.rs:302-309), the common
        // outcome of an iteration's first trial; same coordinates, same
        // results, one round fewer
        j->want_spec = speculate_ && std::isfinite(2.0 * j->L) && std::isfinite(2.0 * j->L * mmax);
        if (j->want_spec) {
          j->spec_L = 2.0 * j->L;
          j->evals.push_back({r, j->d_vec[j->ic2], j->d_grad[3], true, j->gamma, false, false, scalar_of_slot(r), false, 2});
        }
        return;
      }
      case Phase::RoundEnd: {
        // place() after optimize_density_once (eplace.rs:384-419)
        if (!j->area_ready || j->exhausted) {
          j->phase = Phase::Done;
          return;
        }
        area_round(j);
        continue;
      }
      default:
        return;
    }
  }
}

// propose_area_adjustment (eplace.rs:627-729): RUDY and register clustering
// on the device, the order-dependent charge bookkeeping on the host.
Engine::AreaProposal Engine::propose_areas(Job *j, const double *coords, hipStream_t s) {
  HIP(hipMemsetAsync(j->d_rudy_err, 0, sizeof(unsigned int), s));
  launch_rudy((int)j->n_nets, j->d_net_start, j->d_pin_unit, j->js.pin_xy, coords, (int)j->W, (int)j->H, j->d_rudy_fx,
              j->d_rudy, j->d_rudy_err, s);
  const size_t G = j->G, U = j->U;
  std::vector<double> demand(2 * G), x(2 * U);
  unsigned int err = 0;
  HIP(hipMemcpyAsync(demand.data(), j->d_rudy, 2 * G * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP(hipMemcpyAsync(x.data(), coords, 2 * U * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP(hipMemcpyAsync(&err, j->d_rudy_err, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
  HIP(hipStreamSynchronize(s));
  if (err) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  std::map<uint32_t, double> clustering;
  if (!j->controls.empty()) {
    // propose_register_clustering_areas (eplace.rs:760-791)
    std::vector<std::pair<uint32_t, std::pair<double, double>>> regs;
    for (size_t i = 0; i < j->DU; ++i) {
      const uint32_t u = j->du_unit[i];
      for (uint32_t m = j->du_mstart[i]; m < j->du_mstart[i + 1]; ++m)
        if (j->field_kind[j->m_field[m]] == TEXO_KIND_REGISTER)
          regs.push_back({j->m_cell[m], {x[2 * u] + j->m_off[2 * m], x[2 * u + 1] + j->m_off[2 * m + 1]}});
    }
    std::vector<double> areas;
    int ok = register_clustering_device(regs, j->controls, j->movable_instance_count, areas, blocks_.get(), s);
    if (ok < 0) ok = register_clustering(regs, j->controls, j->movable_instance_count, areas) ? 1 : 0;
    if (!ok) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    for (size_t i = 0; i < regs.size(); ++i) clustering[regs[i].first] = areas[i];
  }
  // ContinuousRoutingDemand::utilization (eplace.rs:822-854)
  auto utilization = [&](const double *d, const std::vector<uint32_t> &cap, uint32_t px, uint32_t py) {
    const size_t idx = (size_t)py * j->W + px;
    if (cap[idx] == 0) return d[idx] == 0.0 ? 0.0 : kSqrt2;
    const double u = d[idx] / (double)cap[idx];
    if (!(std::isfinite(u) && u >= 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
    return u;
  };
  std::vector<AreaMember> members(j->M);
  for (size_t i = 0; i < j->DU; ++i) {
    const uint32_t u = j->du_unit[i];
    for (uint32_t m = j->du_mstart[i]; m < j->du_mstart[i + 1]; ++m) {
      const uint32_t px = rounded_coordinate(x[2 * u] + j->m_off[2 * m], j->W);
      const uint32_t py = rounded_coordinate(x[2 * u + 1] + j->m_off[2 * m + 1], j->H);
      const double h = std::min(utilization(demand.data(), j->routing_h, px, py), kSqrt2);
      const double v = std::min(utilization(demand.data() + G, j->routing_v, px, py), kSqrt2);
      const uint32_t kind = j->field_kind[j->m_field[m]];
      // routability_adjusted_member_area (eplace.rs:741-758, instance_area.rs:64-84)
      double area = j->m_charge[m];
      if (is_lut(kind) || kind == TEXO_KIND_REGISTER) {
        const double uu = std::max(h, v);
        area = j->m_charge[m] * std::min(uu * uu, 2.0);
      }
      auto it = clustering.find(j->m_cell[m]);
      members[m] = {kind, j->m_charge[m], area, it == clustering.end() ? 0.0 : it->second};
    }
  }
  std::vector<std::pair<uint32_t, double>> fillers(j->Fl);
  for (size_t k = 0; k < j->Fl; ++k) fillers[k] = {j->field_kind[j->fl_field[k]], j->fl_charge[k]};
  AreaProposal p;
  p.requires_reset = adjust_areas(members, fillers, p.member_areas, p.filler_areas);
  std::map<uint32_t, double> before, inc;
  for (size_t m = 0; m < j->M; ++m) {
    before[members[m].kind] += members[m].current;
    inc[members[m].kind] += p.member_areas[m] - members[m].current;
  }
  double bt = 0.0, it = 0.0;
  for (auto &kv : before) bt += kv.second;
  for (auto &kv : inc) it += kv.second;
  p.total_relative_increase = it / bt;
  if (!std::isfinite(p.total_relative_increase)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  return p;
}

// The area round of place() (eplace.rs:384-419): propose, then either
// restart with the same areas (no growth) or install them
// (apply_area_adjustment, eplace.rs:793-819).
void Engine::area_round(Job *j) {
  await_uploads(cur_);
  const AreaProposal p = propose_areas(j, j->d_vec[j->ix], cur());
  const std::vector<double> &ma = p.member_areas, &fa = p.filler_areas;
  const double total_inc = p.total_relative_increase;
  if (!p.requires_reset) {
    j->has_prev_change = true;
    j->prev_change = 0.0;
    j->reset_after_area = true;
    j->begin_round();
    return;
  }
  if (j->area_round >= j->round_limit) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
  j->apply_charges(ma, fa, cur());
  j->p_valid = false;
  j->has_prev_change = true;
  j->prev_change = total_inc;
  j->reset_after_area = true;
  j->area_round += 1;
  j->area_rounds += 1;
  j->begin_round();
}

namespace {

// Restart from the origin (global_placement.rs:291-301): momentum 1, y = origin.
void restart_from_origin(Job *j) {
  j->restart = true;
  j->next_momentum = 0.5 * (1.0 + std::sqrt(1.0 + 4.0));
  j->y_is_origin = true;
  j->y_s = j->origin_s;
  j->y_value = j->origin_value;
}

void consider_best(Job *j, hipStream_t s) {
  // PlacementCheckpoint::consider (eplace.rs:107-120): lexicographic total_cmp
  const auto sc = j->excess(j->current);
  const bool better = !j->has_best || sc.first < j->best_max || (sc.first == j->best_max && sc.second < j->best_total);
  if (better) {
    // rounded into d_best by the engine's batched launch once the round's
    // jobs have advanced (before any later work on the stream)
    j->best_src = j->d_vec[j->ix];
    j->best_max = sc.first;
    j->best_total = sc.second;
    j->has_best = true;
  }
}

void iteration_end(Job *j, bool accepted, hipStream_t s) {
  // eplace.rs:530-573
  j->last_wa = j->current.wa_value;
  j->validate_fields(j->current);
  consider_best(j, s);
  if (!j->no_stop) {
    if (j->stop_reached(j->current, j->area_ready)) {
      j->phase = Phase::RoundEnd;
      return;
    }
    // stationary_target_is_fixed (eplace.rs:973-985)
    if (accepted) {
      j->has_prev_stationary = false;
    } else {
      std::vector<double> x(2 * j->U);
      HIP(hipMemcpyAsync(x.data(), j->d_vec[j->ix], 2 * j->U * sizeof(double), hipMemcpyDeviceToHost, s));
      HIP(hipStreamSynchronize(s));
      std::vector<uint32_t> t(2 * j->U);
      for (size_t i = 0; i < 2 * j->U; ++i) t[i] = rounded_coordinate(x[i], (i & 1) ? j->H : j->W);
      const bool fixed = j->has_prev_stationary && j->prev_stationary == t;
      j->prev_stationary = t;
      j->has_prev_stationary = true;
      if (fixed) {
        if (!j->best_effort) {
          Failure f;
          f.status = TEXO_GP_ERR_DID_NOT_CONVERGE;
          f.index = (int64_t)j->iterations;
          throw f;
        }
        j->exhausted = true;
        j->phase = Phase::RoundEnd;
        return;
      }
    }
  }
  j->update_multipliers();
  j->phase = Phase::IterStart;
}

// Copies the current coordinates over the previous iterate, first rounding
// a pending best checkpoint that still lives in the overwritten buffer.
void copy_to_previous(Job *j, hipStream_t s) {
  if (j->best_src && j->best_src == j->d_vec[j->ixp]) launch_round_targets(std::vector<Job *>{j}, s);
  HIP(hipMemcpyAsync(j->d_vec[j->ixp], j->d_vec[j->ix], j->C * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void stationary(Job *j, bool set_l, double l, hipStream_t s) {
  // NumericallyStationary return (global_placement.rs:185-208, 259-271)
  copy_to_previous(j, s);
  j->momentum = 1.0;
  if (set_l) {
    j->has_lipschitz = true;
    j->lipschitz = l;
  }
  j->current = j->origin_s;
  iteration_end(j, false, s);
}

// y-stationarity decision after y is known (global_placement.rs:182-211).
void decide_y(Job *j, hipStream_t s) {
  for (;;) {
    if (j->y_is_origin && !j->origin_full) {
      // y is the origin, whose gradient is not evaluated yet
      j->held_y = false;
      j->phase = Phase::OriginFull;
      return;
    }
    const double y_stat = kkt_of(j->y_s);
    if (!std::isfinite(y_stat)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
    if (y_stat == 0.0) {
      if (!j->y_is_origin) {
        restart_from_origin(j);
        continue;
      }
      stationary(j, false, 0.0, s);
      return;
    }
    j->L = half_nonzero(j->has_wl ? j->wl : y_stat);
    j->may_reduce = true;
    j->phase = Phase::Trial;
    return;
  }
}

// One trial's outcome (global_placement.rs:241-311) from its step scalars
// and its candidate's evaluation in working slot `slot` (coordinates
// d_vec[ic], or d_vec[ic2] for the speculative candidate).  Returns true
// when the trial only doubled L, the next trial being due.
bool trial_outcome(Job *j, const StepScalars &st, const EvalScalars &ce, int slot, bool spec, hipStream_t s) {
  if (!st.changed) {
    // candidate == y (global_placement.rs:241-272)
    if (j->may_reduce) {
      const double smaller = half_nonzero(j->L);
      if (smaller < j->L) {
        j->L = smaller;
        return false;
      }
    }
    if (!j->y_is_origin) {
      j->has_wl = true;
      j->wl = j->L;
      restart_from_origin(j);
      decide_y(j, s);
      return false;
    }
    stationary(j, true, j->L, s);
    return false;
  }
  j->count_evaluation();
  const double value = j->combined_value(ce);
  check_evaluation(ce, value);
  const bool skipped = ce.gradient_skipped != 0u;
  // majorizer and monotonicity (global_placement.rs:274-290)
  const double upper = j->y_value + st.gdelta + 0.5 * j->L * st.msq;
  if (!std::isfinite(upper)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
  const bool smooth = value <= upper + rounding_margin(value, upper);
  const bool monotone = value <= j->origin_value + rounding_margin(value, j->origin_value);
  if (smooth && monotone) {
    // the device's test accepts whatever the host accepts: its gradient ran
    if (skipped) fail(TEXO_GP_ERR_DEVICE);
    const double stat = kkt_of(ce);
    if (!std::isfinite(stat)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
    // previous <- origin, coordinates <- candidate (global_placement.rs:322-325)
    const int old_prev = j->ixp;
    j->ixp = j->ix;
    int &cand = spec ? j->ic2 : j->ic;
    j->ix = cand;
    cand = old_prev;
    j->momentum = j->next_momentum;
    j->has_lipschitz = true;
    j->lipschitz = j->L;
    j->current = ce;
    j->p_slot = slot;
    j->p_valid = true;
    if (j->opt.trace)
      std::fprintf(stderr,
                   "TEXO_PNR_METRICS eplace iteration=%llu objective=%.9e stationarity=%.9e "
                   "coordinate_change=%.9e line_search_trials=%llu gamma=%.9e wa=%.9e\n",
                   (unsigned long long)j->iterations, value, stat, bits_to_double(st.change_bits),
                   (unsigned long long)j->trials, j->gamma, ce.wa_value);
    iteration_end(j, true, s);
    return false;
  }
  if (smooth && !j->restart) {
    j->has_wl = true;
    j->wl = j->L;
    restart_from_origin(j);
    decide_y(j, s);
    return false;
  }
  j->L *= 2.0;
  j->may_reduce = false;
  if (!std::isfinite(j->L)) fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
  return true;
}

// y's evaluation of an iteration (global_placement.rs:162-182), after the
// origin's: `has_y` when y was evaluated and differs from the origin.
void take_y(Job *j, bool has_y, const EvalScalars &ys, hipStream_t s) {
  j->y_is_origin = !has_y;
  if (has_y) {
    j->count_evaluation();
    j->y_s = ys;
    j->y_value = j->combined_value(j->y_s);
    check_evaluation(j->y_s, j->y_value);
  } else {
    j->y_s = j->origin_s;
    j->y_value = j->origin_value;
  }
  decide_y(j, s);
}

}  // namespace

void Engine::advance(Job *j) {
  const JobScalars &hs = *j->h_scalars;
  switch (j->phase) {
    case Phase::Probe: {
      j->count_evaluation();
      const EvalScalars probe = hs.eval[0];
      j->validate_initial(probe);
      j->g0 = j->adaptive_gamma(probe);
      j->phase = Phase::Init;
      return;
    }
    case Phase::Init: {
      // eplace.rs:459-499
      j->count_evaluation();
      const EvalScalars init = hs.eval[0];
      j->validate_initial(init);
      if (j->has_fixed_scales) {
        j->scales = j->fixed_scales;
      } else {
        j->scales.assign(j->F, 1.0);
        for (int f = 0; f < j->F; ++f) j->scales[f] = std::max(init.energy[f], 1.0);
        j->fixed_scales = j->scales;
        j->has_fixed_scales = true;
      }
      j->validate_fields(init);
      // DynamicNesterovState::new: previous = coordinates, momentum 1, no L
      copy_to_previous(j, cur());
      j->momentum = 1.0;
      j->has_lipschitz = false;
      j->lipschitz = 0.0;
      j->has_best = false;
      j->current = init;
      j->last_wa = init.wa_value;
      consider_best(j, cur());
      j->has_prev_stationary = false;
      j->p_slot = 0;
      j->p_valid = true;
      j->local_iterations = 0;
      if (!j->no_stop && j->stop_reached(init, j->area_ready)) {
        j->phase = Phase::RoundEnd;
        return;
      }
      if (j->reset_after_area) {
        // area_adjusted_density_multipliers / _step (eplace.rs:1183-1232)
        const std::vector<double> s = j->subgradient(init);
        const double wf = init.wa_l1;
        double df = 0.0;
        for (int f = 0; f < j->F; ++f) df += init.force_l1[f] * s[f];
        if (!std::isfinite(wf) || wf <= 0.0 || !std::isfinite(df) || df <= 0.0) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
        const double scale = AREA_RESET_DENSITY_WEIGHT * wf / df;
        bool any = false;
        double norm = 0.0;
        j->mult.assign(j->F, 0.0);
        for (int f = 0; f < j->F; ++f) {
          j->mult[f] = scale * s[f];
          if (!(std::isfinite(j->mult[f]) && j->mult[f] >= 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
          any = any || j->mult[f] > 0.0;
          norm += j->mult[f] * j->mult[f];
        }
        if (!any) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
        j->mstep = (MULTIPLIER_ALPHA_HIGH - 1.0) * std::sqrt(norm);
        if (!(std::isfinite(j->mstep) && j->mstep > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
      } else {
        // initial_density_multiplier (eplace.rs:1126-1148)
        const double wf = init.wa_l1;
        double df = 0.0;
        for (int f = 0; f < j->F; ++f) df += init.raw_l1[f];
        const double m0 = (wf == 0.0 || df == 0.0) ? 1.0 : INITIAL_DENSITY_WEIGHT * wf / df;
        if (!(std::isfinite(m0) && m0 > 0.0)) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
        j->mult.assign(j->F, m0);
        j->mstep = MULTIPLIER_ALPHA_HIGH - 1.0;
      }
      j->phase = Phase::IterStart;
      return;
    }
    case Phase::IterStart: {
      // origin (and y) evaluated (global_placement.rs:143-182)
      if (hs.metric.nonfinite) fail(TEXO_GP_ERR_INVALID_NORMALIZATION);
      j->count_evaluation();
      const int p = j->p_slot, q = (p + 1) % WORKING_SLOTS;
      j->origin_s = hs.eval[scalar_of_slot(p)];
      j->origin_value = j->combined_value(j->origin_s);
      j->p_valid = true;
      const bool has_y = j->y_evaluated && hs.step.differs;
      if (j->origin_value_only) {
        // checked_evaluation of the origin (global_placement.rs:143-149):
        // the value here; the gradient is finite when the bound says so,
        // else an OriginFull round checks it before y is taken
        if (!std::isfinite(j->origin_value)) fail(TEXO_GP_ERR_INVALID_OBJECTIVE);
        j->origin_full = false;
        if (origin_mode_ == 2 || !j->gradient_certainly_finite(j->origin_s, j->gamma)) {
          j->held_y = true;
          j->held_y_has = has_y;
          if (has_y) j->held_y_s = hs.eval[scalar_of_slot(q)];
          j->phase = Phase::OriginFull;
          return;
        }
      } else {
        check_evaluation(j->origin_s, j->origin_value);
        j->origin_full = true;
      }
      take_y(j, has_y, hs.eval[scalar_of_slot(q)], cur());
      return;
    }
    case Phase::OriginFull: {
      const EvalScalars full = hs.eval[scalar_of_slot(j->p_slot)];
      const double value = j->combined_value(full);
      // the value-only evaluation's value, bit for bit
      if (std::memcmp(&value, &j->origin_value, sizeof value) != 0) fail(TEXO_GP_ERR_DEVICE);
      check_evaluation(full, value);
      j->origin_s = full;
      j->origin_full = true;
      if (j->held_y) {
        j->held_y = false;
        take_y(j, j->held_y_has, j->held_y_s, cur());
        return;
      }
      // a decision on y at the origin (restart or y == origin)
      j->y_s = j->origin_s;
      j->y_value = j->origin_value;
      decide_y(j, cur());
      return;
    }
    case Phase::Trial: {
      const int q = (j->p_slot + 1) % WORKING_SLOTS, r = (j->p_slot + 2) % WORKING_SLOTS;
      const bool spec = j->want_spec;  // planned and evaluated in the round just finished
      if (!trial_outcome(j, hs.step, hs.eval[scalar_of_slot(q)], q, false, cur())) return;
      // the trial only grew L: the speculative candidate is the next trial's
      if (!spec || j->L != j->spec_L) return;
      const double mmax = bits_to_double(hs.metric.max_bits), mmin = bits_to_double(hs.metric.min_bits);
      // the per-trial steps of plan (global_placement.rs:213-240); L * mmin
      // cannot vanish here (the trial's L / 2 did not)
      if (!std::isfinite(j->L * mmax) || !std::isfinite(j->L * mmin)) {
        j->trials += 1;
        fail(TEXO_GP_ERR_LINE_SEARCH_OVERFLOW);
      }
      if (j->L * mmin == 0.0) return;
      j->trials += 1;
      (void)trial_outcome(j, hs.step_spec, hs.eval[scalar_of_slot(r)], r, true, cur());
      return;
    }
    default:
      return;
  }
}


void Engine::round(const std::vector<Job *> &jobs) {
  if (launch(0, jobs)) finish(0);
}

bool Engine::launch(int c, const std::vector<Job *> &jobs) {
  RoundCtx &cx = ctx_[c];
  HIP(hipSetDevice(device_));
  cur_ = c;
  g_timer = &timer;
  std::vector<Job *> &active = cx.active;
  active.clear();
  size_t live = 0, live_charges = 0;
  for (Job *j : jobs)
    if (j->phase != Phase::Done && j->phase != Phase::Failed && j->phase != Phase::Paused) {
      live += 1;
      live_charges += j->M + j->Fl;
    }
  speculate_ = live <= SPECULATE_MAX_JOBS || live_charges <= SPECULATE_MAX_CHARGES;
  for (Job *j : jobs) {
    if (j->phase == Phase::Done || j->phase == Phase::Failed || j->phase == Phase::Paused) continue;
    try {
      plan(j);
      if (!j->evals.empty()) active.push_back(j);
    } catch (...) {
      j->failure = current_failure();
      j->status = j->failure.status;
      j->phase = Phase::Failed;
    }
  }
  if (active.empty()) {
    g_timer = nullptr;
    return false;
  }

  // ---- tables ----
  std::vector<OptTask> opt;
  std::vector<SortTask> sorts;
  std::vector<int> l_sort, l_pre, l_cand;
  Plan p_sort_c, p_sort_k, p_pre, p_ext, p_cand;
  std::vector<Task> tasks;
  int n_density = 0, n_gradient = 0;
  double wa_evaluations = 0.0;
  Plan p_dep, p_conv, p_wa, p_gr;
  std::map<std::tuple<int, bool, bool>, DctGroup> groups;
  for (Job *j : active) {
    if (j->want_sort && j->F > 0 && j->M + j->Fl > 0) {
      // by tile at the coordinates of the round's first density evaluation
      // (computed by the optimizer launch that precedes the sort)
      const double *at = j->d_vec[j->ix];
      for (const Job::EvalReq &r : j->evals)
        if (!r.skip) {
          at = r.coords;
          break;
        }
      j->sorted = true;
      l_sort.push_back((int)sorts.size());
      sorts.push_back(j->sort_task(at));
      p_sort_c.add(j->M + j->Fl, BLOCK);
      p_sort_k.add((size_t)j->n_keys + 1, BLOCK);
    }
    if (j->want_precond || j->want_extrapolate || j->want_candidate) {
      OptTask o{};
      o.n = j->C;
      o.n_units = (int)j->U;
      o.n_fillers = (int)j->Fl;
      o.F = j->F;
      o.DU = (int)j->DU;
      o.lower = j->d_lower;
      o.upper = j->d_upper;
      o.wl_diag = j->d_wl_diag;
      o.du_fcharge = j->d_du_fcharge;
      o.fl_charge = j->d_fl_charge;
      o.unit_du = j->d_unit_du;
      o.fl_field = j->d_fl_field;
      for (int f = 0; f < MAXF; ++f) o.lambda[f] = f < j->F ? j->mult[f] : 0.0;
      o.metric = j->d_metric;
      o.part_metric = j->d_part_metric;
      o.metric_scalars = &j->d_scalars->metric;
      o.x = j->d_vec[j->ix];
      o.xp = j->d_vec[j->ixp];
      o.y = j->d_vec[j->iy];
      o.beta = j->beta;
      o.yp = j->y_is_origin ? j->d_vec[j->ix] : j->d_vec[j->iy];
      o.gy = j->y_is_origin ? j->d_grad[0] : j->d_grad[1];
      o.origin = j->d_vec[j->ix];
      o.cand = j->d_vec[j->ic];
      o.lipschitz = j->L;
      o.part_step = j->d_part_step;
      o.step = &j->d_scalars->step;
      o.do_precond = j->want_precond;
      o.do_extrapolate = j->want_extrapolate;
      o.do_candidate = j->want_candidate;
      const int oi = (int)opt.size();
      opt.push_back(o);
      if (j->want_precond) {
        p_pre.add(j->U + j->Fl, BLOCK);
        l_pre.push_back(oi);
      } else {
        p_pre.none();
      }
      if (j->want_extrapolate)
        p_ext.add(j->C, BLOCK);
      else
        p_ext.none();
      if (j->want_candidate) {
        p_cand.add(j->C, BLOCK);
        l_cand.push_back(oi);
      } else {
        p_cand.none();
      }
      if (j->want_spec) {
        // the speculative candidate at 2L, its own step partials and scalars
        OptTask sp = o;
        sp.do_precond = sp.do_extrapolate = 0;
        sp.do_candidate = 1;
        sp.cand = j->d_vec[j->ic2];
        sp.lipschitz = j->spec_L;
        sp.part_step = j->d_part_step_spec;
        sp.step = &j->d_scalars->step_spec;
        const int si = (int)opt.size();
        opt.push_back(sp);
        p_pre.none();
        p_ext.none();
        p_cand.add(j->C, BLOCK);
        l_cand.push_back(si);
      }
    }
    // the nets of fixed pins run when their value at this round's gamma is
    // not held yet (every evaluation of a job in a round has one gamma)
    bool one_gamma = true;
    for (const Job::EvalReq &r : j->evals) one_gamma = one_gamma && r.gamma == j->evals.front().gamma;
    const bool fixed_nets = !j->evals.empty() && (!one_gamma || !j->fixed_wa_valid ||
                                                  j->evals.front().gamma != j->fixed_wa_gamma);
    if (fixed_nets) {
      j->fixed_wa_valid = one_gamma;
      j->fixed_wa_gamma = j->evals.front().gamma;
    }
    for (const Job::EvalReq &r : j->evals) {
      Task t = j->task(r.slot, r.coords, r.grad, r.metric ? j->d_metric : nullptr, r.gamma, r.skip,
                       &j->d_scalars->eval[r.scalar_slot], fixed_nets, r.value_only);
      if (r.trial && skip_rejected_) {
        // the acceptance test of the trial planned at L (or the speculative
        // one at 2L), against this iteration's y and origin values
        t.trial = 1;
        t.trial_y_value = j->y_value;
        t.trial_origin_value = j->origin_value;
        t.trial_L = r.trial == 2 ? j->spec_L : j->L;
        t.trial_step = r.trial == 2 ? &j->d_scalars->step_spec : &j->d_scalars->step;
        t.trial_wa_bound = j->wa_bound_ok ? 3.0 * j->wa_bound_w * (1.0 + 3.0 * j->wa_bound_ext / r.gamma) : INFINITY;
        for (int f = 0; f < MAXF; ++f) t.trial_charge[f] = f < j->F ? j->field_charge[f] : 0.0;
        if (poisson_fast_shape((int)j->W, (int)j->H) && j->F > 0) {
          t.phi_amax = j->slot[r.slot].amax_part + (size_t)j->nb_conv * MAXF + 3 * (size_t)j->nb_tile;
          t.phi_tiles = (int)(j->G / 1024);
        }
      }
      if (r.parity) {
        t.raw_units = j->d_raw_units;
        t.raw_fillers = j->d_raw_fillers;
        t.wa_grad_out = j->d_wa_grad;
      }
      const int ti = (int)tasks.size();
      tasks.push_back(t);
      const bool density = j->F > 0 && !r.skip;
      // a skipped task reuses the slot's field (same point)
      if (density) {
        n_density += 1;
        p_dep.add((size_t)j->n_keys, 1);
        p_conv.add((size_t)j->F * j->G, BLOCK);
        dct_group(groups, *j).add((size_t)ti, *j);
      } else {
        p_dep.none();
        p_conv.none();
      }
      p_wa.add((size_t)t.nb_wa, 1);
      wa_evaluations += r.value_only ? j->wa_value_only_share : 1.0;
      if (r.value_only) {
        p_gr.none();
      } else {
        n_gradient += 1;
        p_gr.add((size_t)(t.nb_ug + t.nb_gf), 1);
      }
    }
  }
  std::vector<char> buf;
  const size_t o_opt = append(buf, opt), o_tasks = append(buf, tasks), o_sorts = append(buf, sorts);
  const size_t o_lpre = append(buf, l_pre), o_lcand = append(buf, l_cand), o_lsort = append(buf, l_sort);
  const size_t o_psc = append(buf, p_sort_c.start), o_psk = append(buf, p_sort_k.start),
               o_ppre = append(buf, p_pre.start), o_pext = append(buf, p_ext.start),
               o_pcand = append(buf, p_cand.start);
  const size_t o_pdep = append(buf, p_dep.start), o_pconv = append(buf, p_conv.start),
               o_pwa = append(buf, p_wa.start), o_pgr = append(buf, p_gr.start);
  std::vector<std::array<size_t, 4>> o_groups;
  for (auto &kv : groups) {
    DctGroup &g = kv.second;
    g.pad(tasks.size());
    o_groups.push_back({append(buf, g.h.start), append(buf, g.w.start), append(buf, g.h_wide.start),
                        append(buf, g.w_wide.start)});
  }
  std::vector<int> pool_index;
  for (Job *j : active) pool_index.push_back(j->pool_index);
  const size_t o_pool = append(buf, pool_index);
  ensure_stage(cx, buf.size());
  ensure_compact(cx, active.size());
  std::memcpy(cx.h_stage, buf.data(), buf.size());
  await_uploads(c);
  HIP(hipMemcpyAsync(cx.d_stage, cx.h_stage, buf.size(), hipMemcpyHostToDevice, cx.stream));
  char *const d_stage_ = cx.d_stage;
  auto dptr = [&](size_t off) { return d_stage_ + off; };
  const Task *d_tasks = reinterpret_cast<const Task *>(dptr(o_tasks));

  // ---- launches ----
  OptMaps om{};
  om.precond = p_pre.device(d_stage_, o_ppre);
  om.extrapolate = p_ext.device(d_stage_, o_pext);
  om.candidate = p_cand.device(d_stage_, o_pcand);
  om.n_tasks = (int)opt.size();
  const OptTask *d_opt = reinterpret_cast<const OptTask *>(dptr(o_opt));
  Maps mp{};
  mp.deposit = p_dep.device(d_stage_, o_pdep);
  mp.convert = p_conv.device(d_stage_, o_pconv);
  mp.wa = p_wa.device(d_stage_, o_pwa);
  mp.gradient = p_gr.device(d_stage_, o_pgr);
  mp.n_tasks = (int)tasks.size();
  mp.n_density = n_density;
  mp.n_gradient = n_gradient;
  mp.wa_evaluations = wa_evaluations;
  SortMaps sm{};
  if (!sorts.empty()) {
    sm.charges = p_sort_c.device(d_stage_, o_psc);
    sm.keys = p_sort_k.device(d_stage_, o_psk);
    sm.list = reinterpret_cast<const int *>(dptr(o_lsort));
    sm.n_list = (int)l_sort.size();
  }
  std::vector<DctLaunch> dct;
  {
    size_t gi = 0;
    for (auto &kv : groups) {
      const size_t off[4] = {o_groups[gi][0], o_groups[gi][1], o_groups[gi][2], o_groups[gi][3]};
      dct.push_back(kv.second.device(d_stage_, off));
      ++gi;
    }
  }
  int opt_passes = 0;
  for (const OptTask &o : opt) opt_passes += 2 * o.do_precond + 5 * o.do_extrapolate + 7 * o.do_candidate;
  const int *d_pool_index = reinterpret_cast<const int *>(dptr(o_pool));
  const SortTask *d_sorts = reinterpret_cast<const SortTask *>(dptr(o_sorts));
  auto issue = [&] {
    if (!opt.empty()) launch_optimizer(d_opt, om, cx.stream, opt_passes);
    // The second stream forks after the optimizer's vectors: it runs WA (it
    // depends only on coordinates) and then reduces the optimizer's scalars
    // (read by the trial tests and the host) beside the sort and the density
    // chain, and joins before the coefficients.
    HIP(hipEventRecord(cx.fork, cx.stream));
    HIP(hipStreamWaitEvent(cx.wa_stream, cx.fork, 0));
    launch_wa(d_tasks, mp, cx.wa_stream);
    if (!opt.empty()) launch_optimizer_finalize(d_opt, om, cx.wa_stream);
    HIP(hipEventRecord(cx.join, cx.wa_stream));
    if (!sorts.empty()) launch_sort(d_sorts, sm, cx.stream);
    launch_density(d_tasks, mp, cx.stream);
    for (const DctLaunch &d : dct) launch_poisson(d_tasks, d, cx.stream);
    // the coefficients' trial tests read the WA value
    HIP(hipStreamWaitEvent(cx.stream, cx.join, 0));
    launch_gather(d_tasks, mp, cx.stream);
    launch_combine(d_tasks, mp, cx.stream);
    launch_collect_scalars(d_pool_index, (int)active.size(), d_pool_, cx.d_compact, cx.stream);
  };
  // A round's launch sequence is fully described by the by-value kernel
  // arguments below (every table it reads sits in the staging buffer, copied
  // above).  Sequences that recur (a single job's iteration rounds, a stable
  // batch) replay as one captured graph: the same kernels with the same
  // arguments, so results do not change, with one launch call instead of a
  // dozen.  Kernel timing needs its events between launches: no graphs then.
  if (use_graphs_ && !timer.enabled) {
    std::string key;
    key.reserve(512);
    auto add = [&](const void *p, size_t n) { key.append(static_cast<const char *>(p), n); };
    auto add_plan = [&](const BlockPlan &b) {
      add(&b.start, sizeof b.start);
      add(&b.n_tasks, sizeof b.n_tasks);
      add(&b.total, sizeof b.total);
      add(&b.per, sizeof b.per);
    };
    const size_t n_opt = opt.size(), n_sorts = sorts.size(), n_active = active.size();
    add(&n_opt, sizeof n_opt);
    add(&n_sorts, sizeof n_sorts);
    add(&n_active, sizeof n_active);
    add(&d_stage_, sizeof d_stage_);
    add(&cx.d_compact, sizeof cx.d_compact);
    for (const BlockPlan *b : {&om.precond, &om.extrapolate, &om.candidate, &mp.deposit, &mp.convert, &mp.wa,
                               &mp.gradient, &sm.charges, &sm.keys})
      add_plan(*b);
    add(&mp.n_tasks, sizeof mp.n_tasks);
    add(&mp.n_density, sizeof mp.n_density);
    add(&sm.list, sizeof sm.list);
    add(&sm.n_list, sizeof sm.n_list);
    for (const DctLaunch &d : dct) {
      for (const BlockPlan *b : {&d.h, &d.w, &d.h_wide, &d.w_wide}) add_plan(*b);
      add(&d.precision, sizeof d.precision);
      add(&d.kmax, sizeof d.kmax);
      add(&d.n_tasks, sizeof d.n_tasks);
      const char flags[2] = {d.fast ? '1' : '0', d.wide_ok ? '1' : '0'};
      add(flags, 2);
    }
    auto hit = cx.graphs.find(key);
    if (hit == cx.graphs.end() && ++cx.seen[key] >= GRAPH_AFTER_ROUNDS) {
      if (cx.graphs.size() >= GRAPH_CACHE) {
        for (auto &kv : cx.graphs) (void)hipGraphExecDestroy(kv.second);
        cx.graphs.clear();
      }
      if (cx.seen.size() > 64 * GRAPH_CACHE) cx.seen.clear();
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIP(hipStreamBeginCapture(cx.stream, hipStreamCaptureModeThreadLocal));
      try {
        issue();
      } catch (...) {
        (void)hipStreamEndCapture(cx.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      HIP(hipStreamEndCapture(cx.stream, &graph));
      const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP(ie);
      hit = cx.graphs.emplace(std::move(key), exec).first;
    }
    if (hit != cx.graphs.end()) {
      HIP(hipGraphLaunch(hit->second, cx.stream));
      graph_rounds_ += 1;
    } else {
      issue();
    }
  } else {
    issue();
  }
  HIP(hipGetLastError());
  HIP(hipEventRecord(cx.done, cx.stream));
  cx.pending = true;
  g_timer = nullptr;
  return true;
}

void Engine::finish(int c) {
  RoundCtx &cx = ctx_[c];
  if (!cx.pending) return;
  HIP(hipSetDevice(device_));
  cur_ = c;
  // the next round's planning waits on this: poll rather than sleep
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t q = hipEventQuery(cx.done);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) HIP(q);
  }
  last_wait_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  cx.pending = false;
  g_timer = &timer;
  for (size_t k = 0; k < cx.active.size(); ++k) {
    Job *j = cx.active[k];
    j->release_image();  // the round ran after the job's upload
    j->h_scalars = cx.h_compact + k;
    j->moved = 0;
    for (const Job::EvalReq &r : j->evals) {
      if (!r.skip) j->moved = std::max(j->moved, j->h_scalars->eval[r.scalar_slot].moved);
      // the timed gradient launch did not cover a skipped trial
      if (r.trial && j->h_scalars->eval[r.scalar_slot].gradient_skipped && timer.enabled && timer.name == "gather")
        timer.tasks -= 1.0;
    }
    try {
      advance(j);
    } catch (...) {
      j->failure = current_failure();
      j->status = j->failure.status;
      j->phase = Phase::Failed;
    }
  }
  launch_round_targets(cx.active, cx.stream);
  g_timer = nullptr;
}

void Engine::evaluate_once(Job *j, const double *coords, double g, bool density_out) {
  HIP(hipSetDevice(device_));
  if (!j->sorted) sort_now(j);
  const hipStream_t st = stream();
  HIP(hipMemcpyAsync(j->d_eval_coords, coords, j->C * sizeof(double), hipMemcpyHostToDevice, st));
  // slot 0 is overwritten: the optimizer's cached field there is stale now
  if (j->p_slot == 0) j->p_valid = false;
  const Phase saved = j->phase;
  j->evals.clear();
  j->evals.push_back({0, j->d_eval_coords, j->d_grad[0], false, g, false, true, 2, false, 0});
  j->want_precond = j->want_extrapolate = j->want_candidate = false;
  std::vector<Job *> one{j};
  // run exactly the planned request: mark the phase so plan() leaves it alone
  j->phase = Phase::Paused;
  std::vector<Task> tasks;
  const bool fixed_nets = !j->fixed_wa_valid || g != j->fixed_wa_gamma;
  j->fixed_wa_valid = true;
  j->fixed_wa_gamma = g;
  Task t = j->task(0, j->d_eval_coords, j->d_grad[0], nullptr, g, false, &j->d_scalars->eval[2], fixed_nets);
  t.raw_units = j->d_raw_units;
  t.raw_fillers = j->d_raw_fillers;
  t.wa_grad_out = j->d_wa_grad;
  if (density_out) t.density_out = j->d_density_out;
  tasks.push_back(t);
  Plan p_dep, p_conv, p_wa, p_gr;
  std::map<std::tuple<int, bool, bool>, DctGroup> groups;
  if (j->F > 0) {
    p_dep.add((size_t)j->n_keys, 1);
    p_conv.add((size_t)j->F * j->G, BLOCK);
    dct_group(groups, *j).add(0, *j);
  } else {
    p_dep.none();
    p_conv.none();
  }
  p_wa.add((size_t)t.nb_wa, 1);
  p_gr.add((size_t)(t.nb_ug + t.nb_gf), 1);
  std::vector<char> buf;
  const size_t o_tasks = append(buf, tasks), o_dep = append(buf, p_dep.start), o_conv = append(buf, p_conv.start),
               o_wa = append(buf, p_wa.start), o_gr = append(buf, p_gr.start);
  std::vector<std::array<size_t, 4>> o_groups;
  for (auto &kv : groups) {
    DctGroup &dg = kv.second;
    dg.pad(tasks.size());
    o_groups.push_back({append(buf, dg.h.start), append(buf, dg.w.start), append(buf, dg.h_wide.start),
                        append(buf, dg.w_wide.start)});
  }
  HIP(hipStreamSynchronize(st));
  ensure_stage(aux_, buf.size());
  std::memcpy(aux_.h_stage, buf.data(), buf.size());
  HIP(hipMemcpyAsync(aux_.d_stage, aux_.h_stage, buf.size(), hipMemcpyHostToDevice, st));
  char *const d_stage_ = aux_.d_stage;
  const Task *d_tasks = reinterpret_cast<const Task *>(d_stage_ + o_tasks);
  Maps mp{};
  mp.deposit = p_dep.device(d_stage_, o_dep);
  mp.convert = p_conv.device(d_stage_, o_conv);
  mp.wa = p_wa.device(d_stage_, o_wa);
  mp.gradient = p_gr.device(d_stage_, o_gr);
  mp.n_tasks = 1;
  mp.n_density = j->F > 0 ? 1 : 0;
  mp.n_gradient = 1;
  mp.wa_evaluations = 1.0;
  g_timer = &timer;
  launch_wa(d_tasks, mp, st);
  launch_density(d_tasks, mp, st);
  size_t gi = 0;
  for (auto &kv : groups) {
    const size_t off[4] = {o_groups[gi][0], o_groups[gi][1], o_groups[gi][2], o_groups[gi][3]};
    launch_poisson(d_tasks, kv.second.device(d_stage_, off), st);
    ++gi;
  }
  launch_gather(d_tasks, mp, st);
  launch_combine(d_tasks, mp, st);
  g_timer = nullptr;
  HIP(hipGetLastError());
  HIP(hipMemcpyAsync(h_pool_ + j->pool_index, d_pool_ + j->pool_index, sizeof(JobScalars), hipMemcpyDeviceToHost,
                     st));
  HIP(hipStreamSynchronize(st));
  j->h_scalars = h_pool_ + j->pool_index;
  j->phase = saved;
  j->evals.clear();
}

bool Engine::precondition_once(Job *j, const std::vector<double> &lambda) {
  HIP(hipSetDevice(device_));
  const hipStream_t st = stream();
  OptTask o{};
  o.n = j->C;
  o.n_units = (int)j->U;
  o.n_fillers = (int)j->Fl;
  o.F = j->F;
  o.DU = (int)j->DU;
  o.lower = j->d_lower;
  o.upper = j->d_upper;
  o.wl_diag = j->d_wl_diag;
  o.du_fcharge = j->d_du_fcharge;
  o.fl_charge = j->d_fl_charge;
  o.unit_du = j->d_unit_du;
  o.fl_field = j->d_fl_field;
  for (int f = 0; f < MAXF; ++f) o.lambda[f] = f < j->F ? lambda[f] : 0.0;
  o.metric = j->d_metric;
  o.part_metric = j->d_part_metric;
  o.metric_scalars = &j->d_scalars->metric;
  o.do_precond = 1;
  Plan pre, none;
  pre.add(j->U + j->Fl, BLOCK);
  none.none();
  std::vector<OptTask> tasks{o};
  std::vector<char> buf;
  const size_t o_t = append(buf, tasks), o_pre = append(buf, pre.start), o_none = append(buf, none.start);
  HIP(hipStreamSynchronize(st));
  ensure_stage(aux_, buf.size());
  std::memcpy(aux_.h_stage, buf.data(), buf.size());
  HIP(hipMemcpyAsync(aux_.d_stage, aux_.h_stage, buf.size(), hipMemcpyHostToDevice, st));
  OptMaps om{};
  om.precond = pre.device(aux_.d_stage, o_pre);
  om.extrapolate = none.device(aux_.d_stage, o_none);
  om.candidate = none.device(aux_.d_stage, o_none);
  om.n_tasks = 1;
  const OptTask *d_opt = reinterpret_cast<const OptTask *>(aux_.d_stage + o_t);
  launch_optimizer(d_opt, om, st);
  launch_optimizer_finalize(d_opt, om, st);
  HIP(hipGetLastError());
  MetricScalars ms{};
  HIP(hipMemcpyAsync(&ms, &j->d_scalars->metric, sizeof(ms), hipMemcpyDeviceToHost, st));
  HIP(hipStreamSynchronize(st));
  return ms.nonfinite == 0;
}

}  // namespace texo
