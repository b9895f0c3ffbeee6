// spx_engine.hip -- C ABI of the MI355X SPH engine (libspheral_amd.so).
//
// spx_run: one FluidNodeList evolved by SPH (+ MonaghanGingold Q + SPHSmoothingScale,
// + TreeGravity when enabled) under CheapSynchronousRK2 -- a one-slot batch.
// spx_batch: many independent runs of one problem type stepped together.
#include <cstring>

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/spx_abi.h"
#include "spx_batch.hpp"

namespace spx {
static thread_local std::string g_last_error;

// Plain streaming copy, 16 bytes per lane per iteration: the reference point for the
// achievable HBM bandwidth and for scaling the memory-traffic counters.
__global__ void __launch_bounds__(kBlock) k_copy(const double2* __restrict__ src, double2* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x*blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x*blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Device-resident scratch of spx_run_num_pairs_noalloc: per-block partial counts plus the total.
// Each GPU the library runs on holds one copy (512 KiB) until the process ends.  Host code never
// passes `scratch` itself to a kernel: it passes the address hipGetSymbolAddress returns for the
// current device (noalloc_scratch).  Every partial a call reads is written by the same call, so
// nothing needs clearing between calls.
constexpr int kNoAllocBlocks = 65536;            // up to 65536*256 = 16,777,216 particles
__device__ __align__(256) unsigned long long scratch[kNoAllocBlocks + 1];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with that device
// current (the caller's current device is restored afterwards).
static unsigned long long* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) raise(SPX_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    HIP_CHECK(hipSetDevice(prev));
    HIP_CHECK(e);
    if (!p) raise(SPX_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = static_cast<unsigned long long*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace spx

struct spx_run {
  std::unique_ptr<spx::Batch> impl;
  int64_t n = 0;
  bool haveState = false;
};

struct spx_batch {
  std::unique_ptr<spx::Batch> impl;
};

using spx::g_last_error;
using spx::SpxError;
using spx::raise;

template <typename F>
static int guarded(F&& fn) {
  try {
    fn();
    return SPX_OK;
  } catch (const SpxError& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SPX_ERR_ARG;
  } catch (...) {
    g_last_error = "unknown error";
    return SPX_ERR_ARG;
  }
}

#define RUN_OR_FAIL(run)                                                    \
  do {                                                                      \
    if (!(run) || !(run)->impl) { g_last_error = "null run"; return SPX_ERR_ARG; } \
  } while (0)

extern "C" {

const char* spx_last_error(void) { return g_last_error.c_str(); }

int spx_config_default(int32_t dim, spx_config* c) {
  if (!c || dim < 1 || dim > 3) { g_last_error = "bad arguments"; return SPX_ERR_ARG; }
  std::memset(c, 0, sizeof(*c));
  c->dim = dim;
  c->kernel = SPX_KERNEL_BSPLINE;
  c->kernel_points = 1000;
  c->precision = SPX_PREC_FP64;
  c->min_nperh = 0.25;
  c->max_nperh = 64.0;
  c->nperh = 2.01;
  c->neighbor_extent = 2.0;
  c->hmin = 1.0e-20;
  c->hmax = 1.0e20;
  c->rho_min = 1.0e-10;
  c->rho_max = 1.0e10;
  c->gamma = 5.0/3.0;
  c->Cl = 1.0;
  c->Cq = 1.0;
  c->cfl = 0.25;
  c->compatible_energy = 1;
  c->gradh_correction = 1;
  c->correct_velocity_gradient = 1;
  c->dt_min = 0.0;
  c->dt_max = 1.7976931348623157e308;
  c->dt_growth = 2.0;
  c->last_dt = 1.0e-5;
  for (int a = 0; a < 3; ++a) { c->xmin[a] = 0.0; c->xmax[a] = 1.0; }
  c->reorder_interval = 1;
  c->G = 1.0;
  c->softening = 0.01;
  c->opening = 0.5;
  c->ftimestep = 0.1;
  c->kernel_order = 5;
  c->viscosity = SPX_Q_MONAGHAN_GINGOLD;
  c->eta_crit_frac = 1.0;
  c->eta_fold_frac = 0.2;
  c->integrator = SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2;
  c->gravity_only = 0;
  c->gravity_time_step = SPX_GRAVITY_DT_ACCELERATION_RATIO;
  return SPX_OK;
}

int spx_device_count(int32_t* out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (out) *out = n;
  return SPX_OK;
}

int spx_table_eval(const spx_config* cfg, const double* eta, int64_t n, double* W, double* gW, double* nperh) {
  return guarded([&] {
    if (!cfg || !eta || !W || !gW) spx::raise(SPX_ERR_ARG, "null argument");
    spx::KernelTable t;
    t.build(cfg->kernel, cfg->dim, (unsigned)cfg->kernel_points, cfg->min_nperh, cfg->max_nperh, cfg->kernel_order);
    for (int64_t i = 0; i < n; ++i) {
      t.kernelAndGrad(eta[i], 1.0, W[i], gW[i]);
      if (nperh) nperh[i] = t.nodesPerh(eta[i]);
    }
  });
}

int spx_table_coefficients(const spx_config* cfg, int64_t* ncoef, double* wc, double* gwc, double* xstep) {
  return guarded([&] {
    if (!cfg || !ncoef) spx::raise(SPX_ERR_ARG, "null argument");
    spx::KernelTable t;
    t.build(cfg->kernel, cfg->dim, (unsigned)cfg->kernel_points, cfg->min_nperh, cfg->max_nperh, cfg->kernel_order);
    *ncoef = (int64_t)t.W.coef.size();
    if (wc) std::memcpy(wc, t.W.coef.data(), t.W.coef.size()*sizeof(double));
    if (gwc) std::memcpy(gwc, t.gW.coef.data(), t.gW.coef.size()*sizeof(double));
    if (xstep) *xstep = t.W.xstep;
  });
}

static void check_device(int32_t device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) spx::raise(SPX_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) spx::raise(SPX_ERR_ARG, "device index out of range");
}

static void kernel_times(spx::Batch& B, double* ms, int32_t max_entries, int32_t* n_entries, char* names,
                         int64_t names_cap) {
  if (max_entries < 0) { B.timing = true; B.tnames.clear(); B.tms.clear(); return; }   // enable timing, clear sums
  const int32_t k = (int32_t)std::min<size_t>(B.tms.size(), (size_t)max_entries);
  std::string all;
  for (int32_t i = 0; i < k; ++i) { if (ms) ms[i] = B.tms[i]; all += (i ? "," : "") + B.tnames[i]; }
  if (n_entries) *n_entries = k;
  if (names && names_cap > 0) {
    const size_t c = std::min<size_t>(all.size(), (size_t)names_cap - 1);
    std::memcpy(names, all.data(), c);
    names[c] = 0;
  }
}

#define RUN_STATE(run) \
  do { if (!(run)->haveState) spx::raise(SPX_ERR_STATE, "no state uploaded"); } while (0)

int spx_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps) {
  return guarded([&] {
    if (bytes < 16 || reps < 1 || !gbps) spx::raise(SPX_ERR_ARG, "bad arguments");
    check_device(device);
    HIP_CHECK(hipSetDevice(device));
    const int64_t n = bytes/16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (st) (void)hipStreamDestroy(st);
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
    };
    try {
      HIP_CHECK(hipMalloc(&a, n*16));
      HIP_CHECK(hipMalloc(&b, n*16));
      HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      HIP_CHECK(hipEventCreate(&e0));
      HIP_CHECK(hipEventCreate(&e1));
      HIP_CHECK(hipMemsetAsync(a, 0, n*16, st));
      const unsigned grid = 256*32;
      hipLaunchKernelGGL(spx::k_copy, dim3(grid), dim3(spx::kBlock), 0, st, (const double2*)a, (double2*)b, n);  // warm-up
      HIP_CHECK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(spx::k_copy, dim3(grid), dim3(spx::kBlock), 0, st, (const double2*)a, (double2*)b, n);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipEventRecord(e1, st));
      HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      *gbps = 2.0*(double)n*16.0*reps/(ms*1.0e-3)/1.0e9;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int spx_run_create(const spx_config* cfg, int64_t n, int32_t device, spx_run** out) {
  return guarded([&] {
    if (!cfg || !out) spx::raise(SPX_ERR_ARG, "null argument");
    check_device(device);
    if (n < 1 || n > (int64_t)spx::kIdxMask) spx::raise(SPX_ERR_CAPACITY, "node count must be in [1, 2^26)");
    auto r = std::make_unique<spx_run>();
    r->impl = std::make_unique<spx::Batch>(*cfg, n, 1, device, false);
    r->n = n;
    *out = r.release();
  });
}

void spx_run_destroy(spx_run* run) { delete run; }

int spx_run_set_state(spx_run* run, const double* x, const double* v, const double* m, const double* rho,
                      const double* eps, const double* H) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    auto& B = *run->impl;
    HIP_CHECK(hipStreamSynchronize(B.stream));
    run->haveState = false;
    B.runs[0] = spx::Batch::RunInfo();
    B.nTotal = B.nActive = 0;
    B.connValid = B.initialized = B.wrapPending = false;
    B.addRun(x, v, m, rho, eps, H, run->n, std::numeric_limits<double>::infinity(), 0);
    run->haveState = true;
  });
}

int spx_run_update_state(spx_run* run, const double* x, const double* v, const double* m, const double* rho,
                         const double* eps, const double* H) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    run->impl->updateState(x, v, m, rho, eps, H);
  });
}

int spx_run_get_field(spx_run* run, const char* name, double* out, int64_t capacity, int64_t* count) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!name || !out) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    const int64_t c = run->impl->getField(0, name, out, capacity);
    if (count) *count = c;
  });
}

int spx_run_initialize(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    run->impl->initialize();
  });
}

int spx_run_iterate_ideal_h(spx_run* run, int32_t max_iterations, double tolerance, int32_t* passes) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    const int k = run->impl->iterateIdealH((int)max_iterations, tolerance);
    if (passes) *passes = k;
  });
}

int spx_run_step(spx_run* run, double max_time, double* dt_out) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    auto& B = *run->impl;
    B.setEndTime(0, max_time);
    B.step();
    if (dt_out) *dt_out = B.hsc[0].lastDt;
  });
}

int spx_run_advance(spx_run* run, double max_time, int64_t nsteps, int64_t* steps_done, double* t_out) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    auto& B = *run->impl;
    B.setEndTime(0, max_time);
    int64_t k = 0;
    while (k < nsteps && !spx::reached_end(B.hsc[0].t, max_time)) { B.step(); ++k; }
    if (steps_done) *steps_done = k;
    if (t_out) *t_out = B.hsc[0].t;
  });
}

int spx_run_time(spx_run* run, double* t, int64_t* cycle, double* last_dt) {
  RUN_OR_FAIL(run);
  const spx::Scalars& s = run->impl->hsc[0];
  if (t) *t = s.t;
  if (cycle) *cycle = s.cycle;
  if (last_dt) *last_dt = s.lastDt;
  return SPX_OK;
}

int spx_run_update_connectivity(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    run->impl->buildConnectivity(false, true);
    HIP_CHECK(hipStreamSynchronize(run->impl->stream));
  });
}

int spx_run_neighbor_counts(spx_run* run, int32_t* counts) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!counts) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    run->impl->neighbors(0, nullptr, nullptr, nullptr, 0, counts);
  });
}

int spx_run_neighbors(spx_run* run, int64_t* offsets, int64_t* nbr, int32_t* img, int64_t capacity) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!offsets || !nbr || !img) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    run->impl->neighbors(0, offsets, nbr, img, capacity, nullptr);
  });
}

int spx_run_sum_density(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] { RUN_STATE(run); run->impl->sumDensity(); });
}

int spx_run_update_eos(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    auto& B = *run->impl;
    B.eos();
    HIP_CHECK(hipStreamSynchronize(B.stream));
  });
}

int spx_run_post_state_update(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    auto& B = *run->impl;
    B.requireConn();
    B.copyQVelocityGradient();
    B.pack();
    B.omegaPass(false);
    HIP_CHECK(hipStreamSynchronize(B.stream));
  });
}

int spx_run_evaluate_derivatives(spx_run* run, double, double) {
  RUN_OR_FAIL(run);
  return guarded([&] { RUN_STATE(run); run->impl->evaluateDerivatives(); });
}

int spx_run_dt(spx_run* run, double* hydro_dt, double* gravity_dt) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    RUN_STATE(run);
    double h = 0, g = 0;
    run->impl->hydroDtVote(0, &h, &g);
    if (hydro_dt) *hydro_dt = h;
    if (gravity_dt) *gravity_dt = g;
  });
}

int spx_run_compatible_energy(spx_run* run, double multiplier) {
  RUN_OR_FAIL(run);
  return guarded([&] { RUN_STATE(run); run->impl->compatibleEnergy(multiplier); });
}

int spx_run_conservation(spx_run* run, double* out9) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!out9) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    run->impl->conservation(0, out9);
  });
}

int spx_run_kernel_times(spx_run* run, double* ms, int32_t max_entries, int32_t* n_entries, char* names, int64_t names_cap) {
  RUN_OR_FAIL(run);
  return guarded([&] { kernel_times(*run->impl, ms, max_entries, n_entries, names, names_cap); });
}

int spx_run_num_pairs(spx_run* run, int64_t* npairs) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!npairs) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    *npairs = run->impl->numPairs();
  });
}

int spx_run_num_pairs_noalloc(spx_run* run, int64_t* npairs) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!npairs) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    auto& B = *run->impl;
    const int64_t nb = (B.nActive + spx::kBlock - 1)/spx::kBlock;
    if (nb > spx::kNoAllocBlocks)
      spx::raise(SPX_ERR_CAPACITY, "spx_run_num_pairs_noalloc: more than 16,777,216 active particles "
                                   "(the scratch array's limit); use spx_run_num_pairs");
    unsigned long long* p = spx::noalloc_scratch(B.device);
    *npairs = B.numPairs(p);
  });
}

int spx_run_device_bytes(spx_run* run, int64_t* bytes) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!bytes) spx::raise(SPX_ERR_ARG, "null argument");
    *bytes = run->impl->deviceBytes + run->impl->grav.bytes;
  });
}

int spx_run_gravity_interactions(spx_run* run, int64_t* out2) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!out2) spx::raise(SPX_ERR_ARG, "null argument");
    auto& B = *run->impl;
    if (!B.cfg.gravity) spx::raise(SPX_ERR_STATE, "the run has no TreeGravity");
    B.grav.countInteractions(B.cfg.softening, B.cfg.opening, out2);
  });
}

int spx_run_set_step_graphs(spx_run* run, int32_t enable) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    run->impl->useGraphs = enable != 0;
    if (!enable) run->impl->dropGraphs();
  });
}

int spx_run_set_end_of_step_omega(spx_run* run, int32_t eager) {
  RUN_OR_FAIL(run);
  return guarded([&] { run->impl->eagerOmega = eager != 0; });
}

int spx_run_set_group_sizes(spx_run* run, const int32_t* sizes) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    int s[3] = {0, 0, 0};
    if (sizes) for (int p = 0; p < 3; ++p) s[p] = sizes[p];
    run->impl->setGroupSizes(s);
  });
}

int spx_run_step_stats(spx_run* run, int64_t* out9) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!out9) spx::raise(SPX_ERR_ARG, "null argument");
    run->impl->stepStats(out9);
  });
}

int spx_run_neighbor_builds(spx_run* run, int64_t* builds) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!builds) spx::raise(SPX_ERR_ARG, "null argument");
    RUN_STATE(run);
    *builds = run->impl->runs[0].builds;
  });
}

// ---------------------------------------------------------------- batches
#define BATCH_OR_FAIL(b)                                                           \
  do {                                                                             \
    if (!(b) || !(b)->impl) { g_last_error = "null batch"; return SPX_ERR_ARG; }   \
  } while (0)

int spx_batch_create(const spx_config* cfg, int64_t capacity, int32_t max_runs, int32_t device, spx_batch** out) {
  return guarded([&] {
    if (!cfg || !out) spx::raise(SPX_ERR_ARG, "null argument");
    // a TreeGravity-only problem is stepped as a single run (spx_run): the batch's start-up and
    // end-time retirement ride on the hydro passes it does not have
    if (cfg->gravity_only) spx::raise(SPX_ERR_UNSUPPORTED, "TreeGravity-only problems run as single runs (spx_run)");
    check_device(device);
    auto b = std::make_unique<spx_batch>();
    b->impl = std::make_unique<spx::Batch>(*cfg, capacity, max_runs, device, true);
    *out = b.release();
  });
}

void spx_batch_destroy(spx_batch* batch) { delete batch; }

int spx_batch_add_run(spx_batch* batch, const double* x, const double* v, const double* m, const double* rho,
                      const double* eps, const double* H, int64_t n, double t_end, int32_t* run) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    const int slot = batch->impl->addRun(x, v, m, rho, eps, H, n, t_end);
    if (run) *run = slot;
  });
}

int spx_batch_remove_run(spx_batch* batch, int32_t run) {
  BATCH_OR_FAIL(batch);
  return guarded([&] { batch->impl->removeRun(run); });
}

int spx_batch_step(spx_batch* batch, int32_t* active_runs) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    auto& B = *batch->impl;
    if (B.activeRuns() > 0) B.step();
    if (active_runs) *active_runs = B.activeRuns();
  });
}

int spx_batch_run_status(spx_batch* batch, int32_t run, int32_t* state, double* t, int64_t* cycle, double* last_dt,
                         int64_t* n) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    auto& B = *batch->impl;
    B.checkSlot(run);
    if (state) *state = B.runs[run].state;
    if (t) *t = B.hsc[run].t;
    if (cycle) *cycle = B.hsc[run].cycle;
    if (last_dt) *last_dt = B.hsc[run].lastDt;
    if (n) *n = B.runs[run].n;
  });
}

int spx_batch_get_field(spx_batch* batch, int32_t run, const char* name, double* out, int64_t capacity, int64_t* count) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    if (!name || !out) spx::raise(SPX_ERR_ARG, "null argument");
    const int64_t c = batch->impl->getField(run, name, out, capacity);
    if (count) *count = c;
  });
}

int spx_batch_conservation(spx_batch* batch, int32_t run, double* out9) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    if (!out9) spx::raise(SPX_ERR_ARG, "null argument");
    batch->impl->conservation(run, out9);
  });
}

int spx_batch_usage(spx_batch* batch, int64_t* particles, int32_t* runs) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    // runs that hold state (active or done); removed runs' slots and particles are reclaimed by
    // the next admission or step, so they count as free
    auto& B = *batch->impl;
    int32_t k = 0;
    int64_t np = 0;
    for (auto& r : B.runs)
      if (r.state == spx::RUN_ACTIVE || r.state == spx::RUN_DONE) { ++k; np += r.n; }
    if (particles) *particles = np;
    if (runs) *runs = k;
  });
}

int spx_batch_device_bytes(spx_batch* batch, int64_t* bytes) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    if (!bytes) spx::raise(SPX_ERR_ARG, "null argument");
    *bytes = batch->impl->deviceBytes + batch->impl->grav.bytes;
  });
}

int spx_batch_set_step_graphs(spx_batch* batch, int32_t enable) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    batch->impl->useGraphs = enable != 0;
    if (!enable) batch->impl->dropGraphs();
  });
}

int spx_batch_set_group_sizes(spx_batch* batch, const int32_t* sizes) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    int s[3] = {0, 0, 0};
    if (sizes) for (int p = 0; p < 3; ++p) s[p] = sizes[p];
    batch->impl->setGroupSizes(s);
  });
}

int spx_batch_step_stats(spx_batch* batch, int64_t* out9) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    if (!out9) spx::raise(SPX_ERR_ARG, "null argument");
    batch->impl->stepStats(out9);
  });
}

int spx_batch_neighbor_builds(spx_batch* batch, int32_t run, int64_t* builds) {
  BATCH_OR_FAIL(batch);
  return guarded([&] {
    if (!builds) spx::raise(SPX_ERR_ARG, "null argument");
    batch->impl->checkSlot(run);
    *builds = batch->impl->runs[run].builds;
  });
}

int spx_batch_kernel_times(spx_batch* batch, double* ms, int32_t max_entries, int32_t* n_entries, char* names,
                           int64_t names_cap) {
  BATCH_OR_FAIL(batch);
  return guarded([&] { kernel_times(*batch->impl, ms, max_entries, n_entries, names, names_cap); });
}

}  // extern "C"
