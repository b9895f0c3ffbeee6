// spx_batch.hpp -- host orchestration: a batch of independent runs of one problem type,
// resident in HBM and advanced together by CheapSynchronousRK2, each with its own dt.
//
// A single spx_run is a batch with one slot whose end time follows the caller's
// maxTime.  Runs of a batch share the physics configuration (kernel, EOS, viscosity,
// boundary planes) and differ in particle count, state, cell grid, dt and end time.
// Per step the host enqueues, on the batch's stream, the reference's call order
// (Integrator/Integrator.cc:67-111, Integrator/CheapSynchronousRK2.cc:40-132) and
// reads back three small things: the per-run extents (to size the cell grids), the
// neighbour-capacity overflow word, and the per-run times.
//
// Particle order in the arrays: the active runs first, sorted by (run, cell Morton
// code), then the finished runs, then removed runs until they are squeezed out.
//
// Neighbour lists are filtered every step from per-run candidate lists (k_candidates,
// k_neighbor_filter).  A run re-sorts and rebuilds its candidates only when its own nodes
// have moved or grown enough that the list might miss a pair (measured in k_full_update), so the
// rebuild schedule of a run, and with it every bit of its results, does not depend on
// the other runs of the batch.
#pragma once

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "spx_error.hpp"
#include "spx_gravity.hpp"
#include "spx_sph.hpp"
#include "spx_table.hpp"

namespace spx {

static inline int nblocks(int64_t n) { return (int)((n + kBlock - 1)/kBlock); }

// Largest kernel table a pair pass stages in LDS (three 256-thread blocks per CU still fit).
constexpr size_t kLdsTableMax = 48*1024;

// Particles per block of the extent reduction.
constexpr int64_t kExtentTile = 8192;

// Widening of the candidate criterion (ext (1 + skin)) of a run's first candidate list,
// and the range the per-run choice stays in.
constexpr double kSkinInit = 0.1, kSkinMin = 0.03, kSkinMax = 0.5;

// Skin that minimises the modelled per-step cost of a run whose nodes use up the skin at
// `rate` per step: filtering (1 + s)^D candidates every step, plus a sort and candidate
// search (fixed part + (1 + s)^D part, about 1 and 5 filters' worth) every s/rate steps.
static inline double choose_skin(double rate, int D) {
  double best = kSkinInit, bestCost = 1e300;
  for (double s = kSkinMin; s <= kSkinMax + 1e-12; s += 0.01) {
    const double vol = std::pow(1.0 + s, D);
    const double cost = vol + (1.0 + 5.0*vol)*rate/s;
    if (cost < bestCost) { bestCost = cost; best = s; }
  }
  return best;
}

// Value of an order-preserving bit key written by a max-reduction (0: nothing measured).
static inline double skin_value(unsigned long long b) { return b == 0ull ? 0.0 : ord_value(b); }

// A run has reached its end time (a clamped last step may land an ulp short).
static inline bool reached_end(double t, double tEnd) { return t >= tEnd || tEnd - t <= 1.0e-15*std::fabs(tEnd); }

// EOS of the nodes of newly admitted runs (runMask[run] != 0).
__global__ void __launch_bounds__(kBlock) k_eos_masked(Fields f, Phys ph, const int* runMask, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n || !runMask[f.run[i]]) return;
  f.P[i] = ph.g1*f.rho[i]*f.eps[i];
  f.cs[i] = sqrt(fmax(0.0, ph.gamma*ph.g1*f.eps[i]));
}

class Batch {
 public:
  struct RunInfo {
    int state = RUN_FREE;
    int64_t n = 0;
    int64_t offset = 0;     // first slot of the run's contiguous particle segment
    bool needInit = false;
    bool stepOut = false;   // "new H", "mass zeroth moment" of this run's last evaluation are in place
    bool fullOut = false;   // the evaluate-only outputs (sum density, M, ...) are in place
    bool candValid = false; // the run's candidate list covers its nodes (see k_candidates)
    double skin = kSkinInit;
    long long builtCycle = 0;   // the run's step count when its candidate list was built
    int64_t builds = 0;         // candidate-list builds of the run
  };

  spx_config cfg;
  int D;
  int device;
  int64_t cap;           // particle capacity of the batch
  int maxRuns;
  bool retire;           // batch mode: a run that reaches its end time stops; single-run mode: never
  hipStream_t stream = nullptr;
  KernelTable table;
  DevTable dtab{};
  Phys ph{};
  Geo geo{};
  Fields f{};
  Fields alt{};          // reorder targets of the carried arrays
  GravityTree grav;
  std::vector<RunInfo> runs;
  std::vector<void*> allocs;
  std::vector<size_t> allocBytes;   // size of each entry of allocs
  int64_t deviceBytes = 0;          // device memory the batch holds (besides the tree's)
  // device side
  RunGrid* dgrids = nullptr;
  Scalars* dsc = nullptr;
  unsigned long long* dext = nullptr;
  int* dstate = nullptr;
  int* dorder = nullptr;         // place of each active run in the batch (segment order)
  int* horder = nullptr;
  int2* dseg = nullptr;          // rebuilding runs' segments: (first position, compacted start)
  int2* hseg = nullptr;
  int* dinit = nullptr;
  int* dOverflow = nullptr;
  int* dbuild = nullptr;      // per run: 1 if it builds its candidate list this step
  int* dsub = nullptr;        // per run: segment start, compacted start, segment index
  uint64_t *skeys = nullptr, *skeysAlt = nullptr;   // compacted sort of the rebuilding runs
  uint32_t *spos = nullptr, *sposAlt = nullptr;
  int* ddelta = nullptr;      // per run: shift of its segment in a reorder (k_remap_candidates)
  uint32_t* candAlt = nullptr;
  int32_t* ccntAlt = nullptr;
  int candAltCap = 0;
  uint32_t* tlist = nullptr;  // candidate entries handed over by partners (k_candidates)
  int* tcnt = nullptr;
  int tcap = 0;
  uint64_t *keys = nullptr, *keysAlt = nullptr;
  uint32_t *vals = nullptr, *valsAlt = nullptr;
  void* sortTmp = nullptr;
  size_t sortTmpBytes = 0;
  double* stage = nullptr;    // caller-order staging of one state component (updateState)
  int* cellStart = nullptr;
  int* cellEnd = nullptr;
  int64_t cellsAlloc = 0;
  int64_t nCells = 0;
  int* dtiles = nullptr;      // (run, begin, end) particle tiles of the active runs
  int* htiles = nullptr;
  int64_t tilesAlloc = 0;
  int64_t* oidAlt = nullptr;
  int32_t* runAlt = nullptr;
  // pinned host mirrors
  RunGrid* hgrids = nullptr;
  Scalars* hsc = nullptr;
  unsigned long long* hext = nullptr;
  int* hstate = nullptr;
  int* hinit = nullptr;
  int* hOverflow = nullptr;
  int* hbuild = nullptr;
  int* hdelta = nullptr;
  int* hsub = nullptr;
  int ld = 0;
  int mortonBits = 0, runBits = 0;
  int64_t nActive = 0;    // particles of active runs (front of the arrays after a sort)
  int64_t nTotal = 0;     // particles resident (active, done, removed-but-not-yet-dropped)
  bool initialized = false, connValid = false, wrapPending = false;
  bool orderDirty = false;  // a run was added, finished or removed since the last reorder
  // The end-of-step grad-h Omega (SPHBase::postStateUpdate) of the last CheapSynchronousRK2 step
  // is not yet computed: nothing in the next such step reads it, so it is computed only when
  // something else needs it first (ensureOmega), from the end-of-step positions and the step's
  // neighbour lists, exactly as at the end of the step.
  bool omegaStale = false;
  bool eagerOmega = false;           // compute it at the end of every step anyway
  long long inPlaceRetirements = 0;   // steps whose finished runs retired without a reorder
  long long fullReorders = 0;         // reorders of the whole batch (composition changes)
  int64_t candBuilds = 0;   // candidate-list builds, summed over runs
  int64_t npairsCache = -1;
  unsigned long long* pairScratch = nullptr;   // numPairs' partial counts (library-allocated)
  int pairScratchBlocks = 0;
  bool mixed = false;
  // Neighbour-group sizes of the three per-step gather passes (how many neighbours' loads a lane
  // issues before it uses any of them): 0 the filtering sum-density + dt pass, 1 the grad-h
  // passes, 2 the compatible-energy update.  Every size performs the same arithmetic in the same
  // order, so the choice never changes a result bit; it trades latency hiding against registers
  // (waves per SIMD), and which size wins depends on the dimension, the precision, the neighbour
  // count and the particle count.  The batch measures instead of guessing: once a composition
  // has run two warm-up steps it alternates sizes 4 and 8 of every pass over six steps (HIP
  // events around each pass, three samples per size), and moves a pass off its default size
  // (4, 8, 8: measured on Taylor-Green 128^3) only when the other size's fastest sample is at
  // least 3 % faster; it measures again once the active particle count has changed by more
  // than 2x.  Callers may fix the sizes (setGroupSizes).
  struct GroupTuner {
    static constexpr int kPasses = 3;
    static constexpr int kSteps = 6;            // measured steps: sizes 4, 8, 4, 8, 4, 8
    static constexpr int kWarm = 2;             // steps run before measuring
    static constexpr int kDefault[kPasses] = {kGroup, kGroupOmega, kGroupUpdate};
    int size[kPasses] = {kGroup, kGroupOmega, kGroupUpdate};
    bool fixed = false;
    int warm = 0;                   // > 0: warm-up steps left before measuring
    int phase = -1;                 // -1: not measuring; else the measurement step (0..kSteps-1)
    int64_t tunedFor = 0;           // active particles when the sizes were last measured
    float best[kPasses][2];         // fastest time seen per pass and size (4, 8)
    int64_t tunings = 0;
  } tuner;
  hipEvent_t tuneEv[8] = {};
  // Steps replayed from a captured HIP graph (see graphStep) and the captured graphs.
  struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
  };
  StepGraph graphs[2];
  int graphNext = 0;
  std::vector<std::vector<unsigned char>> seenKeys = std::vector<std::vector<unsigned char>>(2);
  int seenNext = 0;
  bool useGraphs = true;
  int64_t graphCaptures = 0, graphReplays = 0, eagerSteps = 0;
  int* dOverflowG = nullptr;  // the filter's overflow flag in graph-replayed steps
  // per-class kernel time (ms) summed over the steps since timing was enabled
  bool timing = false;
  // TreeGravity's own stream and its fork / join events (gravity runs only)
  hipStream_t gstream = nullptr;
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  bool gravPending = false;
  std::vector<std::string> tnames;
  std::vector<double> tms;
  std::vector<hipEvent_t> evPool;       // reused across steps (recording costs no allocation)
  std::vector<const char*> evNames;
  size_t evCount = 0;

  Batch(const spx_config& c, int64_t capacity, int nruns, int dev, bool retireDone)
      : cfg(c), D(c.dim), device(dev), cap(capacity), maxRuns(nruns), retire(retireDone) {
    validate();
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    table.build(cfg.kernel, D, (unsigned)cfg.kernel_points, cfg.min_nperh, cfg.max_nperh, cfg.kernel_order);
    uploadTable();
    setPhys();
    mixed = cfg.precision == SPX_PREC_MIXED;
    ld = (int)((cap + 255)/256*256);
    runs.resize(maxRuns);
    runBits = 0;
    while ((1 << runBits) < maxRuns) ++runBits;
    allocate();
    grav.init(D, cap, ld, stream, cfg.gravity != 0);
    if (cfg.gravity) {
      // the midpoint tree build and walk run on a stream of their own (default priority) beside
      // the midpoint grad-h and SPH passes, which neither read nor write what they use
      HIP_CHECK(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
      HIP_CHECK(hipEventCreateWithFlags(&evFork, hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&evJoin, hipEventDisableTiming));
    }
  }
  ~Batch() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto e : evPool) (void)hipEventDestroy(e);
    for (auto e : tuneEv) if (e) (void)hipEventDestroy(e);
    dropGraphs();
    for (void* p : allocs) (void)hipFree(p);
    for (void* p : {(void*)hgrids, (void*)hsc, (void*)hext, (void*)hstate, (void*)hinit, (void*)hOverflow,
                    (void*)htiles, (void*)hbuild, (void*)hdelta, (void*)hsub})
      if (p) (void)hipHostFree(p);
    if (gstream) (void)hipStreamSynchronize(gstream);
    grav.release();
    if (evFork) (void)hipEventDestroy(evFork);
    if (evJoin) (void)hipEventDestroy(evJoin);
    if (gstream) (void)hipStreamDestroy(gstream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void validate() {
    if (cfg.dim < 1 || cfg.dim > 3) raise(SPX_ERR_ARG, "dim must be 1, 2 or 3");
    if (cap < 1 || cap > (int64_t)kIdxMask) raise(SPX_ERR_CAPACITY, "particle capacity must be in [1, 2^26)");
    if (maxRuns < 1 || maxRuns > 32768) raise(SPX_ERR_CAPACITY, "run slots must be in [1, 32768]");
    if (cfg.kernel != SPX_KERNEL_BSPLINE && cfg.kernel != SPX_KERNEL_WENDLANDC4 && cfg.kernel != SPX_KERNEL_NBSPLINE)
      raise(SPX_ERR_ARG, "unknown kernel");
    if (cfg.kernel == SPX_KERNEL_NBSPLINE && (cfg.kernel_order < 2 || cfg.kernel_order > 9))
      raise(SPX_ERR_ARG, "NBSplineKernel order must be in [2, 9]");
    if (cfg.viscosity != SPX_Q_MONAGHAN_GINGOLD && cfg.viscosity != SPX_Q_LIMITED_MONAGHAN_GINGOLD)
      raise(SPX_ERR_ARG, "unknown viscosity");
    if (cfg.viscosity == SPX_Q_LIMITED_MONAGHAN_GINGOLD && !(cfg.eta_fold_frac > 0.0))
      raise(SPX_ERR_ARG, "eta_fold_frac must be positive");
    if (cfg.kernel_points < 3) raise(SPX_ERR_ARG, "kernel_points must be >= 3");
    if (!cfg.compatible_energy) raise(SPX_ERR_UNSUPPORTED, "only compatibleEnergyEvolution=True is implemented");
    if (cfg.precision != SPX_PREC_FP64 && cfg.precision != SPX_PREC_MIXED) raise(SPX_ERR_ARG, "unknown precision");
    if (cfg.neighbor_extent <= 0.0) raise(SPX_ERR_ARG, "neighbor_extent must be positive");
    if (cfg.reorder_interval > 1)
      raise(SPX_ERR_UNSUPPORTED, "the engine re-sorts particles every step (reorder_interval must be 0 or 1)");
    for (int a = 0; a < 3; ++a) {
      const bool per = cfg.bc_lo[a] == SPX_BC_PERIODIC || cfg.bc_hi[a] == SPX_BC_PERIODIC;
      if (per && cfg.bc_lo[a] != cfg.bc_hi[a]) raise(SPX_ERR_ARG, "periodic must be set on both sides of an axis");
      if (a < cfg.dim && (cfg.bc_lo[a] != SPX_BC_NONE || cfg.bc_hi[a] != SPX_BC_NONE) && !(cfg.xmax[a] > cfg.xmin[a]))
        raise(SPX_ERR_ARG, "boundary planes need xmax > xmin");
    }
    if (cfg.gravity && (cfg.G <= 0.0 || cfg.softening <= 0.0 || cfg.opening <= 0.0 || cfg.ftimestep <= 0.0))
      raise(SPX_ERR_ARG, "TreeGravity needs G, softening, opening, ftimestep > 0");
    if (cfg.gravity && maxRuns != 1) raise(SPX_ERR_UNSUPPORTED, "a TreeGravity batch holds one run");
    // TreeGravity's DynamicalTime vote (TreeGravity.cc:239-300, 497-518) needs every point's
    // interaction buckets and the reference's partial-sort order; the engine votes
    // AccelerationRatio only
    if (cfg.gravity && cfg.gravity_time_step != SPX_GRAVITY_DT_ACCELERATION_RATIO)
      raise(SPX_ERR_UNSUPPORTED, "TreeGravity timeStepChoice DynamicalTime is not implemented on the GPU "
                                 "(AccelerationRatio is)");
    if (cfg.integrator != SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2 && cfg.integrator != SPX_INTEGRATOR_SYNCHRONOUS_RK2)
      raise(SPX_ERR_ARG, "unknown integrator");
    if (cfg.gravity_only && (!cfg.gravity || cfg.integrator != SPX_INTEGRATOR_SYNCHRONOUS_RK2))
      raise(SPX_ERR_UNSUPPORTED, "TreeGravity-only problems (gravity_only = 1) need gravity = 1 and SynchronousRK2");
  }

  template <typename T>
  T* dalloc(size_t count) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1)*sizeof(T);
    HIP_CHECK(hipMalloc(&p, bytes));
    allocs.push_back(p);
    allocBytes.push_back(bytes);
    deviceBytes += (int64_t)bytes;
    return static_cast<T*>(p);
  }
  template <typename T>
  T* halloc(size_t count) {
    void* p = nullptr;
    HIP_CHECK(hipHostMalloc(&p, std::max<size_t>(count, 1)*sizeof(T), hipHostMallocDefault));
    std::memset(p, 0, std::max<size_t>(count, 1)*sizeof(T));
    return static_cast<T*>(p);
  }
  // Host -> device on the batch's stream, complete on return.
  void upload(void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  void download(void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  void uploadTable() {
    const size_t nint = table.W.N1 + 1;
    std::vector<double> coef(6*nint);
    std::vector<float> coef32(6*nint);
    for (size_t k = 0; k < nint; ++k)
      for (int c = 0; c < 3; ++c) { coef[6*k + c] = table.W.coef[3*k + c]; coef[6*k + 3 + c] = table.gW.coef[3*k + c]; }
    for (size_t k = 0; k < coef.size(); ++k) coef32[k] = (float)coef[k];
    double* dc = dalloc<double>(coef.size());
    float* dc32 = dalloc<float>(coef32.size());
    double* dh = dalloc<double>(table.nperh.vals.size());
    upload(dc, coef.data(), coef.size()*sizeof(double));
    upload(dc32, coef32.data(), coef32.size()*sizeof(float));
    upload(dh, table.nperh.vals.data(), table.nperh.vals.size()*sizeof(double));
    dtab.coef = dc;
    dtab.coef32 = dc32;
    dtab.herm = dh;
    dtab.n1 = (int)table.W.N1;
    dtab.inv_xstep = 1.0/table.W.xstep;
    dtab.extent = table.extent;
    dtab.W0 = table.W(0.0);
    dtab.hn = (int)table.nperh.N;
    dtab.hxmin = table.nperh.xmin;
    dtab.hxmax = table.nperh.xmax;
    dtab.hxstep = table.nperh.xstep;
  }

  void setPhys() {
    ph.gamma = cfg.gamma;
    ph.g1 = cfg.gamma - 1.0;
    ph.Cl = cfg.Cl;
    ph.Cq = cfg.Cq;
    ph.cfl = cfg.cfl;
    ph.nperh = cfg.nperh;
    ph.ext = cfg.neighbor_extent;
    ph.ext2 = cfg.neighbor_extent*cfg.neighbor_extent;
    ph.hmin = cfg.hmin;
    ph.hmax = cfg.hmax;
    ph.rho_min = cfg.rho_min;
    ph.rho_max = cfg.rho_max;
    ph.hmaxInv = 1.0/cfg.hmax;
    ph.hminInv = 1.0/cfg.hmin;
    ph.correctVG = cfg.correct_velocity_gradient;
    ph.qImages = 1;
    ph.etaCrit = cfg.eta_crit_frac/cfg.nperh;
    ph.etaFold = cfg.eta_fold_frac/cfg.nperh;
    ph.useVmagDt = cfg.use_velocity_magnitude_for_dt;
    ph.gravity = cfg.gravity;
    geo.permask = 0;
    geo.reflmask = 0;
    for (int a = 0; a < 3; ++a) {
      geo.bc_lo[a] = a < D ? cfg.bc_lo[a] : BC_NONE;
      geo.bc_hi[a] = a < D ? cfg.bc_hi[a] : BC_NONE;
      geo.xmin[a] = cfg.xmin[a];
      geo.xmax[a] = cfg.xmax[a];
      if (geo.bc_lo[a] == BC_PERIODIC) geo.permask |= 1 << a;
      if (geo.bc_lo[a] == BC_REFLECTING) geo.reflmask |= 1 << (2*a);
      if (geo.bc_hi[a] == BC_REFLECTING) geo.reflmask |= 1 << (2*a + 1);
    }
  }

  // Arrays that persist across steps and follow the reorder.
  std::vector<double**> carried(Fields& F) {
    std::vector<double**> v;
    for (int a = 0; a < D; ++a) { v.push_back(&F.x[a]); v.push_back(&F.v[a]); }
    for (double** p : {&F.m, &F.rho, &F.eps, &F.hinv, &F.P, &F.cs, &F.omega, &F.DepsDt, &F.DrhoDt, &F.DHDt, &F.maxQ, &F.pot, &F.potS})
      v.push_back(p);
    for (int a = 0; a < D; ++a) { v.push_back(&F.DxDt[a]); v.push_back(&F.DvDt[a]); v.push_back(&F.DvDtG[a]); }
    for (int a = 0; a < D; ++a) v.push_back(&F.xb[a]);
    v.push_back(&F.hb);
    return v;
  }

  void allocate() {
    std::vector<double**> all;
    Fields& F = f;
    for (int a = 0; a < D; ++a)
      for (double** p : {&F.x[a], &F.v[a], &F.xm[a], &F.vm[a], &F.v1[a], &F.DxDt[a], &F.DvDt[a], &F.DvDtG[a],
                         &F.gradRho[a], &F.m1[a], &F.xb[a]})
        all.push_back(p);
    for (double** p : {&F.m, &F.rho, &F.eps, &F.hinv, &F.P, &F.cs, &F.omega, &F.rhom, &F.epsm, &F.hinvm, &F.Pm,
                       &F.csm, &F.omegam, &F.prhom, &F.DepsDt, &F.DrhoDt, &F.DHDt, &F.maxQ, &F.Hideal, &F.m0,
                       &F.rhoSum, &F.norm, &F.effQ, &F.ns, &F.pot, &F.potS, &F.hb})
      all.push_back(p);
    const size_t total = (all.size() + 2*(size_t)D*D)*(size_t)ld;   // + DvDx, Mcorr
    double* slab = dalloc<double>(total);
    HIP_CHECK(hipMemsetAsync(slab, 0, total*sizeof(double), stream));
    size_t off = 0;
    for (double** p : all) { *p = slab + off; off += ld; }
    F.DvDx = slab + off; off += (size_t)D*D*ld;
    F.Mcorr = slab + off;
    for (int a = D; a < 3; ++a)
      F.x[a] = F.v[a] = F.xm[a] = F.vm[a] = F.v1[a] = F.DxDt[a] = F.DvDt[a] = F.DvDtG[a] = F.gradRho[a] =
          F.m1[a] = F.xb[a] = nullptr;
    alt = f;
    auto cl = carried(alt);
    const size_t altTotal = (cl.size() + (size_t)D*D)*(size_t)ld;
    double* slabAlt = dalloc<double>(altTotal);
    HIP_CHECK(hipMemsetAsync(slabAlt, 0, altTotal*sizeof(double), stream));
    off = 0;
    for (double** p : cl) { *p = slabAlt + off; off += ld; }
    alt.DvDx = slabAlt + off;
    F.DvDxQ = alt.DvDxQ = nullptr;
    if (limitedQ()) {
      const size_t q = (size_t)D*D*ld;
      F.DvDxQ = dalloc<double>(q);
      alt.DvDxQ = dalloc<double>(q);
      HIP_CHECK(hipMemsetAsync(F.DvDxQ, 0, q*sizeof(double), stream));
      HIP_CHECK(hipMemsetAsync(alt.DvDxQ, 0, q*sizeof(double), stream));
    }
    f.xh = dalloc<double>((size_t)4*ld);
    f.vmr = dalloc<double>((size_t)4*ld);
    f.xhm = dalloc<double>((size_t)4*ld);
    f.pkm = dalloc<double>((size_t)8*ld);
    f.pkv = dalloc<double>((size_t)8*ld);
    f.oid = dalloc<int64_t>(ld);
    oidAlt = dalloc<int64_t>(ld);
    f.run = dalloc<int32_t>(ld);
    runAlt = dalloc<int32_t>(ld);
    f.cnt = dalloc<int32_t>(ld);
    HIP_CHECK(hipMemsetAsync(f.cnt, 0, ld*sizeof(int32_t), stream));
    f.ld = ld;
    f.cap = 0;
    f.nbr = nullptr;
    f.ccap = 0;
    f.cand = nullptr;
    f.ccnt = dalloc<int32_t>(ld);
    HIP_CHECK(hipMemsetAsync(f.ccnt, 0, ld*sizeof(int32_t), stream));
    keys = dalloc<uint64_t>(ld);
    keysAlt = dalloc<uint64_t>(ld);
    vals = dalloc<uint32_t>(ld);
    valsAlt = dalloc<uint32_t>(ld);
    size_t tmp = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys, keysAlt, vals, valsAlt, (size_t)ld, 0, 64, stream));
    sortTmpBytes = tmp;
    sortTmp = dalloc<char>(tmp);
    dgrids = dalloc<RunGrid>(maxRuns);
    dsc = dalloc<Scalars>(maxRuns);
    dext = dalloc<unsigned long long>((size_t)maxRuns*kExt);
    dstate = dalloc<int>(maxRuns);
    dseg = dalloc<int2>(maxRuns);
    hseg = halloc<int2>(maxRuns);
    dorder = dalloc<int>(maxRuns);
    horder = halloc<int>(maxRuns);
    for (int r = 0; r < maxRuns; ++r) horder[r] = r;
    upload(dorder, horder, sizeof(int)*maxRuns);
    dinit = dalloc<int>(maxRuns);
    dOverflow = dalloc<int>(2);
    dOverflowG = dalloc<int>(1);
    dbuild = dalloc<int>(maxRuns);
    dsub = dalloc<int>(3*(size_t)maxRuns);
    skeys = dalloc<uint64_t>(ld);
    skeysAlt = dalloc<uint64_t>(ld);
    spos = dalloc<uint32_t>(ld);
    sposAlt = dalloc<uint32_t>(ld);
    ddelta = dalloc<int>(maxRuns);
    hgrids = halloc<RunGrid>(maxRuns);
    hsc = halloc<Scalars>(maxRuns);
    hext = halloc<unsigned long long>((size_t)maxRuns*kExt);
    hstate = halloc<int>(maxRuns);
    hinit = halloc<int>(maxRuns);
    hOverflow = halloc<int>(2);
    hbuild = halloc<int>(maxRuns);
    hsub = halloc<int>(3*(size_t)maxRuns);
    hdelta = halloc<int>(maxRuns);
    tilesAlloc = cap/kExtentTile + maxRuns + 1;
    dtiles = dalloc<int>(3*tilesAlloc);
    htiles = halloc<int>(3*tilesAlloc);
    upload(dsc, hsc, sizeof(Scalars)*maxRuns);
    ensureNbrCapacity(D == 1 ? 16 : (D == 2 ? 32 : 64));
    ensureCandCapacity(D == 1 ? 16 : (D == 2 ? 48 : 96));
    tcnt = dalloc<int>(ld);
    ensureHandoverCapacity(16);
  }

  // Room for c handed-over candidate entries per node (contents are rebuilt on every use).
  void ensureHandoverCapacity(int c) {
    if (c <= tcap) return;
    HIP_CHECK(hipStreamSynchronize(stream));
    dfree(tlist);
    tlist = dalloc<uint32_t>((size_t)c*ld);
    tcap = c;
  }

  void dfree(void* p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) {
      const size_t k = (size_t)(it - allocs.begin());
      deviceBytes -= (int64_t)allocBytes[k];
      allocBytes.erase(allocBytes.begin() + (std::ptrdiff_t)k);
      allocs.erase(it);
    }
    HIP_CHECK(hipFree(p));
  }

  // Grows the candidate lists to c entries per node, keeping every run's entries (the
  // blocked layout does not depend on the capacity).
  void ensureCandCapacity(int c) {
    if (c <= f.ccap) return;
    HIP_CHECK(hipStreamSynchronize(stream));
    uint32_t* p = dalloc<uint32_t>((size_t)c*ld);
    if (f.cand) HIP_CHECK(hipMemcpyAsync(p, f.cand, (size_t)f.ccap*ld*sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    dfree(f.cand);
    dfree(candAlt);
    f.cand = p;
    f.ccap = c;
    candAlt = nullptr;
    candAltCap = 0;
  }
  void ensureCandAlt() {
    if (candAlt && candAltCap == f.ccap) return;
    HIP_CHECK(hipStreamSynchronize(stream));
    dfree(candAlt);
    candAlt = dalloc<uint32_t>((size_t)f.ccap*ld);
    candAltCap = f.ccap;
    if (!ccntAlt) ccntAlt = dalloc<int32_t>(ld);
  }

  void ensureNbrCapacity(int c) {
    if (c <= f.cap) return;
    HIP_CHECK(hipStreamSynchronize(stream));
    for (void* old : {(void*)f.nbr, f.pdd}) dfree(old);
    f.nbr = nullptr;
    f.pdd = nullptr;
    f.nbr = dalloc<uint32_t>((size_t)c*ld);
    f.pdd = mixed ? (void*)dalloc<float>((size_t)c*ld) : (void*)dalloc<double>((size_t)c*ld);
    f.cap = c;
    npairsCache = -1;
  }

  template <typename F>
  void launchD(F&& fn) {
    switch (D) {
      case 1: fn(std::integral_constant<int, 1>{}); break;
      case 2: fn(std::integral_constant<int, 2>{}); break;
      default: fn(std::integral_constant<int, 3>{}); break;
    }
  }

  // ---------------------------------------------------------------- run admission / retirement
  int addRun(const double* x, const double* v, const double* m, const double* rho, const double* eps, const double* H,
             int64_t n, double tEnd, int slot = -1) {
    if (!x || !v || !m || !rho || !eps || !H) raise(SPX_ERR_ARG, "null state pointer");
    if (n < 1) raise(SPX_ERR_ARG, "a run needs at least one node");
    if (!(tEnd >= 0.0)) raise(SPX_ERR_ARG, "end time must be >= 0");
    // validate the whole input before touching the device
    std::vector<double> hbuf((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      const double* Hi = H + i*D*D;
      const double h = Hi[0];
      for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
          if (Hi[a*D + b] != (a == b ? h : 0.0))
            raise(SPX_ERR_UNSUPPORTED, "H must be isotropic (h^-1 I); ASPH smoothing tensors are not implemented");
      if (!(h > 0.0) || !std::isfinite(h)) raise(SPX_ERR_ARG, "H must be positive definite");
      if (!(m[i] > 0.0) || !(rho[i] > 0.0)) raise(SPX_ERR_ARG, "mass and mass density must be positive");
      hbuf[i] = h;
    }
    bool anyRemoved = false;
    for (auto& r : runs) anyRemoved |= r.state == RUN_REMOVED;
    if (slot < 0) {
      for (int s = 0; s < maxRuns && slot < 0; ++s) if (runs[s].state == RUN_FREE) slot = s;
      if (slot < 0 && anyRemoved) {
        compact();
        anyRemoved = false;
        for (int s = 0; s < maxRuns && slot < 0; ++s) if (runs[s].state == RUN_FREE) slot = s;
      }
      if (slot < 0) raise(SPX_ERR_CAPACITY, "no free run slot");
    } else if (slot >= maxRuns || runs[slot].state != RUN_FREE) {
      raise(SPX_ERR_STATE, "run slot in use");
    }
    if (nTotal + n > cap && anyRemoved) compact();
    if (nTotal + n > cap) raise(SPX_ERR_CAPACITY, "particle capacity of the batch exceeded");
    HIP_CHECK(hipStreamSynchronize(stream));
    const int64_t o = nTotal;
    std::vector<double> buf((size_t)n);
    auto up = [&](double* dst, const double* src) { upload(dst + o, src, n*sizeof(double)); };
    auto zero = [&](double* dst) { HIP_CHECK(hipMemsetAsync(dst + o, 0, n*sizeof(double), stream)); };
    for (int a = 0; a < D; ++a) {
      for (int64_t i = 0; i < n; ++i) buf[i] = x[i*D + a];
      up(f.x[a], buf.data());
      for (int64_t i = 0; i < n; ++i) buf[i] = v[i*D + a];
      up(f.v[a], buf.data());
      zero(f.DxDt[a]); zero(f.DvDt[a]); zero(f.DvDtG[a]);
    }
    up(f.hinv, hbuf.data());
    up(f.m, m);
    up(f.rho, rho);
    up(f.eps, eps);
    for (int64_t i = 0; i < n; ++i) buf[i] = 1.0;
    up(f.omega, buf.data());
    // derivatives start at zero (SPHBase::initializeProblemStartup)
    for (double* p : {f.DepsDt, f.DrhoDt, f.DHDt, f.maxQ, f.pot, f.potS, f.P, f.cs}) zero(p);
    for (int c = 0; c < D*D; ++c) HIP_CHECK(hipMemsetAsync(f.DvDx + (size_t)c*ld + o, 0, n*sizeof(double), stream));
    if (limitedQ())
      for (int c = 0; c < D*D; ++c) HIP_CHECK(hipMemsetAsync(f.DvDxQ + (size_t)c*ld + o, 0, n*sizeof(double), stream));
    std::vector<int64_t> ids(n);
    std::vector<int32_t> rr(n, slot);
    for (int64_t i = 0; i < n; ++i) ids[i] = ((int64_t)slot << 40) | i;
    upload(f.oid + o, ids.data(), n*sizeof(int64_t));
    upload(f.run + o, rr.data(), n*sizeof(int32_t));
    RunInfo info;
    info.state = RUN_ACTIVE;
    info.n = n;
    info.offset = o;
    info.needInit = true;
    runs[slot] = info;
    Scalars s;
    std::memset(&s, 0, sizeof(Scalars));
    s.lastDt = cfg.last_dt;
    s.tEnd = tEnd;
    s.active = 1;
    hsc[slot] = s;
    upload(dsc + slot, &hsc[slot], sizeof(Scalars));
    nTotal += n;
    if (!retire) nActive = nTotal;     // a single run is always active, in upload order until the first sort
    connValid = false;
    orderDirty = true;
    if (retire && tEnd <= 0.0) markDone(slot);
    return slot;
  }

  void markDone(int slot, bool reorder = true) {
    runs[slot].state = RUN_DONE;
    runs[slot].candValid = false;
    if (reorder) orderDirty = true;
    hsc[slot].active = 0;
    upload(&dsc[slot].active, &hsc[slot].active, sizeof(int));
  }

  // Active runs in the order of their segments.
  std::vector<int> runsByOffset() const {
    std::vector<int> a;
    for (int r = 0; r < maxRuns; ++r)
      if (runs[r].state == RUN_ACTIVE) a.push_back(r);
    std::sort(a.begin(), a.end(), [&](int x, int y) { return runs[x].offset < runs[y].offset; });
    return a;
  }

  // Runs finished by a step: when their segments are the tail of the active range (the
  // estimated-remaining-steps order puts the soonest to finish last), the active range just
  // shrinks and nothing moves; otherwise the next connectivity pass reorders the batch.
  void retireFinished() {
    int64_t na = 0;
    for (int r = 0; r < maxRuns; ++r)
      if (runs[r].state == RUN_ACTIVE) na += runs[r].n;
    bool tail = !orderDirty;
    for (int r = 0; r < maxRuns && tail; ++r)
      if (runs[r].state == RUN_ACTIVE && runs[r].offset + runs[r].n > na) tail = false;
    if (tail) {
      nActive = na;
      npairsCache = -1;
      ++inPlaceRetirements;
    } else {
      orderDirty = true;
    }
  }

  void removeRun(int slot) {
    checkSlot(slot);
    runs[slot].state = RUN_REMOVED;
    orderDirty = true;
    hsc[slot].active = 0;
    upload(&dsc[slot].active, &hsc[slot].active, sizeof(int));
  }

  void checkSlot(int slot) const {
    if (slot < 0 || slot >= maxRuns || runs[slot].state == RUN_FREE || runs[slot].state == RUN_REMOVED)
      raise(SPX_ERR_ARG, "no such run");
  }

  // Squeeze removed runs out of the arrays (no cell grids needed).
  void compact() {
    for (int r = 0; r < maxRuns; ++r) {
      RunGrid& g = hgrids[r];
      for (int a = 0; a < 3; ++a) { g.origin[a] = 0.0; g.inv_cell[a] = 0.0; g.nc[a] = 1; }
      g.cellBase = 0;
    }
    HIP_CHECK(hipMemcpyAsync(dgrids, hgrids, sizeof(RunGrid)*maxRuns, hipMemcpyHostToDevice, stream));
    for (int r = 0; r < maxRuns; ++r) hbuild[r] = 0;
    HIP_CHECK(hipMemcpyAsync(dbuild, hbuild, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
    sortAndGather(0);
    orderDirty = false;
    connValid = false;
  }

  // ---------------------------------------------------------------- connectivity
  // Cell grids of the runs that build candidate lists (hbuild), from their current extents;
  // returns the Morton bits per axis needed.
  int setupGrids() {
    hipLaunchKernelGGL(k_reset_extent, dim3(nblocks((int64_t)maxRuns*kExt)), dim3(kBlock), 0, stream, dext, maxRuns);
    int64_t nt = 0;
    for (int r = 0; r < maxRuns; ++r) {
      if (runs[r].state != RUN_ACTIVE || !hbuild[r]) continue;
      const int64_t b = runs[r].offset, e = b + runs[r].n;
      for (int64_t t = b; t < e; t += kExtentTile) {
        htiles[3*nt] = r;
        htiles[3*nt + 1] = (int)t;
        htiles[3*nt + 2] = (int)std::min(e, t + kExtentTile);
        ++nt;
      }
    }
    if (nt > 0) {
      HIP_CHECK(hipMemcpyAsync(dtiles, htiles, 3*nt*sizeof(int), hipMemcpyHostToDevice, stream));
      launchD([&](auto d) {
        constexpr int DD = decltype(d)::value;
        hipLaunchKernelGGL(k_extent<DD>, dim3((unsigned)nt), dim3(kBlock), 0, stream, f, (const int*)dtiles, dext);
      });
      HIP_CHECK(hipGetLastError());
    }
    download(hext, dext, sizeof(unsigned long long)*maxRuns*kExt);
    const int64_t maxAxis = D == 3 ? (1 << 15) : (D == 2 ? (1 << 20) : (1 << 26));
    int64_t base = 0;
    int bits = 0;
    for (int r = 0; r < maxRuns; ++r) {
      RunGrid& gr = hgrids[r];
      for (int a = 0; a < 3; ++a) { gr.origin[a] = 0.0; gr.inv_cell[a] = 0.0; gr.nc[a] = 1; }
      gr.cellBase = (int)base;
      gr.hinvMin = 0.0;
      gr.ext2c = 0.0;
      gr.support = 0.0;
      gr.levels = 1;
      if (runs[r].state != RUN_ACTIVE || !hbuild[r]) continue;
      const double hinvMin = ord_value(hext[r*kExt]), hinvMax = ord_value(hext[r*kExt + 7]);
      if (!(hinvMin > 0.0) || !std::isfinite(hinvMin)) raise(SPX_ERR_NUMERIC, "non-positive or non-finite H on the device");
      gr.hinvMin = hinvMin;
      const double extc = cfg.neighbor_extent*(1.0 + runs[r].skin);
      gr.ext2c = extc*extc;
      const double support = extc/hinvMin*(1.0 + 1.0e-9);
      gr.support = support;
      // where h varies by 2x or more: with open boundaries, cell levels down to the smallest
      // support (up to kCellLevels, while the finest grid stays within ~8 cells per node); with
      // periodic or reflecting axes, cells shrink toward the smaller supports (by up to 4x, same
      // bound) and each node scans the cells its own support reaches
      bool open = true;
      for (int a = 0; a < D; ++a) open &= geo.bc_lo[a] != BC_PERIODIC && geo.bc_lo[a] != BC_REFLECTING &&
                                          geo.bc_hi[a] != BC_PERIODIC && geo.bc_hi[a] != BC_REFLECTING;
      int F = 1, lev = 1;
      if (hinvMax >= 2.0*hinvMin) {
        if (open) lev = std::min(kCellLevels, 1 + (int)std::floor(std::log2(hinvMax/hinvMin)));
        else F = (int)std::min(4.0, std::floor(hinvMax/hinvMin));
      }
      int64_t total = 1;
      for (;;) {
        total = 1;
        int runBitsAxis = 0;
        for (int a = 0; a < D; ++a) {
          const double lo = ord_value(hext[r*kExt + 1 + 2*a]), hi = ord_value(hext[r*kExt + 2 + 2*a]);
          if (!std::isfinite(lo) || !std::isfinite(hi)) raise(SPX_ERR_NUMERIC, "non-finite position on the device");
          const int blo = geo.bc_lo[a], bhi = geo.bc_hi[a];
          const double L = cfg.xmax[a] - cfg.xmin[a];
          const double target = lev > 1 ? support/(double)(4 << (lev - 1)) : support/F;
          int64_t nc;
          double cell, origin;
          if (blo == BC_PERIODIC) {
            if (std::floor(L/support) < 3 && runs[r].skin > 0.0) {
              // too small a box for the widened radius: fall back to the plain one
              runs[r].skin = 0.0;
              return setupGrids();
            }
            if (std::floor(L/support) < 3) raise(SPX_ERR_UNSUPPORTED, "periodic box must span at least 3 support radii on every axis");
            nc = std::min<int64_t>((int64_t)std::floor(L/target), maxAxis);
            cell = L/nc;
            origin = cfg.xmin[a];
          } else if (blo == BC_REFLECTING && bhi == BC_REFLECTING) {
            nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor(L/target), maxAxis));
            cell = L/nc;
            origin = cfg.xmin[a];
          } else if (blo == BC_REFLECTING) {
            const double top = std::max(hi, cfg.xmin[a]);
            nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor((top - cfg.xmin[a])/target) + 1, maxAxis));
            cell = std::max(target, (top - cfg.xmin[a])/nc*(1.0 + 1e-12));
            origin = cfg.xmin[a];
          } else if (bhi == BC_REFLECTING) {
            const double bot = std::min(lo, cfg.xmax[a]);
            nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor((cfg.xmax[a] - bot)/target) + 1, maxAxis));
            cell = std::max(target, (cfg.xmax[a] - bot)/nc*(1.0 + 1e-12));
            origin = cfg.xmax[a] - nc*cell;
          } else {
            const double span = std::max(0.0, hi - lo);
            nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor(span/target) + 1, maxAxis));
            cell = std::max(target, span/nc*(1.0 + 1e-12));
            origin = lo;
          }
          gr.nc[a] = (int)nc;
          gr.inv_cell[a] = 1.0/cell;
          gr.origin[a] = origin;
          total *= nc;
          int b = 0;
          while (((int64_t)1 << b) < nc) ++b;
          runBitsAxis = std::max(runBitsAxis, b);
        }
        if (lev > 1) {
          // every level's table: level-l cells are 2^(lev-1-l) finest cells per axis
          int64_t all = 0;
          for (int l = 0; l < lev; ++l) {
            int64_t t = 1;
            for (int a = 0; a < D; ++a) t *= ((int64_t)(gr.nc[a] - 1) >> (lev - 1 - l)) + 1;
            gr.levelBase[l] = (int)(base + all);
            all += t;
          }
          total = all;
        }
        if (F > 1 && total > 8*runs[r].n + 4096) { --F; continue; }
        if (lev > 1 && total > 8*runs[r].n + 4096) { --lev; continue; }
        bits = std::max(bits, runBitsAxis);
        break;
      }
      gr.levels = lev;
      if (lev <= 1) gr.levelBase[0] = (int)base;
      base += total;
      if (base > ((int64_t)1 << 28)) raise(SPX_ERR_CAPACITY, "cell grids too large (support radius tiny relative to the domain)");
    }
    if (base > cellsAlloc) {
      HIP_CHECK(hipStreamSynchronize(stream));
      const int64_t c = base + base/4 + 1024;
      for (void* p : {(void*)cellStart, (void*)cellEnd}) dfree(p);
      cellStart = dalloc<int>(c);
      cellEnd = dalloc<int>(c);
      cellsAlloc = c;
    }
    nCells = base;
    HIP_CHECK(hipMemcpyAsync(dgrids, hgrids, sizeof(RunGrid)*maxRuns, hipMemcpyHostToDevice, stream));
    if (base > 0) {
      HIP_CHECK(hipMemsetAsync(cellStart, 0, base*sizeof(int), stream));
      HIP_CHECK(hipMemsetAsync(cellEnd, 0, base*sizeof(int), stream));
    }
    return bits;
  }

  // Sort by [rank | run | Morton] and gather the carried arrays; drops removed runs.  Runs
  // that keep their candidate lists (hbuild 0) keep their order; their lists follow them.
  void sortAndGather(int bitsPerAxis) {
    mortonBits = D*bitsPerAxis;
    const int endBit = mortonBits + runBits + 2;
    if (endBit > 64) raise(SPX_ERR_CAPACITY, "sort key exceeds 64 bits");
    for (int r = 0; r < maxRuns; ++r) hstate[r] = runs[r].state;
    HIP_CHECK(hipMemcpyAsync(dstate, hstate, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
    // active runs in decreasing order of their estimated remaining steps: the run expected to
    // finish first sits last, so it can retire without moving anything (retireFinished)
    std::vector<int> act;
    std::vector<double> est(maxRuns, 0.0);
    for (int r = 0; r < maxRuns; ++r) {
      if (runs[r].state != RUN_ACTIVE) continue;
      act.push_back(r);
      const double dt = hsc[r].dt > 0.0 ? hsc[r].dt : hsc[r].lastDt;
      est[r] = (runs[r].needInit || !(dt > 0.0)) ? HUGE_VAL : (hsc[r].tEnd - hsc[r].t)/dt;
    }
    std::stable_sort(act.begin(), act.end(), [&](int a, int b) { return est[a] > est[b]; });
    ++fullReorders;
    for (size_t k = 0; k < act.size(); ++k) horder[act[k]] = (int)k;
    HIP_CHECK(hipMemcpyAsync(dorder, horder, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
    const int nb = nblocks(nTotal);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_keys<DD>, dim3(nb), dim3(kBlock), 0, stream, f, geo, (const RunGrid*)dgrids, (const int*)dstate,
                         (const int*)dorder, (const int*)dbuild, (int)nTotal, wrapPending ? 1 : 0, 1, mortonBits, runBits, keys, vals);
    });
    HIP_CHECK(hipGetLastError());
    wrapPending = false;
    size_t tmp = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys, keysAlt, vals, valsAlt, (size_t)nTotal, 0, endBit, stream));
    if (tmp > sortTmpBytes) {
      HIP_CHECK(hipStreamSynchronize(stream));
      dfree(sortTmp);
      sortTmp = dalloc<char>(tmp);
      sortTmpBytes = tmp;
    }
    tmp = sortTmpBytes;
    HIP_CHECK(rocprim::radix_sort_pairs(sortTmp, tmp, keys, keysAlt, vals, valsAlt, (size_t)nTotal, 0, endBit, stream));
    auto src = carried(f);
    auto dst = carried(alt);
    GatherSet gs{};
    gs.count = 0;
    for (size_t k = 0; k < src.size(); ++k) { gs.src[gs.count] = *src[k]; gs.dst[gs.count] = *dst[k]; ++gs.count; }
    for (int c = 0; c < D*D; ++c) { gs.src[gs.count] = f.DvDx + (size_t)c*ld; gs.dst[gs.count] = alt.DvDx + (size_t)c*ld; ++gs.count; }
    if (limitedQ())
      for (int c = 0; c < D*D; ++c) { gs.src[gs.count] = f.DvDxQ + (size_t)c*ld; gs.dst[gs.count] = alt.DvDxQ + (size_t)c*ld; ++gs.count; }
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(kBlock), 0, stream, gs, valsAlt, (int)nTotal);
    hipLaunchKernelGGL(k_gather_ids, dim3(nb), dim3(kBlock), 0, stream, f.oid, oidAlt, f.run, runAlt, valsAlt, (int)nTotal);
    HIP_CHECK(hipGetLastError());
    for (size_t k = 0; k < src.size(); ++k) std::swap(*src[k], *dst[k]);
    std::swap(f.DvDx, alt.DvDx);
    std::swap(f.DvDxQ, alt.DvDxQ);
    std::swap(f.oid, oidAlt);
    std::swap(f.run, runAlt);
    // segments in key order: active runs by place, then finished runs by slot
    std::vector<int64_t> oldOffset(maxRuns);
    for (int r = 0; r < maxRuns; ++r) oldOffset[r] = runs[r].offset;
    nActive = 0;
    int64_t live = 0;
    for (int r : act) {
      RunInfo& ri = runs[r];
      ri.offset = live;
      live += ri.n;
      nActive += ri.n;
    }
    for (int r = 0; r < maxRuns; ++r) {
      RunInfo& ri = runs[r];
      if (ri.state == RUN_DONE) { ri.offset = live; live += ri.n; }
      if (ri.state == RUN_REMOVED) ri = RunInfo();
      ri.stepOut = ri.fullOut = false;    // the per-step outputs were not carried through the reorder
    }
    nTotal = live;
    npairsCache = -1;
    // candidate lists of the runs that keep them move with their segments
    bool moved = false;
    for (int r = 0; r < maxRuns; ++r) {
      hdelta[r] = kNoRemap;
      if (runs[r].state == RUN_ACTIVE && !hbuild[r] && runs[r].candValid) {
        hdelta[r] = (int)(runs[r].offset - oldOffset[r]);
        moved |= hdelta[r] != 0;
      }
    }
    if (moved && nActive > 0) {
      ensureCandAlt();
      HIP_CHECK(hipMemcpyAsync(ddelta, hdelta, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
      hipLaunchKernelGGL(k_remap_candidates, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, (const uint32_t*)f.cand,
                         (const int32_t*)f.ccnt, candAlt, ccntAlt, (const int32_t*)f.run, (const int*)ddelta, (size_t)ld,
                         (int)nActive);
      HIP_CHECK(hipGetLastError());
      std::swap(f.cand, candAlt);
      std::swap(f.ccnt, ccntAlt);
    }
  }

  // Re-sort only the runs that rebuild (hbuild), each within its own segment, and gather
  // only their particles; the other runs, their candidate lists and every offset stay where
  // they are.  For steps on which no run was added, finished or removed.  The rebuilding
  // runs' particles are compacted with keys [segment | cell Morton code], sorted with one
  // device-wide radix sort, and the k-th sorted particle goes to the k-th of their slots.
  void sortRebuildingRuns(int bitsPerAxis) {
    mortonBits = D*bitsPerAxis;
    if (mortonBits + runBits + 2 > 64) raise(SPX_ERR_CAPACITY, "sort key exceeds 64 bits");
    for (int r = 0; r < maxRuns; ++r) hstate[r] = runs[r].state;
    HIP_CHECK(hipMemcpyAsync(dstate, hstate, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_keys<DD>, dim3(nblocks(nTotal)), dim3(kBlock), 0, stream, f, geo, (const RunGrid*)dgrids,
                         (const int*)dstate, (const int*)dorder, (const int*)dbuild, (int)nTotal, wrapPending ? 1 : 0, 1,
                         mortonBits, runBits, keys, vals);
    });
    HIP_CHECK(hipGetLastError());
    wrapPending = false;
    int nseg = 0;
    int64_t nsub = 0;
    for (int r = 0; r < maxRuns; ++r) {
      hsub[3*r] = (int)runs[r].offset;
      hsub[3*r + 1] = (int)nsub;
      hsub[3*r + 2] = nseg;
      if (!hbuild[r]) continue;
      nsub += runs[r].n;
      ++nseg;
    }
    if (nseg == 0) return;
    int segBits = 0;
    while ((1 << segBits) < nseg) ++segBits;
    if (mortonBits + segBits > 64) raise(SPX_ERR_CAPACITY, "sort key exceeds 64 bits");
    HIP_CHECK(hipMemcpyAsync(dsub, hsub, sizeof(int)*3*maxRuns, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_subset_keys, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, (const uint64_t*)keys,
                       (const int32_t*)f.run, (const int*)dbuild, (const int*)dsub, mortonBits, (int)nActive, skeys, spos);
    HIP_CHECK(hipGetLastError());
    const int endBit = std::max(1, mortonBits + segBits);
    size_t tmp = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, skeys, skeysAlt, spos, sposAlt, (size_t)nsub, 0, endBit, stream));
    if (tmp > sortTmpBytes) {
      HIP_CHECK(hipStreamSynchronize(stream));
      dfree(sortTmp);
      sortTmp = dalloc<char>(tmp);
      sortTmpBytes = tmp;
    }
    tmp = sortTmpBytes;
    HIP_CHECK(rocprim::radix_sort_pairs(sortTmp, tmp, skeys, skeysAlt, spos, sposAlt, (size_t)nsub, 0, endBit, stream));
    // gather into the spare arrays at the compacted slots, then copy those slots back
    auto src = carried(f);
    auto dst = carried(alt);
    GatherSet there{}, back{};
    for (size_t k = 0; k < src.size(); ++k) {
      there.src[there.count] = *src[k]; there.dst[there.count++] = *dst[k];
      back.src[back.count] = *dst[k]; back.dst[back.count++] = *src[k];
    }
    for (int c = 0; c < D*D; ++c) {
      there.src[there.count] = f.DvDx + (size_t)c*ld; there.dst[there.count++] = alt.DvDx + (size_t)c*ld;
      back.src[back.count] = alt.DvDx + (size_t)c*ld; back.dst[back.count++] = f.DvDx + (size_t)c*ld;
    }
    if (limitedQ())
      for (int c = 0; c < D*D; ++c) {
        there.src[there.count] = f.DvDxQ + (size_t)c*ld; there.dst[there.count++] = alt.DvDxQ + (size_t)c*ld;
        back.src[back.count] = alt.DvDxQ + (size_t)c*ld; back.dst[back.count++] = f.DvDxQ + (size_t)c*ld;
      }
    const int nb = nblocks(nsub);
    hipLaunchKernelGGL(k_gather_list, dim3(nb), dim3(kBlock), 0, stream, there, (const uint32_t*)spos,
                       (const uint32_t*)sposAlt, (const int64_t*)f.oid, oidAlt, (const uint64_t*)skeysAlt, keysAlt,
                       (int)nsub);
    hipLaunchKernelGGL(k_gather_list, dim3(nb), dim3(kBlock), 0, stream, back, (const uint32_t*)spos,
                       (const uint32_t*)nullptr, (const int64_t*)oidAlt, f.oid, (const uint64_t*)nullptr,
                       (uint64_t*)nullptr, (int)nsub);
    HIP_CHECK(hipGetLastError());
    for (int r = 0; r < maxRuns; ++r)
      if (hbuild[r]) runs[r].stepOut = runs[r].fullOut = false;   // per-step outputs were not carried
    npairsCache = -1;
  }

  // The candidate list of run r no longer provably holds every neighbour pair (or the
  // measurement is not finite).
  bool skinTripped(int r) const {
    const double d = skin_value(hsc[r].skinD), g = skin_value(hsc[r].skinG);
    return !(g + 2.0*d <= (1.0 + runs[r].skin)*(1.0 - 1.0e-9));
  }

  // ConnectivityMap::computeConnectivity for every active run: runs whose candidate list is
  // missing or stale (or all, with force) re-sort and rebuild it; then every active node's
  // neighbour list is filtered from its candidates.  withDensityDt: follow with the step's
  // sum density and dt vote.
  void buildConnectivity(bool withDensityDt, bool force = false) {
    connectivityPrep(force);
    if (nActive == 0) return;
    connectivityFilter(withDensityDt);
  }
  // The part of the connectivity pass that decides and does candidate-list builds: boundary
  // enforcement, the reorders, the candidate search (with its own capacity checks, which may
  // resize buffers).  Leaves connValid false.
  void connectivityPrep(bool force) {
    if (nTotal == 0) raise(SPX_ERR_STATE, "no state uploaded");
    ensureOmega();                   // a CheapSynchronousRK2 step drops it first (it never reads it)
    npairsCache = -1;                // every pass re-derives the neighbour lists
    bool any = false;
    for (int r = 0; r < maxRuns; ++r) {
      RunInfo& ri = runs[r];
      hbuild[r] = 0;
      if (ri.state != RUN_ACTIVE) continue;
      const bool tripped = ri.candValid && skinTripped(r);
      if (!(force || !ri.candValid || tripped)) continue;
      if (tripped && ri.skin > 0.0) {
        // the skin was used up in (cycle - builtCycle) steps: re-choose it from that rate
        const long long k = std::max<long long>(1, hsc[r].cycle - ri.builtCycle);
        const double used = skin_value(hsc[r].skinG) + 2.0*skin_value(hsc[r].skinD) - 1.0;
        if (std::isfinite(used)) ri.skin = choose_skin(std::max(used, 0.0)/(double)k, D);
      }
      hbuild[r] = 1;
      any = true;
    }
    if (any || orderDirty) {
      HIP_CHECK(hipMemcpyAsync(dbuild, hbuild, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
      const int bits = any ? setupGrids() : 0;
      // a change of the batch's composition moves segments: full reorder; otherwise only the
      // rebuilding runs are re-sorted, in place
      if (orderDirty) sortAndGather(bits);
      else sortRebuildingRuns(bits);
      orderDirty = false;
      mark("reorder");
    } else {
      applyPendingBoundaries();
    }
    connValid = false;
    if (nActive == 0) return;
    const int nb = nblocks(nActive);
    if (any) {
      launchD([&](auto d) {
        constexpr int DD = decltype(d)::value;
        hipLaunchKernelGGL(k_cell_bounds<DD>, dim3(nb), dim3(kBlock), 0, stream, (const uint64_t*)keysAlt,
                           (const int32_t*)f.run, (int)nActive, (const RunGrid*)dgrids, (const int*)dbuild, mortonBits,
                           cellStart, cellEnd);
      });
      HIP_CHECK(hipGetLastError());
      // the rebuilding runs' segments, in position order, with their compacted starts
      int nseg = 0;
      int64_t nBuild = 0;
      for (int r : runsByOffset()) {
        if (!hbuild[r]) continue;
        hseg[nseg++] = make_int2((int)runs[r].offset, (int)nBuild);
        nBuild += runs[r].n;
      }
      HIP_CHECK(hipMemcpyAsync(dseg, hseg, sizeof(int2)*nseg, hipMemcpyHostToDevice, stream));
      bool ok = false;
      for (int attempt = 0; attempt < 4 && !ok; ++attempt) {
        HIP_CHECK(hipMemsetAsync(dOverflow, 0, 2*sizeof(int), stream));
        HIP_CHECK(hipMemsetAsync(tcnt, 0, nActive*sizeof(int), stream));
        launchD([&](auto d) {
          constexpr int DD = decltype(d)::value;
          hipLaunchKernelGGL(k_candidates<DD>, dim3(xcd_grid(nBuild, kBlock)), dim3(kBlock), 0, stream, f, geo,
                             (const RunGrid*)dgrids, (const int2*)dseg, nseg, (int)nBuild, (const int*)cellStart,
                             (const int*)cellEnd, tlist, tcnt, tcap, dOverflow);
        });
        hipLaunchKernelGGL(k_candidates_merge, dim3(nblocks(nBuild)), dim3(kBlock), 0, stream, f, (const int2*)dseg,
                           nseg, (int)nBuild, (const uint32_t*)tlist, (const int*)tcnt, tcap, dOverflow);
        HIP_CHECK(hipGetLastError());
        download(hOverflow, dOverflow, 2*sizeof(int));
        if (hOverflow[0] == 0 && hOverflow[1] == 0) { ok = true; break; }
        if (hOverflow[0] > 0) {
          const int need = hOverflow[0];
          if (need > 8192) raise(SPX_ERR_CAPACITY, "more than 8192 neighbour candidates for one node");
          ensureCandCapacity(((need + need/4) + 7)/8*8);
        }
        if (hOverflow[1] > 0) {
          const int need = hOverflow[1];
          if (need > 8192) raise(SPX_ERR_CAPACITY, "more than 8192 handed-over candidates for one node");
          ensureHandoverCapacity(((need + need/4) + 7)/8*8);
        }
      }
      if (!ok) raise(SPX_ERR_CAPACITY, "candidate list capacity not reached");
      mark("candidates");
      for (int r = 0; r < maxRuns; ++r) {
        if (!hbuild[r]) continue;
        RunInfo& ri = runs[r];
        ri.candValid = true;
        ri.builtCycle = hsc[r].cycle;
        ++ri.builds;
        ++candBuilds;
        hsc[r].skinD = hsc[r].skinG = 0ull;
      }
    }
  }
  // The step's exact neighbour lists from the candidate lists, with the connectivity records
  // and, withDensityDt, the sum density and the dt vote; grows the lists on overflow.
  void connectivityFilter(bool withDensityDt) {
    pack();
    bool ok = false;
    for (int attempt = 0; attempt < 3 && !ok; ++attempt) {
      HIP_CHECK(hipMemsetAsync(dOverflow, 0, sizeof(int), stream));
      const unsigned nbx = xcd_grid(nActive, kBlock);
      if (withDensityDt) {
        // the step's sum density and dt vote ride on the filter
        hipLaunchKernelGGL(k_reset_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns);
        tuneMark(0);
        launchFilter(nbx, dOverflow);
        tuneMark(1);
      } else {
        launchD([&](auto d) {
          constexpr int DD = decltype(d)::value;
          hipLaunchKernelGGL((k_neighbor_filter<DD, kGroup>), dim3(nbx), dim3(kBlock), 0, stream, f, geo, ph,
                             (int)nActive, dOverflow);
        });
      }
      HIP_CHECK(hipGetLastError());
      download(hOverflow, dOverflow, sizeof(int));
      if (*hOverflow == 0) { ok = true; break; }
      const int need = *hOverflow;
      if (need > 4096) raise(SPX_ERR_CAPACITY, "more than 4096 neighbours for one node");
      ensureNbrCapacity(((need + need/4) + 7)/8*8);
    }
    if (!ok) raise(SPX_ERR_CAPACITY, "neighbour list capacity not reached");
    connValid = true;
  }

  // Boundary enforcement without a neighbour rebuild (before a state download).
  void applyPendingBoundaries() {
    if (!wrapPending) return;
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_keys<DD>, dim3(nblocks(nTotal)), dim3(kBlock), 0, stream, f, geo, (const RunGrid*)nullptr,
                         (const int*)nullptr, (const int*)nullptr, (const int*)nullptr, (int)nTotal, 1, 0, 0, 0, keys,
                         vals);
    });
    HIP_CHECK(hipGetLastError());
    wrapPending = false;
  }

  // ---------------------------------------------------------------- passes (over the active runs)
  template <typename K64, typename K32, typename... Args>
  void launchPrec(K64 k64, K32 k32, int nb, Args... args) {
    if (mixed) hipLaunchKernelGGL(k32, dim3(nb), dim3(kBlock), 0, stream, args...);
    else hipLaunchKernelGGL(k64, dim3(nb), dim3(kBlock), 0, stream, args...);
    HIP_CHECK(hipGetLastError());
  }
  // As launchPrec for pair passes that stage the kernel table in LDS (their last argument
  // is the staging flag).
  template <typename K64, typename K32, typename... Args>
  void launchPrecTable(K64 k64, K32 k32, int nb, Args... args) {
    const size_t bytes = tableBytes();
    const int lds = bytes <= kLdsTableMax ? 1 : 0;
    const size_t shm = lds ? bytes : 0;
    if (mixed) hipLaunchKernelGGL(k32, dim3(nb), dim3(kBlock), shm, stream, args..., lds);
    else hipLaunchKernelGGL(k64, dim3(nb), dim3(kBlock), shm, stream, args..., lds);
    HIP_CHECK(hipGetLastError());
  }
  size_t tableBytes() const { return (size_t)6*(dtab.n1 + 1)*(mixed ? sizeof(float) : sizeof(double)); }

  void eos() {
    if (nActive == 0) return;
    hipLaunchKernelGGL(k_eos, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, f, ph, (int)nActive);
    HIP_CHECK(hipGetLastError());
  }
  void ensureOmega() {
    if (!omegaStale) return;
    omegaStale = false;
    omegaPass(false);
  }
  void omegaPass(bool mid, const int* runMask = nullptr) {
    if (nActive == 0) return;
    if (!cfg.gradh_correction) {
      if (mid) packMid(true);
      return;
    }
    const int nb = (int)xcd_grid(nActive, kBlock);
    // midpoint: the record k_midstep wrote; state: the connectivity record (k_pack), or the
    // end-of-step position k_full_update left in it
    if (mid) tuneMark(2);
    if (groupSize(1) == 8) launchOmega<8>(mid, nb, runMask);
    else launchOmega<4>(mid, nb, runMask);
    if (mid) tuneMark(3);
  }
  template <int G>
  void launchOmega(bool mid, int nb, const int* runMask) {
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (mid)
        launchPrecTable(k_omega<DD, double, true, G>, k_omega<DD, float, true, G>, nb, f, geo, dtab, (int)nActive,
                        (const double*)f.xhm, f.omegam, runMask);
      else
        launchPrecTable(k_omega<DD, double, false, G>, k_omega<DD, float, false, G>, nb, f, geo, dtab, (int)nActive,
                        (const double*)f.xh, f.omega, runMask);
    });
  }
  // The step's filtering sum-density + dt pass (see buildConnectivity).
  void launchFilter(unsigned nbx, int* overflow) {
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (groupSize(0) == 8)
        launchPrecTable(k_density_dt<DD, true, true, double, 8, true>, k_density_dt<DD, true, true, float, 8, true>,
                        (int)nbx, f, geo, ph, dtab, (int)nActive, dsc, overflow);
      else
        launchPrecTable(k_density_dt<DD, true, true, double, 4, true>, k_density_dt<DD, true, true, float, 4, true>,
                        (int)nbx, f, geo, ph, dtab, (int)nActive, dsc, overflow);
    });
  }

  // ---------------------------------------------------------------- neighbour-group sizes
  int groupSize(int pass) const {
    if (tuner.phase >= 0) return (tuner.phase & 1) ? 8 : 4;
    return tuner.size[pass];
  }
  // Event k of the measurement step (0/1 filter, 2/3 and 4/5 the two grad-h passes, 6/7 the
  // energy update).
  void tuneMark(int k) {
    if (tuner.phase < 0) return;
    if (!tuneEv[k]) HIP_CHECK(hipEventCreate(&tuneEv[k]));
    HIP_CHECK(hipEventRecord(tuneEv[k], stream));
  }
  void maybeStartTuning() {
    if (tuner.fixed || tuner.phase >= 0 || tuner.warm > 0 || nActive == 0) return;
    if (tuner.tunedFor && nActive <= 2*tuner.tunedFor && 2*nActive >= tuner.tunedFor) return;
    for (auto& b : tuner.best) b[0] = b[1] = 3.0e38f;
    tuner.warm = GroupTuner::kWarm;
    tuner.tunedFor = nActive;
  }
  // End of any step: count down the warm-up; the step after it is measured.
  void tunerStepDone() {
    if (tuner.warm > 0 && --tuner.warm == 0) tuner.phase = 0;
  }
  // After a measured step (the run scalars' download has synchronised the stream).
  void finishTuningStep() {
    if (tuner.phase < 0) return;
    const int v = tuner.phase & 1;
    float ms[3] = {0.f, 0.f, 0.f};
    auto el = [&](int a, int b) {
      float t = 0.f;
      if (tuneEv[a] && tuneEv[b]) HIP_CHECK(hipEventElapsedTime(&t, tuneEv[a], tuneEv[b]));
      return t;
    };
    ms[0] = el(0, 1);
    ms[1] = el(2, 3);
    ms[2] = el(6, 7);
    for (int p = 0; p < 3; ++p) tuner.best[p][v] = std::min(tuner.best[p][v], ms[p]);
    if (++tuner.phase < GroupTuner::kSteps) return;
    for (int p = 0; p < 3; ++p) {
      const int d = GroupTuner::kDefault[p] == 8 ? 1 : 0;
      tuner.size[p] = tuner.best[p][1 - d] < 0.97f*tuner.best[p][d] ? (d ? 4 : 8) : GroupTuner::kDefault[p];
    }
    tuner.phase = -1;
    tuner.tunedFor = nActive;
    ++tuner.tunings;
  }
  // Counters: graph captures, graph replays, eager steps, group-size measurements, and the
  // current sizes of the three passes.
  void stepStats(int64_t* out) const {
    out[0] = graphCaptures;
    out[1] = graphReplays;
    out[2] = eagerSteps;
    out[3] = tuner.tunings;
    for (int p = 0; p < 3; ++p) out[4 + p] = tuner.size[p];
    out[7] = fullReorders;
    out[8] = inPlaceRetirements;
  }
  // Fixed sizes (each 4 or 8), or sizes <= 0: measure.
  void setGroupSizes(const int* sizes) {
    tuner.phase = -1;
    tuner.warm = 0;
    if (!sizes || sizes[0] <= 0) {
      tuner.fixed = false;
      tuner.tunedFor = 0;
      return;
    }
    for (int p = 0; p < 3; ++p)
      if (sizes[p] != 4 && sizes[p] != 8) raise(SPX_ERR_ARG, "neighbour-group sizes must be 4 or 8");
    for (int p = 0; p < 3; ++p) tuner.size[p] = sizes[p];
    tuner.fixed = true;
  }
  // Packed neighbour records and nodeScale of the current state (active runs).
  void pack() {
    ensureOmega();
    if (nActive == 0) return;
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_pack<DD>, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, f, ph, (int)nActive);
    });
    HIP_CHECK(hipGetLastError());
  }
  // Sum density and/or dt vote over the current neighbour lists; reads the packed records,
  // which must describe the current state (pack()).
  void densityDt(bool dens, bool dtv) {
    if (nActive == 0) return;
    const int nb = (int)xcd_grid(nActive, kBlock);
    if (dtv) hipLaunchKernelGGL(k_reset_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (dens && dtv)
        launchPrecTable(k_density_dt<DD, true, true, double, kGroup>, k_density_dt<DD, true, true, float, kGroup>, nb, f,
                        geo, ph, dtab, (int)nActive, dsc, (int*)nullptr);
      else if (dens)
        launchPrecTable(k_density_dt<DD, true, false, double, kGroup>, k_density_dt<DD, true, false, float, kGroup>, nb,
                        f, geo, ph, dtab, (int)nActive, dsc, (int*)nullptr);
      else
        launchPrecTable(k_density_dt<DD, false, true, double, kGroup>, k_density_dt<DD, false, true, float, kGroup>, nb,
                        f, geo, ph, dtab, (int)nActive, dsc, (int*)nullptr);
    });
  }
  // withV12: the step's pass, which also writes the half-step velocity records.
  // stepMid: the integrator's midpoint evaluation, where the Limited Q's image gradients are
  // what Integrator::postStateUpdate left: tensor images with grad-h on (SPH asks for boundaries
  // after the Q's copy), zero with it off.  Package calls see boundaries applied.
  void sphPass(bool full, bool withV12 = false, bool stepMid = false) {
    // lanes per node: the kernel's own choice, from dimension and precision
    const int64_t lanes = sph_lanes(D, mixed);
    const int nb = (int)xcd_grid(nActive*lanes, kBlock);
    Phys ph = this->ph;
    ph.qImages = stepMid ? (cfg.gradh_correction ? 1 : 0) : 1;
    const size_t tabBytes = tableBytes();
    const int useLds = tabBytes <= kLdsTableMax ? 1 : 0;
    const size_t shm = useLds ? tabBytes : 0;
    const Scalars* sc = withV12 ? (const Scalars*)dsc : nullptr;
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (limitedQ()) {
        if (full) {
          if (mixed) hipLaunchKernelGGL((k_sph<DD, true, float, true>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
          else hipLaunchKernelGGL((k_sph<DD, true, double, true>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
        } else {
          if (mixed) hipLaunchKernelGGL((k_sph<DD, false, float, true>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
          else hipLaunchKernelGGL((k_sph<DD, false, double, true>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
        }
      } else if (full) {
        if (mixed) hipLaunchKernelGGL((k_sph<DD, true, float>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
        else hipLaunchKernelGGL((k_sph<DD, true, double>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
      } else {
        if (mixed) hipLaunchKernelGGL((k_sph<DD, false, float>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
        else hipLaunchKernelGGL((k_sph<DD, false, double>), dim3(nb), dim3(kBlock), shm, stream, f, geo, ph, dtab, (int)nActive, useLds, sc);
      }
    });
    HIP_CHECK(hipGetLastError());
  }
  // TreeGravity::initialize (tree on the given positions) + evaluateDerivatives + dt vote.
  // toState: the evaluation's potential also becomes the State's (potS), as it does for the
  // start-up dry run, a standalone evaluation and SynchronousRK2's start-of-step evaluation; a
  // midpoint evaluation's potential is replaced by the state0 copy's in the reference.
  // forked: the gravity stream has been forked from the main stream by forkGravity(); the pass
  // runs there, and joinGravity() makes the main stream wait for it.  Otherwise it runs in order
  // on the main stream.
  void gravityPass(bool mid, bool toState = false, bool forked = false) {
    if (!cfg.gravity || nActive == 0) return;
    grav.stream = forked ? gstream : stream;
    double* xs[3] = {mid ? f.xm[0] : f.x[0], mid ? f.xm[1] : f.x[1], mid ? f.xm[2] : f.x[2]};
    grav.build(xs, f.m, (int)nActive);
    if (!forked) mark("tree_build");
    grav.evaluate(f.DvDtG, f.pot, (int)nActive, cfg.G, cfg.softening, cfg.opening, dsc, cfg.ftimestep);
    if (toState)
      HIP_CHECK(hipMemcpyAsync(f.potS, f.pot, nActive*sizeof(double), hipMemcpyDeviceToDevice, grav.stream));
    if (forked) {
      HIP_CHECK(hipEventRecord(evJoin, gstream));
      gravPending = true;
      grav.stream = stream;
    }
  }
  // The gravity stream starts from everything queued on the main stream so far.
  void forkGravity() {
    HIP_CHECK(hipEventRecord(evFork, stream));
    HIP_CHECK(hipStreamWaitEvent(gstream, evFork, 0));
  }
  void joinGravity() {
    if (!gravPending) return;
    HIP_CHECK(hipStreamWaitEvent(stream, evJoin, 0));
    gravPending = false;
  }
  void copyStateToMid() {
    ensureOmega();
    auto cp = [&](double* d, const double* s) {
      HIP_CHECK(hipMemcpyAsync(d, s, nActive*sizeof(double), hipMemcpyDeviceToDevice, stream));
    };
    for (int a = 0; a < D; ++a) { cp(f.xm[a], f.x[a]); cp(f.vm[a], f.v[a]); }
    cp(f.rhom, f.rho); cp(f.epsm, f.eps); cp(f.hinvm, f.hinv); cp(f.Pm, f.P); cp(f.csm, f.cs); cp(f.omegam, f.omega);
    packMid(false);
  }
  // Midpoint records for the SPH pass (k_pack_mid).
  void packMid(bool keepOmega) {
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_pack_mid<DD>, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, f, (int)nActive,
                         keepOmega ? 1 : 0);
    });
    HIP_CHECK(hipGetLastError());
  }

  void mark(const char* name) {
    if (!timing) return;
    if (evCount == evPool.size()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      evPool.push_back(e);
      evNames.push_back(nullptr);
    }
    HIP_CHECK(hipEventRecord(evPool[evCount], stream));
    evNames[evCount] = name;
    ++evCount;
  }
  void collectTimes() {
    if (!timing || evCount == 0) return;
    HIP_CHECK(hipEventSynchronize(evPool[evCount - 1]));
    // summed per class over the steps since timing was enabled
    for (size_t k = 1; k < evCount; ++k) {
      float v = 0.f;
      HIP_CHECK(hipEventElapsedTime(&v, evPool[k - 1], evPool[k]));
      auto it = std::find(tnames.begin(), tnames.end(), std::string(evNames[k]));
      if (it == tnames.end()) { tnames.push_back(evNames[k]); tms.push_back(v); }
      else tms[it - tnames.begin()] += v;
    }
    evCount = 0;
  }

  // ---------------------------------------------------------------- lifecycle
  // Start-up of runs admitted since the last step (the reinitializeProblem subset):
  // EOS on the uploaded density before anything replaces it (prepare), then, on this
  // step's connectivity, grad-h Omega and the TreeGravity dry run (finish).
  bool prepareNewRuns() {
    bool any = false;
    for (int r = 0; r < maxRuns; ++r) {
      hinit[r] = runs[r].needInit && runs[r].state == RUN_ACTIVE ? 1 : 0;
      any |= hinit[r] != 0;
    }
    if (!any) return false;
    HIP_CHECK(hipMemcpyAsync(dinit, hinit, sizeof(int)*maxRuns, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_eos_masked, dim3(nblocks(nTotal)), dim3(kBlock), 0, stream, f, ph, (const int*)dinit, (int)nTotal);
    HIP_CHECK(hipGetLastError());
    return true;
  }
  void finishNewRuns() {
    omegaPass(false, dinit);
    if (cfg.gravity) gravityPass(false, true);
    HIP_CHECK(hipStreamSynchronize(stream));
    for (int r = 0; r < maxRuns; ++r) if (hinit[r]) runs[r].needInit = false;
  }

  // Explicit start-up of a single run (spx_run_initialize).
  void initialize() {
    if (nTotal == 0) raise(SPX_ERR_STATE, "no state uploaded");
    // derivatives start at zero (SPHBase::initializeProblemStartup)
    auto z = [&](double* p) { HIP_CHECK(hipMemsetAsync(p, 0, nTotal*sizeof(double), stream)); };
    for (int a = 0; a < D; ++a) { z(f.DxDt[a]); z(f.DvDt[a]); z(f.DvDtG[a]); }
    for (double* p : {f.DepsDt, f.DrhoDt, f.DHDt, f.maxQ}) z(p);
    for (int c = 0; c < D*D; ++c) z(f.DvDx + (size_t)c*ld);
    if (limitedQ()) for (int c = 0; c < D*D; ++c) z(f.DvDxQ + (size_t)c*ld);
    for (auto& r : runs) if (r.state == RUN_ACTIVE) r.needInit = true;
    const bool fresh = prepareNewRuns();
    buildConnectivity(false);
    if (fresh) finishNewRuns();
    initialized = true;
  }

  // SpheralController.iterateIdealH at start-up (SimulationControl/SpheralController.py:231-232,
  // 909-931; Utilities/iterateIdealH.cc:28-230) for a single run: each pass rebuilds the images
  // and the connectivity on the current H, then k_ideal_h replaces H on the nodes not yet done.
  // Ends with the connectivity of the final H.  Returns the number of passes.
  int iterateIdealH(int maxIterations, double tolerance) {
    if (!initialized) raise(SPX_ERR_STATE, "spx_run_initialize must be called first");
    if (maxRuns != 1) raise(SPX_ERR_UNSUPPORTED, "iterateIdealH runs on a single run");
    if (maxIterations < 0 || !(tolerance >= 0.0)) raise(SPX_ERR_ARG, "bad iteration count or tolerance");
    int* ddone = nullptr;
    unsigned long long* dmax = nullptr;
    HIP_CHECK(hipMalloc(&ddone, std::max<int64_t>(nTotal, 1)*sizeof(int)));
    HIP_CHECK(hipMalloc(&dmax, sizeof(unsigned long long)));
    int itr = 0;
    try {
      HIP_CHECK(hipMemsetAsync(ddone, 0, nTotal*sizeof(int), stream));
      double maxDeltaH = 2.0*tolerance;
      while (itr < maxIterations && maxDeltaH > tolerance) {
        ++itr;
        buildConnectivity(false, true);
        HIP_CHECK(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), stream));
        launchD([&](auto d) {
          constexpr int DD = decltype(d)::value;
          hipLaunchKernelGGL(k_ideal_h<DD>, dim3(xcd_grid(nActive, kBlock)), dim3(kBlock), 0, stream, f, geo, ph,
                             dtab, (int)nActive, tolerance, ddone, dmax);
        });
        HIP_CHECK(hipGetLastError());
        unsigned long long bits = 0;
        download(&bits, dmax, sizeof(bits));
        std::memcpy(&maxDeltaH, &bits, sizeof(double));
      }
      buildConnectivity(false, true);
    } catch (...) {
      (void)hipStreamSynchronize(stream);
      (void)hipFree(ddone);
      (void)hipFree(dmax);
      throw;
    }
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipFree(ddone));
    HIP_CHECK(hipFree(dmax));
    return itr;
  }

  // Integrator::step with SynchronousRK2::step when TreeGravity is the only package
  // (Integrator/SynchronousRK2.cc:40-135): dt from the previous evaluation's vote, the tree and its
  // walk at the start and at the midpoint, the full step with the midpoint derivatives.  No
  // neighbours are needed: nothing but TreeGravity reads the state.
  void stepSynchronousGravity() {
    if (timing) {
      evCount = 0;
      mark("start");
    }
    if (nActive == 0) return;
    const int nb = (int)xcd_grid(nActive, kBlock);
    hipLaunchKernelGGL(k_reset_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns);
    hipLaunchKernelGGL(k_select_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns, ph, cfg.dt_min,
                       cfg.dt_max, cfg.dt_growth, 1);
    HIP_CHECK(hipGetLastError());
    mark("select_dt");
    gravityPass(false, true);
    mark("gravity");
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_grav_mid<DD>, dim3(nb), dim3(kBlock), 0, stream, f, (const Scalars*)dsc, (int)nActive);
    });
    HIP_CHECK(hipGetLastError());
    mark("update");
    gravityPass(true);
    mark("gravity");
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_grav_full<DD>, dim3(nb), dim3(kBlock), 0, stream, f, (const Scalars*)dsc, (int)nActive);
    });
    HIP_CHECK(hipGetLastError());
    mark("update");
    hipLaunchKernelGGL(k_advance_time, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns);
    HIP_CHECK(hipGetLastError());
    download(hsc, dsc, sizeof(Scalars)*maxRuns);
    collectTimes();
  }

  // One Integrator::step of every active run.
  void step() {
    if (!retire && !initialized) raise(SPX_ERR_STATE, "spx_run_initialize must be called before stepping");
    if (cfg.gravity_only) { stepSynchronousGravity(); return; }
    if (cfg.integrator == SPX_INTEGRATOR_SYNCHRONOUS_RK2) { stepSynchronousHydro(); return; }
    omegaStale = false;              // this step never reads the last step's end-of-step Omega
    maybeStartTuning();
    if (graphsAllowed()) { graphStep(); return; }
    ++eagerSteps;
    if (timing) {
      evCount = 0;
      mark("start");
    }
    const bool fresh = prepareNewRuns();
    // Integrator::step -> boundary enforcement, connectivity; preStepInitialize
    // (RigorousSumDensity) and the hydro dt vote ride on the neighbour pass
    buildConnectivity(true);
    if (nActive == 0) return;
    mark("connectivity");
    if (fresh) finishNewRuns();
    const int nb = (int)xcd_grid(nActive, kBlock);
    // Integrator::selectDt; TreeGravity's vote is the previous evaluation's
    hipLaunchKernelGGL(k_select_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns, ph, cfg.dt_min,
                       cfg.dt_max, cfg.dt_growth, cfg.gravity);
    HIP_CHECK(hipGetLastError());
    mark("select_dt");
    // half step with the previous derivatives
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_midstep<DD>, dim3(nb), dim3(kBlock), 0, stream, f, ph, (const Scalars*)dsc, (int)nActive);
    });
    HIP_CHECK(hipGetLastError());
    mark("update");
    // TreeGravity at the midpoint reads the midpoint positions and the masses and writes the
    // gravity acceleration, potential and dt vote, which only the full update reads: outside
    // per-class timing it runs on its own stream beside the grad-h and SPH passes (forked here,
    // its launches queued after the SPH pass's so the host's waits inside the tree build do not
    // hold them back, joined before the update)
    const bool overlap = cfg.gravity && !timing;
    if (overlap) forkGravity();
    omegaPass(true);
    mark("omega");
    // derivatives at the midpoint
    if (cfg.gravity && !overlap) { gravityPass(true); mark("gravity"); }
    // the Limited Q reads its gradient (the previous evaluation's DvDx, ArtificialViscosityHandle::
    // postStateUpdate) while this evaluation writes DvDx; the end-of-step postStateUpdate then
    // copies the new DvDx into it, as the reference's state holds it after the step
    sphPass(false, true, true);
    copyQVelocityGradient();
    if (overlap) gravityPass(true, false, true);
    mark("sph");
    joinGravity();
    fullStep(nb);
  }

  // Integrator::step with SynchronousRK2::step and the SPH package (Integrator/SynchronousRK2.cc:
  // 40-135): the connectivity pass (sum density, the hydro dt vote) and the dt selection as in
  // step(); derivatives at the start of the step, whose half-step velocity records carry
  // multiplier dt/2; the trial half step with the full policies (k_half_update) into the midpoint
  // arrays; the midpoint grad-h Omega and the Q's gradient copy (postStateUpdate); derivatives at
  // the midpoint; then the full step from the start-of-step state as in step().
  void stepSynchronousHydro() {
    if (timing) {
      evCount = 0;
      mark("start");
    }
    const bool fresh = prepareNewRuns();
    buildConnectivity(true);
    if (nActive == 0) return;
    mark("connectivity");
    if (fresh) finishNewRuns();
    const int nb = (int)xcd_grid(nActive, kBlock);
    hipLaunchKernelGGL(k_select_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns, ph, cfg.dt_min,
                       cfg.dt_max, cfg.dt_growth, cfg.gravity);
    HIP_CHECK(hipGetLastError());
    mark("select_dt");
    // derivatives at the start of the step (boundaries applied: the Q sees its gradient's images)
    copyStateToMid();
    if (cfg.gravity) { gravityPass(false, true); mark("gravity"); }
    sphPass(false);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_v12<DD>, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, f, (const Scalars*)dsc, 0.25,
                         (int)nActive);
    });
    HIP_CHECK(hipGetLastError());
    copyQVelocityGradient();
    mark("sph");
    // trial half step, then the midpoint postStateUpdate (grad-h Omega and the records)
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      launchPrec(k_half_update<DD, double, kGroupUpdate>, k_half_update<DD, float, kGroupUpdate>, nb, f, geo, ph,
                 (const Scalars*)dsc, (int)nActive);
    });
    mark("update");
    omegaPass(true);
    mark("omega");
    // derivatives at the midpoint
    if (cfg.gravity) { gravityPass(true); mark("gravity"); }
    sphPass(false, true, true);
    copyQVelocityGradient();
    mark("sph");
    fullStep(nb);
  }

  // The full step from the start-of-step state with the midpoint derivatives (compatible energy
  // first), the end-of-step grad-h Omega and the time advance; shared by both integrators.
  void fullStep(int nb) {
    fullStepLaunches(nb);
    omegaStale = cfg.gradh_correction != 0;
    if (eagerOmega) ensureOmega();
    download(hsc, dsc, sizeof(Scalars)*maxRuns);
    collectTimes();
    finishTuningStep();
    tunerStepDone();
    endOfStep();
  }
  // The full step's launches: update (compatible energy first), velocity buffer swap, end-of-step
  // grad-h Omega, time advance.
  void fullStepLaunches(int nb) {
    tuneMark(6);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (groupSize(2) == 8)
        launchPrec(k_full_update<DD, false, double, 8>, k_full_update<DD, false, float, 8>, nb, f, geo, ph, dsc, 0.0,
                   (int)nActive);
      else
        launchPrec(k_full_update<DD, false, double, 4>, k_full_update<DD, false, float, 4>, nb, f, geo, ph, dsc, 0.0,
                   (int)nActive);
    });
    tuneMark(7);
    for (int a = 0; a < D; ++a) {
      std::swap(f.v[a], f.v1[a]);
      // the update wrote only the active runs: the finished runs behind them keep their velocity
      if (nTotal > nActive) {
        hipLaunchKernelGGL(k_copy_doubles, dim3(nblocks(nTotal - nActive)), dim3(kBlock), 0, stream,
                           (const double*)(f.v1[a] + nActive), f.v[a] + nActive, nTotal - nActive);
        HIP_CHECK(hipGetLastError());
      }
    }
    mark("energy_update");
    wrapPending = true;
    hipLaunchKernelGGL(k_advance_time, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns);
    HIP_CHECK(hipGetLastError());
  }
  // Host bookkeeping once the step's run scalars are downloaded.
  void endOfStep() {
    for (int r = 0; r < maxRuns; ++r)
      if (runs[r].state == RUN_ACTIVE && retire && reached_end(hsc[r].t, hsc[r].tEnd)) {
        ensureOmega();               // a finishing run's final state, before it leaves the active range
        break;
      }
    bool finished = false;
    for (int r = 0; r < maxRuns; ++r) {
      if (runs[r].state != RUN_ACTIVE) continue;
      runs[r].stepOut = true;
      if (retire && reached_end(hsc[r].t, hsc[r].tEnd)) {
        markDone(r, false);
        finished = true;
      }
    }
    if (finished) retireFinished();
  }

  // ---------------------------------------------------------------- graph-replayed steps
  // After the connectivity pass's build decisions (candidate-list builds and reorders, which
  // may resize buffers and wait on the device, run eagerly), the rest of a CheapSynchronousRK2
  // step -- connectivity records, the filtering sum-density + dt pass, dt selection, half step,
  // midpoint grad-h, derivatives, full step -- launches the same kernels with the same arguments
  // as the step two before it whenever no buffer was reallocated and the active particle count
  // is unchanged (the velocity arrays alternate between two buffers).  That part is captured as
  // a HIP graph per buffer parity and replayed: one launch instead of about fifteen, which is
  // most of a small run's step time.  The graph holds no memset and no host synchronisation: the
  // overflow flag is cleared by a kernel, and with the neighbour lists sized to the candidate
  // lists' capacity the filter cannot overflow (a node's neighbours are a subset of its
  // candidates).  Results are the eager step's bit for bit.  Steps with newly admitted runs,
  // TreeGravity (host-side tree sizing), the Limited Q (its gradient copy), SynchronousRK2,
  // timing or a group-size measurement run eagerly.
  bool graphsAllowed() const {
    if (!useGraphs || timing || tuner.phase >= 0 || cfg.gravity || limitedQ()) return false;
    for (int r = 0; r < maxRuns; ++r)
      if (runs[r].state == RUN_ACTIVE && runs[r].needInit) return false;
    return true;
  }
  std::vector<unsigned char> graphKey() const {
    std::vector<unsigned char> k;
    auto put = [&](const void* p, size_t n) {
      const unsigned char* c = static_cast<const unsigned char*>(p);
      k.insert(k.end(), c, c + n);
    };
    put(&f, sizeof(f));
    put(&nActive, sizeof(nActive));
    put(&nTotal, sizeof(nTotal));
    put(tuner.size, sizeof(tuner.size));
    put(&ph, sizeof(ph));
    put(&geo, sizeof(geo));
    put(&dtab, sizeof(dtab));
    const void* bufs[] = {keys, vals, dsc, dOverflowG};
    put(bufs, sizeof(bufs));
    return k;
  }
  void dropGraphs() {
    for (auto& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
      g.exec = nullptr;
      g.graph = nullptr;
      g.key.clear();
    }
    for (auto& k : seenKeys) k.clear();
  }
  // The eager step's launches after the connectivity build decisions (connectivity records,
  // filter + sum density + dt vote, dt selection, half step, midpoint grad-h, derivatives, full
  // step), with the overflow flag cleared by a kernel.
  void launchSteadyStep() {
    pack();
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, dOverflowG, 1, 0);
    hipLaunchKernelGGL(k_reset_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns);
    launchFilter(xcd_grid(nActive, kBlock), dOverflowG);
    const int nb = (int)xcd_grid(nActive, kBlock);
    hipLaunchKernelGGL(k_select_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns, ph, cfg.dt_min,
                       cfg.dt_max, cfg.dt_growth, cfg.gravity);
    HIP_CHECK(hipGetLastError());
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_midstep<DD>, dim3(nb), dim3(kBlock), 0, stream, f, ph, (const Scalars*)dsc, (int)nActive);
    });
    HIP_CHECK(hipGetLastError());
    omegaPass(true);
    sphPass(false, true, true);
    fullStepLaunches(nb);
  }
  void graphStep() {
    connectivityPrep(false);
    if (nActive == 0) return;
    ensureNbrCapacity(f.ccap);          // every candidate fits: the filter cannot overflow
    const Fields f0 = f;
    std::vector<unsigned char> key = graphKey();
    StepGraph* g = nullptr;
    for (auto& c : graphs)
      if (c.exec && c.key == key) g = &c;
    if (!g && std::find(seenKeys.begin(), seenKeys.end(), key) == seenKeys.end()) {
      // first sighting of this launch configuration: launch it directly and capture it only if
      // it comes back (a batch whose composition changes every few steps would otherwise pay a
      // capture and an instantiation per step and replay none of them)
      seenKeys[seenNext] = std::move(key);
      seenNext ^= 1;
      launchSteadyStep();
      f = f0;
      for (int a = 0; a < D; ++a) std::swap(f.v[a], f.v1[a]);
      omegaStale = cfg.gradh_correction != 0;
      if (eagerOmega) ensureOmega();
      ++eagerSteps;
      connValid = true;
      npairsCache = -1;
      wrapPending = true;
      HIP_CHECK(hipMemcpyAsync(hOverflow, dOverflowG, sizeof(int), hipMemcpyDeviceToHost, stream));
      download(hsc, dsc, sizeof(Scalars)*maxRuns);
      if (*hOverflow != 0) raise(SPX_ERR_STATE, "neighbour list overflow in a steady step");
      tunerStepDone();
      endOfStep();
      return;
    }
    if (!g) {
      StepGraph& c = graphs[graphNext];
      graphNext ^= 1;
      if (c.exec) HIP_CHECK(hipGraphExecDestroy(c.exec));
      if (c.graph) HIP_CHECK(hipGraphDestroy(c.graph));
      c.exec = nullptr;
      c.graph = nullptr;
      c.key.clear();
      HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        launchSteadyStep();
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(stream, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        f = f0;
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(stream, &c.graph));
      HIP_CHECK(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
      c.key = std::move(key);
      ++graphCaptures;
      g = &c;
    }
    f = f0;
    HIP_CHECK(hipGraphLaunch(g->exec, stream));
    for (int a = 0; a < D; ++a) std::swap(f.v[a], f.v1[a]);
    omegaStale = cfg.gradh_correction != 0;
    if (eagerOmega) ensureOmega();
    ++graphReplays;
    connValid = true;
    npairsCache = -1;
    wrapPending = true;
    HIP_CHECK(hipMemcpyAsync(hOverflow, dOverflowG, sizeof(int), hipMemcpyDeviceToHost, stream));
    download(hsc, dsc, sizeof(Scalars)*maxRuns);
    if (*hOverflow != 0) raise(SPX_ERR_STATE, "neighbour list overflow in a replayed step");
    tunerStepDone();
    endOfStep();
  }

  int activeRuns() const {
    int k = 0;
    for (auto& r : runs) k += r.state == RUN_ACTIVE;
    return k;
  }

  void setEndTime(int slot, double tEnd) {
    checkSlot(slot);
    hsc[slot].tEnd = tEnd;
    upload(&dsc[slot].tEnd, &hsc[slot].tEnd, sizeof(double));
  }

  bool limitedQ() const { return cfg.viscosity == SPX_Q_LIMITED_MONAGHAN_GINGOLD; }
  // ArtificialViscosityHandle::postStateUpdate (ArtificialViscosityHandle.cc:168-183): the
  // Limited Q's gradient becomes a copy of the hydro's DvDx of the last evaluation.
  void copyQVelocityGradient() {
    if (!limitedQ() || nTotal == 0) return;
    for (int c = 0; c < D*D; ++c)
      HIP_CHECK(hipMemcpyAsync(f.DvDxQ + (size_t)c*ld, f.DvDx + (size_t)c*ld, nTotal*sizeof(double),
                               hipMemcpyDeviceToDevice, stream));
  }

  // ---------------------------------------------------------------- single-run package calls
  void requireConn() const {
    if (!connValid) raise(SPX_ERR_STATE, "no connectivity: call spx_run_update_connectivity or spx_run_initialize");
  }
  void evaluateDerivatives() {
    requireConn();
    copyStateToMid();
    if (cfg.gravity) gravityPass(true, true);
    sphPass(true);
    HIP_CHECK(hipStreamSynchronize(stream));
    for (auto& r : runs) if (r.state == RUN_ACTIVE) r.stepOut = r.fullOut = true;
  }
  // State values from the caller (Physics::registerState's fields as the integrator's State holds
  // them: position, velocity, mass, mass density, specific thermal energy, H) replace the run's
  // resident ones in place, in the caller's node order; the EOS policies (pressure, sound speed)
  // follow, as State::update fires them after any change of density or energy.  Derivatives,
  // grad-h Omega, the pair accelerations and the connectivity are kept: within a step the
  // reference evaluates the moved state on the connectivity of the step's start.  Candidate lists
  // are marked stale, so the next connectivity build starts from the new positions.  Boundary
  // enforcement of the positions is the caller's (Integrator::enforceBoundaries).
  void updateState(const double* x, const double* v, const double* m, const double* rho, const double* eps,
                   const double* H) {
    if (retire || maxRuns != 1) raise(SPX_ERR_UNSUPPORTED, "state updates apply to single runs");
    if (!initialized) raise(SPX_ERR_STATE, "spx_run_initialize must be called first");
    if (!x || !v || !m || !rho || !eps || !H) raise(SPX_ERR_ARG, "null argument");
    const int64_t n = runs[0].n;
    std::vector<double> hbuf((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      const double* Hi = H + i*D*D;
      const double h = Hi[0];
      for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
          if (Hi[a*D + b] != (a == b ? h : 0.0))
            raise(SPX_ERR_UNSUPPORTED, "H must be isotropic (h^-1 I); ASPH smoothing tensors are not implemented");
      if (!(h > 0.0) || !std::isfinite(h)) raise(SPX_ERR_ARG, "H must be positive definite");
      if (!(m[i] > 0.0) || !(rho[i] > 0.0)) raise(SPX_ERR_ARG, "mass and mass density must be positive");
      hbuf[i] = h;
    }
    if (!stage) stage = dalloc<double>(ld);
    std::vector<double> buf((size_t)n);
    auto put = [&](double* dst, const double* src) {
      upload(stage, src, n*sizeof(double));
      hipLaunchKernelGGL(k_from_local, dim3(nblocks(n)), dim3(kBlock), 0, stream, dst, (const double*)stage,
                         (const int64_t*)f.oid, (int)n);
      HIP_CHECK(hipGetLastError());
    };
    for (int a = 0; a < D; ++a) {
      for (int64_t i = 0; i < n; ++i) buf[i] = x[i*D + a];
      put(f.x[a], buf.data());
      for (int64_t i = 0; i < n; ++i) buf[i] = v[i*D + a];
      put(f.v[a], buf.data());
    }
    put(f.m, m);
    put(f.rho, rho);
    put(f.eps, eps);
    put(f.hinv, hbuf.data());
    eos();
    HIP_CHECK(hipStreamSynchronize(stream));
    runs[0].candValid = false;
  }
  void sumDensity() {
    requireConn();
    pack();
    densityDt(true, false);
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  void hydroDtVote(int slot, double* hydro, double* gravity) {
    requireConn();
    pack();
    densityDt(false, true);
    hipLaunchKernelGGL(k_select_dt, dim3(nblocks(maxRuns)), dim3(kBlock), 0, stream, dsc, maxRuns, ph, 0.0,
                       1.7976931348623157e308, 1.7976931348623157e308, cfg.gravity);
    HIP_CHECK(hipGetLastError());
    download(hsc, dsc, sizeof(Scalars)*maxRuns);
    *hydro = hsc[slot].hydroDt;
    *gravity = cfg.gravity ? hsc[slot].gravDt : 0.0;
  }
  void compatibleEnergy(double multiplier) {
    requireConn();
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_v12<DD>, dim3(nblocks(nActive)), dim3(kBlock), 0, stream, f, (const Scalars*)nullptr,
                         0.5*multiplier, (int)nActive);
      launchPrec(k_full_update<DD, true, double, kGroupUpdate>, k_full_update<DD, true, float, kGroupUpdate>,
                 (int)xcd_grid(nActive, kBlock), f, geo, ph, dsc, multiplier, (int)nActive);
    });
    eos();
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // Slots of one run in the current order and the run-local index of each.
  // The slots of run `slot` (its contiguous segment [offset, offset + n)) and the run-local
  // node index held in each: one download of the segment's ids, not of the whole batch.
  void runSlots(int slot, std::vector<int64_t>& local, std::vector<int64_t>& slots) {
    const int64_t o = runs[slot].offset, n = runs[slot].n;
    std::vector<int64_t> oid(n);
    if (n) download(oid.data(), f.oid + o, n*sizeof(int64_t));
    local.resize(n);
    slots.resize(n);
    for (int64_t k = 0; k < n; ++k) {
      if ((oid[k] >> 40) != slot) raise(SPX_ERR_STATE, "internal: a run's particles are not one segment");
      slots[k] = o + k;
      local[k] = oid[k] & (((int64_t)1 << 40) - 1);
    }
  }

  void neighbors(int slot, int64_t* offsets, int64_t* nbr, int32_t* img, int64_t capacity, int32_t* countsOnly) {
    checkSlot(slot);
    requireConn();
    if (runs[slot].state != RUN_ACTIVE) raise(SPX_ERR_STATE, "the run holds no connectivity");
    std::vector<int64_t> local, slots;
    runSlots(slot, local, slots);
    const int64_t n = runs[slot].n;
    std::vector<int32_t> cnt(nActive);
    std::vector<int64_t> oid(nActive);
    download(cnt.data(), f.cnt, nActive*sizeof(int32_t));
    download(oid.data(), f.oid, nActive*sizeof(int64_t));
    if (countsOnly) {
      for (size_t k = 0; k < slots.size(); ++k) countsOnly[local[k]] = cnt[slots[k]];
      return;
    }
    std::vector<uint32_t> ell((size_t)f.cap*ld);
    download(ell.data(), f.nbr, ell.size()*sizeof(uint32_t));
    std::vector<int64_t> off(n + 1, 0), slotOf(n);
    for (size_t k = 0; k < slots.size(); ++k) { off[local[k] + 1] = cnt[slots[k]]; slotOf[local[k]] = slots[k]; }
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    if (capacity < off[n]) raise(SPX_ERR_ARG, "neighbour capacity too small");
    const int64_t mask = ((int64_t)1 << 40) - 1;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t p = slotOf[i];
      std::vector<std::pair<int64_t, int32_t>> l;
      for (int k = 0; k < cnt[p]; ++k) {
        const uint32_t e = ell[ell_at(k, (int)p, (size_t)ld)];
        l.push_back({oid[e & kIdxMask] & mask, (int32_t)(e >> kImgShift)});
      }
      std::sort(l.begin(), l.end());
      for (size_t k = 0; k < l.size(); ++k) { nbr[off[i] + k] = l[k].first; img[off[i] + k] = l[k].second; }
      offsets[i] = off[i];
    }
    offsets[n] = off[n];
  }

  // Unique pairs of the active runs: internal pairs once, image pairs once per (node, image) entry.
  // scratch: nblocks(nActive) + 1 counters of device memory for the per-block partial counts and
  // the total (the library's own buffer when null).
  int64_t numPairs(unsigned long long* scratch = nullptr) {
    requireConn();
    if (npairsCache >= 0 && !scratch) return npairsCache;
    const int nb = std::max(1, nblocks(nActive));
    if (!scratch) {
      if (!pairScratch || pairScratchBlocks < nb) {
        HIP_CHECK(hipStreamSynchronize(stream));
        dfree(pairScratch);
        pairScratch = dalloc<unsigned long long>((size_t)nb + 1);
        pairScratchBlocks = nb;
      }
      scratch = pairScratch;
    }
    hipLaunchKernelGGL(k_count_pairs, dim3(nb), dim3(kBlock), 0, stream, (const uint32_t*)f.nbr, (const int32_t*)f.cnt,
                       (size_t)ld, (int)nActive, scratch);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, stream, (const unsigned long long*)scratch, nb, scratch + nb);
    HIP_CHECK(hipGetLastError());
    unsigned long long np = 0;
    download(&np, scratch + nb, sizeof(np));
    npairsCache = (int64_t)np;
    return npairsCache;
  }

  // ---------------------------------------------------------------- field access
  int64_t getField(int slot, const std::string& name, double* out, int64_t capacity) {
    checkSlot(slot);
    applyPendingBoundaries();
    const RunInfo& ri = runs[slot];
    auto needStep = [&] {
      if (!ri.stepOut) raise(SPX_ERR_STATE, "'" + name + "' is not retained after the batch re-sorts; read it right after the run's step");
    };
    auto needFull = [&] {
      if (!ri.fullOut) raise(SPX_ERR_STATE, "'" + name + "' is produced by spx_run_evaluate_derivatives only");
    };
    std::vector<int64_t> local, slots;
    runSlots(slot, local, slots);
    const int64_t n = ri.n;
    // the run's segment of an array: tmp[k] is slot slots[k] = offset + k
    std::vector<double> tmp(n);
    auto dl = [&](const double* src) { if (n) download(tmp.data(), src + ri.offset, n*sizeof(double)); };
    auto scalar = [&](const double* src) -> int64_t {
      if (capacity < n) raise(SPX_ERR_ARG, "output capacity too small");
      dl(src);
      for (size_t k = 0; k < slots.size(); ++k) out[local[k]] = tmp[k];
      return n;
    };
    auto vec = [&](double* const* src) -> int64_t {
      if (capacity < n*D) raise(SPX_ERR_ARG, "output capacity too small");
      for (int a = 0; a < D; ++a) {
        dl(src[a]);
        for (size_t k = 0; k < slots.size(); ++k) out[local[k]*D + a] = tmp[k];
      }
      return n*D;
    };
    auto ten = [&](const double* src) -> int64_t {
      if (capacity < n*D*D) raise(SPX_ERR_ARG, "output capacity too small");
      for (int c = 0; c < D*D; ++c) {
        dl(src + (size_t)c*ld);
        for (size_t k = 0; k < slots.size(); ++k) out[local[k]*D*D + c] = tmp[k];
      }
      return n*D*D;
    };
    auto iso = [&](const double* src) -> int64_t {
      if (capacity < n*D*D) raise(SPX_ERR_ARG, "output capacity too small");
      dl(src);
      for (size_t k = 0; k < slots.size(); ++k)
        for (int a = 0; a < D; ++a)
          for (int b = 0; b < D; ++b) out[local[k]*D*D + a*D + b] = a == b ? tmp[k] : 0.0;
      return n*D*D;
    };
    if (name == "position") return vec(f.x);
    if (name == "velocity") return vec(f.v);
    if (name == "mass") return scalar(f.m);
    if (name == "mass density") return scalar(f.rho);
    if (name == "specific thermal energy") return scalar(f.eps);
    if (name == "H") return iso(f.hinv);
    if (name == "pressure") return scalar(f.P);
    if (name == "sound speed") return scalar(f.cs);
    if (name == "grad h corrections") {
      ensureOmega();
      return scalar(f.omega);
    }
    if (name == "delta velocity hydro") return vec(f.DvDt);
    if (name == "delta velocity") return vec(f.DvDtG);
    if (name == "delta position") return vec(f.DxDt);
    if (name == "delta specific thermal energy") return scalar(f.DepsDt);
    if (name == "delta mass density") return scalar(f.DrhoDt);
    if (name == "velocity gradient") return ten(f.DvDx);
    if (name == "delta H") return iso(f.DHDt);
    if (name == "max viscous pressure") return scalar(f.maxQ);
    if (name == "gravitational potential") return scalar(f.potS);
    if (name == "new H") { needStep(); return iso(f.Hideal); }
    if (name == "mass zeroth moment") { needStep(); return scalar(f.m0); }
    if (name == "effective viscous pressure") { needFull(); return scalar(f.effQ); }
    if (name == "mass first moment") { needFull(); return vec(f.m1); }
    if (name == "M SPH correction") { needFull(); return ten(f.Mcorr); }
    if (name == "mass density gradient") { needFull(); return vec(f.gradRho); }
    if (name == "normalization") { needFull(); return scalar(f.norm); }
    if (name == "new mass density") { needFull(); return scalar(f.rhoSum); }
    raise(SPX_ERR_ARG, "unknown field '" + name + "'");
  }

  // Totals of one run: mass, momentum (3), kinetic, thermal, gravitational potential energy.
  void conservation(int slot, double* out) {
    checkSlot(slot);
    const int64_t n = runs[slot].n;
    std::vector<double> m(n), e(n), v(n*D), pot(n, 0.0);
    getField(slot, "mass", m.data(), n);
    getField(slot, "specific thermal energy", e.data(), n);
    getField(slot, "velocity", v.data(), n*D);
    if (cfg.gravity) {
      // TreeGravity::extraEnergy: 0.5 sum m phi of the last evaluation (a member, not State)
      std::vector<int64_t> local, slots;
      runSlots(slot, local, slots);
      std::vector<double> tmp(n);
      if (n) download(tmp.data(), f.pot + runs[slot].offset, n*sizeof(double));
      for (size_t k = 0; k < slots.size(); ++k) pot[local[k]] = tmp[k];
    }
    double M = 0, P[3] = {0, 0, 0}, KE = 0, TE = 0, PE = 0;
    for (int64_t i = 0; i < n; ++i) {
      M += m[i];
      double v2 = 0.0;
      for (int a = 0; a < D; ++a) { P[a] += m[i]*v[i*D + a]; v2 += v[i*D + a]*v[i*D + a]; }
      KE += 0.5*m[i]*v2;
      TE += m[i]*e[i];
      PE += 0.5*m[i]*pot[i];
    }
    out[0] = M; out[1] = P[0]; out[2] = P[1]; out[3] = P[2]; out[4] = KE; out[5] = TE; out[6] = cfg.gravity ? PE : 0.0;
  }
};

}  // namespace spx
