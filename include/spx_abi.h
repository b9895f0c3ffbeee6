/*
 * spx_abi.h -- C ABI of the MI355X SPH hydro + TreeGravity engine (libspheral_amd.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No C++
 * exceptions and no torch types cross this boundary.  Every call returns 0 on
 * success or a negative SPX_ERR_* code; spx_last_error() returns the message of
 * the last failing call made on the calling thread.
 *
 * Each entry point replaces a piece of the reference's Physics-package surface
 * for the SPH hydro (SPH + MonaghanGingoldViscosity + SPHSmoothingScale) and
 * TreeGravity packages, and of the CheapSynchronousRK2 integrator that drives
 * them.  The reference interface each one stands in for is cited beside it
 * (paths relative to the reference's src/).
 *
 * Threading: a run is not re-entrant; calls on one run must be serialised by the
 * caller.  Different runs may be driven from different host threads.
 *
 * Array layouts at the boundary follow the reference's Field storage: one
 * record per node, vectors as nDim contiguous doubles, H as the full nDim x nDim
 * row-major symmetric matrix, tensors as nDim x nDim row-major.
 */
#ifndef SPX_ABI_H
#define SPX_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define SPX_OK               0
#define SPX_ERR_ARG         -1   /* bad argument / inconsistent configuration   */
#define SPX_ERR_HIP         -2   /* a HIP runtime call failed                   */
#define SPX_ERR_STATE       -3   /* call out of order (e.g. step before state)  */
#define SPX_ERR_UNSUPPORTED -4   /* option the GPU path does not implement      */
#define SPX_ERR_CAPACITY    -5   /* a size limit of the engine was exceeded     */
#define SPX_ERR_NUMERIC     -6   /* non-finite state detected on the device     */

/* ---- enums ---------------------------------------------------------------- */
/* Base kernels tabulated by TableKernel (Kernel/BSplineKernelInline.hh,
   Kernel/WendlandC4KernelInline.hh). */
#define SPX_KERNEL_BSPLINE     0
#define SPX_KERNEL_WENDLANDC4  1
/* NBSplineKernel(order) (Kernel/NBSplineKernel.cc:20-110): the centred B-spline with
   order+1 knots, extent (order+1)/2 in integer arithmetic, normalised by a Simpson
   volume integral (Kernel/VolumeIntegrationFunctions.cc). Order in kernel_order. */
#define SPX_KERNEL_NBSPLINE    2

/* Artificial viscosity of the SPH package.  MonaghanGingold: MonaghanGingoldViscosity
   (ArtificialViscosity/MonaghanGingoldViscosity.cc:41-101).  Limited:
   LimitedMonaghanGingoldViscosity (ArtificialViscosity/LimitedMonaghanGingoldViscosity.cc:
   123-219), the SPH() factory's default Q (SPH/SPHHydros.py:88-94): the pair velocity
   jump is reduced by a van Leer-limited projection of each node's velocity gradient as
   of the previous derivative evaluation (ArtificialViscosityHandle.cc:168-183). */
#define SPX_Q_MONAGHAN_GINGOLD          0
#define SPX_Q_LIMITED_MONAGHAN_GINGOLD  1

/* Time integrator.  CheapSynchronousRK2 (Integrator/CheapSynchronousRK2.cc:40-132): one
   derivative evaluation per step, at the midpoint.  SynchronousRK2 (Integrator/SynchronousRK2.cc:
   40-135): evaluations at the start and at the midpoint of every step, the trial half step with
   the full policies; with the SPH package (plus an optional TreeGravity) or with TreeGravity as
   the only package (gravity_only = 1). */
#define SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2  0
#define SPX_INTEGRATOR_SYNCHRONOUS_RK2        1

/* TreeGravity's time step vote (Gravity/TreeGravity.cc:483-520; GravityTimeStepType):
   AccelerationRatio ftimestep*min sqrt(softening/|a|), or DynamicalTime ftimestep/sqrt(G rhoMax)
   with rhoMax the largest local effective density over the interaction buckets
   (TreeGravity.cc:239-300). */
#define SPX_GRAVITY_DT_ACCELERATION_RATIO  0
#define SPX_GRAVITY_DT_DYNAMICAL_TIME      1

/* Per-side boundary condition of one axis (Boundary/PeriodicBoundary.cc,
   Boundary/ReflectingBoundary.cc).  Periodic must be set on both sides. */
#define SPX_BC_NONE        0
#define SPX_BC_PERIODIC    1
#define SPX_BC_REFLECTING  2

/* Arithmetic of the pair passes. FP64: everything in double.  MIXED: pair
   arithmetic in float on cell-local offsets, double accumulation and state. */
#define SPX_PREC_FP64   0
#define SPX_PREC_MIXED  1

/* ---- configuration -------------------------------------------------------- */
/* One NodeList of gamma-law fluid evolved by SPH (+ optional TreeGravity) under
   CheapSynchronousRK2.  Field meaning and defaults follow the reference:
     kernel/kernel_points/min_nperh/max_nperh : TableKernel ctor (Kernel/TableKernel.cc:169-209)
     nperh, neighbor_extent, hmin, hmax, rho_min, rho_max : makeFluidNodeList (NodeList/FluidNodeLists.py:15-31)
     gamma : GammaLawGas; Cl, Cq : MonaghanGingoldViscosity
     cfl .. correct_velocity_gradient : SPH ctor (SPH/SPH.hh:40-56, SPH/SPHHydros.py:9-33)
     dt_min .. last_dt : Integrator members (Integrator/Integrator.cc:41-47)
     G .. ftimestep : TreeGravity ctor (Gravity/TreeGravity.hh:46-50)
     kernel_order : NBSplineKernel ctor (Kernel/NBSplineKernelInline.hh:19-29)
     viscosity, eta_crit_frac, eta_fold_frac : the Q and the Limited Q's ctor arguments
       (PYB11/ArtificialViscosity/LimitedMonaghanGingoldViscosity.py:24-32; defaults 1.0, 0.2)
     integrator, gravity_only : the integrator class and its package list (an N-body script
       appends TreeGravity alone, tests/functional/Gravity/CollisionlessSphereCollapse.py:178-186)
     gravity_time_step : TreeGravity timeStepChoice (Gravity/TreeGravity.hh:46-50) */
typedef struct spx_config {
  int32_t dim;                       /* 1, 2 or 3                                    */
  int32_t kernel;                    /* SPX_KERNEL_*                                 */
  int32_t kernel_points;             /* TableKernel numPoints                        */
  int32_t precision;                 /* SPX_PREC_*                                   */
  double  min_nperh, max_nperh;      /* TableKernel nperh lookup range               */
  double  nperh;                     /* NodeList nodesPerSmoothingScale              */
  double  neighbor_extent;           /* NodeList neighbor kernelExtent               */
  double  hmin, hmax;                /* NodeList h bounds                            */
  double  rho_min, rho_max;          /* NodeList mass-density bounds                 */
  double  gamma;                     /* GammaLawGas gamma                            */
  double  Cl, Cq;                    /* MonaghanGingoldViscosity coefficients        */
  double  cfl;
  int32_t use_velocity_magnitude_for_dt;
  int32_t compatible_energy;         /* must be 1 (compatibleEnergyEvolution)        */
  int32_t gradh_correction;          /* gradhCorrection                              */
  int32_t correct_velocity_gradient; /* correctVelocityGradient                      */
  double  dt_min, dt_max, dt_growth; /* Integrator dtMin, dtMax, dtGrowth            */
  double  last_dt;                   /* Integrator lastDt at start                   */
  int32_t bc_lo[3], bc_hi[3];        /* SPX_BC_* per axis side                       */
  double  xmin[3], xmax[3];          /* boundary plane positions per axis            */
  int32_t gravity;                   /* 1: TreeGravity package appended              */
  int32_t reorder_interval;          /* 0 or 1: particles are re-sorted every step   */
  double  G, softening, opening, ftimestep;
  int32_t kernel_order;              /* NBSplineKernel order (SPX_KERNEL_NBSPLINE)   */
  int32_t viscosity;                 /* SPX_Q_*                                      */
  double  eta_crit_frac, eta_fold_frac; /* LimitedMonaghanGingoldViscosity ctor     */
  int32_t integrator;                /* SPX_INTEGRATOR_*                             */
  int32_t gravity_only;              /* 1: TreeGravity is the only package (no SPH;  */
                                     /*    needs gravity = 1 and SynchronousRK2)     */
  int32_t gravity_time_step;         /* SPX_GRAVITY_DT_*: TreeGravity timeStepChoice */
} spx_config;

typedef struct spx_run spx_run;
typedef struct spx_batch spx_batch;

/* ---- library / errors ---------------------------------------------------- */
const char* spx_last_error(void);
/* Default configuration (reference defaults for every field). */
int spx_config_default(int32_t dim, spx_config* out);
/* Device count visible to the library (0 => no GPU). */
int spx_device_count(int32_t* out);

/* ---- kernel table (host, FP64) --------------------------------------------
   TableKernel<Dim>(BaseKernel, numPoints, minNperh, maxNperh)
   (Kernel/TableKernel.cc:169-209, Utilities/QuadraticInterpolator.cc:59-100).
   Writes W, dW/deta for eta[0..n) with Hdet=1 via kernelAndGradValue
   (Kernel/TableKernelViewInline.hh:84-99) and the Wsum->nPerh lookup
   equivalentNodesPerSmoothingScale (Kernel/TableKernelViewInline.hh:151-156). */
int spx_table_eval(const spx_config* cfg, const double* eta, int64_t n,
                   double* W, double* gW, double* nperh_of_wsum);
/* Raw interpolator coefficients: 3*(N1+1) each for W and gradW.  *ncoef is
   set to 3*(N1+1); pass coef arrays of at least that size (or NULL to query). */
int spx_table_coefficients(const spx_config* cfg, int64_t* ncoef,
                           double* wcoef, double* gwcoef, double* xstep);

/* ---- run lifetime ---------------------------------------------------------
   A run owns one NodeList's state in HBM plus the packages that evolve it:
   SPH (SPH/SPH.cc:43-80) with MonaghanGingoldViscosity and SPHSmoothingScale
   (SPH/SPHHydros.py:88-140), optional TreeGravity, and CheapSynchronousRK2. */
int  spx_run_create(const spx_config* cfg, int64_t n, int32_t device, spx_run** out);
void spx_run_destroy(spx_run* run);

/* Upload a full state (host pointers; x,v: n*dim; m,rho,eps: n; H: n*dim*dim).
   Replaces NodeList positions/velocity/mass/massDensity/specificThermalEnergy/Hfield. */
int spx_run_set_state(spx_run* run, const double* x, const double* v,
                      const double* m, const double* rho, const double* eps,
                      const double* H);

/* The integrator's State handed to a package hook (Physics::evaluateDerivatives / dt /
   preStepInitialize / postStateUpdate take `const State&`, Physics/Physics.hh:49-73; the fields
   SPHBase::registerState enrolls, SPH/SPHBase.cc:203-262): replaces position, velocity, mass,
   mass density, specific thermal energy and H of an initialized run in place (same layouts and
   node order as spx_run_set_state) and re-evaluates pressure and sound speed (the EOS policies
   State::update fires, Hydro/PressurePolicy.cc, Hydro/SoundSpeedPolicy.cc).  Keeps the
   derivatives, grad-h Omega, the pair accelerations of the last evaluation and the connectivity:
   inside a step the reference evaluates the moved state on the connectivity of the step's start
   (Integrator/CheapSynchronousRK2.cc:56-88).  A driver that keeps its State on the host calls
   this at the top of each hook that reads State, so every hook sees the caller's current state.
   Boundary enforcement of positions stays with the caller (Integrator::enforceBoundaries). */
int spx_run_update_state(spx_run* run, const double* x, const double* v, const double* m,
                         const double* rho, const double* eps, const double* H);

/* Copy one field to host in the caller's original node order.  Names follow
   HydroFieldNames (Hydro/HydroFieldNames.cc:11-40): "position", "velocity",
   "mass", "mass density", "specific thermal energy", "H", "pressure",
   "sound speed", "grad h corrections", "delta velocity hydro" (hydro DvDt),
   "delta velocity" (gravity DvDt), "delta position", "delta specific thermal energy",
   "delta mass density", "velocity gradient", "delta H", "new H",
   "max viscous pressure", "effective viscous pressure", "mass zeroth moment",
   "mass first moment", "M SPH correction", "mass density gradient",
   "normalization", "new mass density", "gravitational potential".
   *count receives the number of doubles written (n * components).
   "new H" and "mass zeroth moment" come from the last step or derivative
   evaluation; "effective viscous pressure", "mass first moment", "M SPH correction",
   "mass density gradient", "normalization" and "new mass density" only from
   spx_run_evaluate_derivatives (SPX_ERR_STATE when not in place). */
int spx_run_get_field(spx_run* run, const char* name, double* out, int64_t capacity,
                      int64_t* count);

/* SpheralController.reinitializeProblem subset (SimulationControl/SpheralController.py:154-300):
   ghost images + connectivity, then initializeProblemStartupDependencies of every
   package: P and cs (SPH/SPHBase.cc:147-161), grad-h Omega, TreeGravity dry run
   (Gravity/TreeGravity.cc:343-356).  Derivatives start at zero. */
int spx_run_initialize(spx_run* run);

/* SpheralController.iterateIdealH at start-up (SimulationControl/SpheralController.py:231-232,
   909-931 -> Utilities/iterateIdealH.cc:28-230, SPHSmoothingScale alone, no spherical start, no
   fixed determinant): repeat { ghost images + connectivity on the current H; the package's ideal
   H from the zeroth moment (SmoothingScale/SPHSmoothingScale.cc:154-245); replace H on every node
   not yet converged } until no node's H moved by more than tolerance, or max_iterations passes
   (the controller uses 50 and 1e-4).  Call after spx_run_initialize and before the first step;
   *passes receives the number of passes.  Single runs only. */
int spx_run_iterate_ideal_h(spx_run* run, int32_t max_iterations, double tolerance, int32_t* passes);

/* Integrator::step(maxTime) with CheapSynchronousRK2::step
   (Integrator/Integrator.cc:67-111, Integrator/CheapSynchronousRK2.cc:40-132).
   *dt_out receives the step taken.  Device work is enqueued on the run's stream;
   the call returns after the step's dt has been read back. */
int spx_run_step(spx_run* run, double max_time, double* dt_out);

/* Advance nsteps steps (or until max_time) without per-step host reads of dt
   where possible; *steps_done, *t_out report progress. */
int spx_run_advance(spx_run* run, double max_time, int64_t nsteps,
                    int64_t* steps_done, double* t_out);

int spx_run_time(spx_run* run, double* t, int64_t* cycle, double* last_dt);

/* ---- package-level entry points (Physics interface, Physics/Physics.hh:49-73) --- */
/* ConnectivityMap rebuild from the current positions and H
   (Neighbor/ConnectivityMap.cc:750-1152 via Integrator::setGhostNodes). */
int spx_run_update_connectivity(spx_run* run);
/* Number of neighbours (gather+scatter) per node, original order
   (ConnectivityMap::numNeighborsForNode). */
int spx_run_neighbor_counts(spx_run* run, int32_t* counts);
/* Neighbour sets: for node i (original order) entries offsets[i]..offsets[i+1]
   hold (neighbour original index, image code); image code packs per axis 2 bits
   (0 none, 1 image beyond the low plane, 2 beyond the high plane), axis a at
   bits 2a..2a+1.  Each node's list is sorted by (index, image).  offsets has n+1
   entries; nbr/img need capacity total = offsets[n]. */
int spx_run_neighbors(spx_run* run, int64_t* offsets, int64_t* nbr, int32_t* img,
                      int64_t capacity);
/* SPHBase::preStepInitialize with RigorousSumDensity: mass density <- SPH sum
   (SPH/computeSPHSumMassDensity.cc, SPH/SPHBase.cc:336-353). */
int spx_run_sum_density(spx_run* run);
/* EOS policies on the current state: pressure, sound speed
   (Hydro/PressurePolicy.cc, Hydro/SoundSpeedPolicy.cc, Material/GammaLawGas.cc). */
int spx_run_update_eos(spx_run* run);
/* SPHBase::postStateUpdate: grad-h Omega (SPH/computeSPHOmegaGradhCorrection.cc). */
int spx_run_post_state_update(spx_run* run);
/* SPH::evaluateDerivatives + SPHSmoothingScale::evaluateDerivatives (+ TreeGravity
   initialize/evaluateDerivatives if enabled) on the current state
   (SPH/SPH.cc:165-555, SmoothingScale/SPHSmoothingScale.cc:101-275,
   Gravity/TreeGravity.cc:143-328).  Every derivative output is overwritten. */
int spx_run_evaluate_derivatives(spx_run* run, double t, double dt);
/* Physics::dt votes: GenericHydro::dt (Physics/GenericHydro.cc:112-381, includes the
   cfl factor) and TreeGravity::dt (Gravity/TreeGravity.cc:483-497). */
int spx_run_dt(spx_run* run, double* hydro_dt, double* gravity_dt);
/* SpecificThermalEnergyPolicy::update (Hydro/SpecificThermalEnergyPolicy.cc:50-175):
   eps <- eps + multiplier * compatible DepsDt, using the current derivatives. */
int spx_run_compatible_energy(spx_run* run, double multiplier);

/* ---- diagnostics --------------------------------------------------------- */
/* SpheralConservation totals (SimulationControl/SpheralConservation.py:37-150):
   out[0]=mass, out[1..3]=linear momentum, out[4]=kinetic, out[5]=thermal,
   out[6]=gravitational extra energy (0.5 sum m phi). */
int spx_run_conservation(spx_run* run, double* out9);
/* Device time per kernel class (ms), measured with HIP events on the run's stream and
   summed over the steps since timing was enabled; max_entries < 0 enables timing and
   clears the sums.  names: comma-separated labels written to buf. */
int spx_run_kernel_times(spx_run* run, double* ms, int32_t max_entries,
                         int32_t* n_entries, char* names, int64_t names_cap);
/* Unique pairs of the current connectivity (NodePairList size): ConnectivityMap::nodePairList()
   (Neighbor/ConnectivityMap.hh:79; pairs admitted by calculatePairInteraction,
   ConnectivityMapInline.hh:296-309), counted on the device into a library-allocated buffer. */
int spx_run_num_pairs(spx_run* run, int64_t* npairs);
/* spx_run_num_pairs without any device allocation.  No counterpart in the reference: for callers
   that cannot allocate device memory (code under stream/graph capture, serving loops that forbid
   hipMalloc).  Counts into a 512 KiB device array that the library holds per GPU for the life of
   the process; runs of more than 16,777,216 active particles get SPX_ERR_CAPACITY.  Calls that
   overlap on one GPU share that array and race; the array keeps the last call's counts.  Calls
   no hipMalloc/hipFree.  Requires the run's connectivity (SPX_ERR_STATE otherwise). */
int spx_run_num_pairs_noalloc(spx_run* run, int64_t* npairs);
/* Candidate-list builds of the run so far (no counterpart in the reference, which rebuilds
   its ConnectivityMap from scratch every step: a measurement aid).  The neighbour sets are
   the reference's every step either way; a build happens at start-up, after
   spx_run_update_connectivity, and when the run's nodes have moved or grown enough that
   the candidate list could miss a pair. */
int spx_run_neighbor_builds(spx_run* run, int64_t* builds);
/* Device memory the run holds, in bytes: its state and work arrays, neighbour and candidate
   lists (which grow with the largest neighbour count seen) and the tree (no counterpart in the
   reference: for admission by memory budget). */
int spx_run_device_bytes(spx_run* run, int64_t* bytes);
/* TreeGravity interactions of one walk over the run's current tree (no counterpart in the
   reference: a measurement aid for the walk's throughput).  out2[0]: accepted-cell (monopole)
   interactions, out2[1]: point-point interactions in opened leaves, summed over the points.
   Runs a counting copy of the walk; allocates and waits on the device. */
int spx_run_gravity_interactions(spx_run* run, int64_t* out2);
/* Step execution (no counterpart in the reference, whose Integrator::step is plain host
   calls).  A CheapSynchronousRK2 step in which no candidate list is rebuilt, no run is
   admitted, finished or removed and no field was read since the previous step is replayed
   from a HIP graph captured on its first occurrence (one launch instead of about fifteen);
   its results are the eager step's bit for bit.  Enabled by default; enable = 0 runs every
   step eagerly (and drops the captured graphs).  TreeGravity, the Limited Q and
   SynchronousRK2 steps always run eagerly. */
int spx_run_set_step_graphs(spx_run* run, int32_t enable);
/* No counterpart in the reference (an evaluation-schedule switch).  The grad-h Omega that
   SPHBase::postStateUpdate leaves at the end of a CheapSynchronousRK2 step (SPHBase.cc:526-551)
   is read by nothing in the next step (which recomputes it at its midpoint), so by default the
   engine computes it when it is first needed -- a state read of "grad h corrections", a package
   hook, SynchronousRK2, a run finishing in a batch -- from the same end-of-step positions and
   neighbour lists: the same bits whenever it is read.  eager != 0 computes it at the end of
   every step, as the reference's integrator does. */
int spx_run_set_end_of_step_omega(spx_run* run, int32_t eager);
/* Neighbour-group sizes of the filtering sum-density pass, the grad-h passes and the
   compatible-energy update (sizes[0..3), each 4 or 8: how many neighbours' loads a lane
   issues before using them).  Every size gives the same result bits.  By default (sizes null
   or sizes[0] <= 0) the engine measures: after two warm-up steps it alternates 4 and 8 per
   pass over six steps, timing each, moves a pass off its default (4, 8, 8) only when the
   other size is at least 3 % faster, and measures again once the active particle count has
   changed by more than 2x. */
int spx_run_set_group_sizes(spx_run* run, const int32_t* sizes);
/* out9: graph captures, graph replays, eager steps, group-size measurements, current sizes of
   the three passes, whole-batch reorders, steps whose finished runs retired without a reorder
   (the last two count batch composition changes; a single run reorders once). */
int spx_run_step_stats(spx_run* run, int64_t* out9);
/* HBM copy probe (no counterpart in the reference: a measurement aid).  Copies `bytes`
   between two device buffers `reps` times with a plain streaming kernel and returns the
   achieved rate, read + write bytes per second, in GB/s. */
int spx_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps);

/* ---- batches: many independent runs of one problem type --------------------
   The reference advances one problem per process: SpheralController.advance
   (SimulationControl/SpheralController.py:383-426) loops Integrator::step on one
   DataBase.  A batch holds many such problems in one HBM-resident set of arrays
   and steps all of them together, each with its own dt (Integrator::selectDt per
   run), neighbour grid and end time; pairs never cross runs.  This serves
   ensembles and request streams of small/medium problems that cannot fill the GPU
   alone.  Every run of a batch shares cfg (kernel, EOS, viscosity, boundary
   planes); TreeGravity batches hold one run.  A run is stepped until its time
   reaches t_end (the last step is clamped to land on t_end), then it is DONE and
   its state stays readable until it is removed. */
#define SPX_RUN_ACTIVE 1
#define SPX_RUN_DONE   2

/* capacity: total particles resident at once; max_runs: run slots (<= 32768). */
int  spx_batch_create(const spx_config* cfg, int64_t capacity, int32_t max_runs,
                      int32_t device, spx_batch** out);
void spx_batch_destroy(spx_batch* batch);
/* Admit a run (state layout as spx_run_set_state).  Its start-up
   (reinitializeProblem subset, as spx_run_initialize) happens at the next step.
   *run receives the run id.  SPX_ERR_CAPACITY when particles or slots run out. */
int spx_batch_add_run(spx_batch* batch, const double* x, const double* v,
                      const double* m, const double* rho, const double* eps,
                      const double* H, int64_t n, double t_end, int32_t* run);
/* Release a run's slot and particles (reclaimed at the next step or admission). */
int spx_batch_remove_run(spx_batch* batch, int32_t run);
/* One Integrator::step of every active run; *active_runs receives the number of
   runs still active afterwards. */
int spx_batch_step(spx_batch* batch, int32_t* active_runs);
/* State of one run: SPX_RUN_ACTIVE or SPX_RUN_DONE, time, cycle, last dt, nodes. */
int spx_batch_run_status(spx_batch* batch, int32_t run, int32_t* state, double* t,
                         int64_t* cycle, double* last_dt, int64_t* n);
/* As spx_run_get_field for one run.  "new H" and "mass zeroth moment" are kept
   only until the batch's next step; the evaluate-only outputs are not produced. */
int spx_batch_get_field(spx_batch* batch, int32_t run, const char* name, double* out,
                        int64_t capacity, int64_t* count);
/* As spx_run_conservation for one run. */
int spx_batch_conservation(spx_batch* batch, int32_t run, double* out9);
/* Occupancy for admission control: particles and run slots held by active and finished runs
   (removed runs' slots and particles count as free: the next admission reclaims them). */
int spx_batch_usage(spx_batch* batch, int64_t* particles, int32_t* runs);
/* As spx_run_neighbor_builds for one run of a batch. */
int spx_batch_neighbor_builds(spx_batch* batch, int32_t run, int64_t* builds);
/* As spx_run_set_step_graphs, spx_run_set_group_sizes and spx_run_step_stats for the batch's
   steps (one measurement and one graph pair per batch: the runs share the launches). */
int spx_batch_set_step_graphs(spx_batch* batch, int32_t enable);
/* As spx_run_device_bytes for the batch (all its runs share the arrays sized by its capacity). */
int spx_batch_device_bytes(spx_batch* batch, int64_t* bytes);
int spx_batch_set_group_sizes(spx_batch* batch, const int32_t* sizes);
int spx_batch_step_stats(spx_batch* batch, int64_t* out9);
/* As spx_run_kernel_times for the batch's steps. */
int spx_batch_kernel_times(spx_batch* batch, double* ms, int32_t max_entries,
                           int32_t* n_entries, char* names, int64_t names_cap);

#ifdef __cplusplus
}
#endif

#endif /* SPX_ABI_H */
